"""LogUp-GKR (crates/tessera-logup: LogUpInstance, multiplicities, prove_logup) and prove_trees
over product forests (crates/tessera-sumcheck/src/trees.rs:761-776, forest.rs) on the MI355X back
end: fraction trees and forests are built and proven level by level on the device."""
import ctypes
from dataclasses import dataclass

import numpy as np

from .device import DeviceBuffer, ptr_of
from .lib import check, lib, u64


class _CInstance(ctypes.Structure):
    _fields_ = [("rows", u64), ("keys", ctypes.c_void_p), ("values", ctypes.c_void_p),
                ("lookups", u64), ("lookup_keys", ctypes.c_void_p),
                ("lookup_values", ctypes.c_void_p), ("multiplicities", ctypes.c_void_p)]


class _CForest(ctypes.Structure):
    _fields_ = [("leaves", ctypes.c_void_p), ("len", u64), ("height", ctypes.c_uint32)]


@dataclass
class Forest:
    """A product forest (Forest::new, forest.rs:30-60): 2^s stacked product trees of height
    `height` over the base leaves (a power-of-two count, host array or DeviceBuffer)."""
    leaves: object
    height: int


@dataclass
class LogUpInstance:
    """A table (keys, optional values) and its lookups (as key/value pairs), with the
    multiplicity column (instance.rs:19-102)."""
    keys: object
    values: object
    lookup_keys: object
    lookup_values: object
    multiplicities: object


@dataclass
class ProvenLogUp:
    alpha: tuple
    beta: tuple
    roots: np.ndarray   # 4 per instance, x 2
    values: np.ndarray
    claims: np.ndarray  # 4 per instance, x 2
    levels: list
    forest_claims: list = None  # one (c0, c1) per forest


def _count(x):
    return x.count if isinstance(x, DeviceBuffer) else np.asarray(x).size


def multiplicities(device, rows: int, lookup_rows, out=None):
    """Exact counts over 2^ceil(log2 rows) entries, row 0 counting the lookup padding. out: a
    DeviceBuffer of at least that many u64 to count into on the device (returned), else a new
    host array."""
    lr = lookup_rows if isinstance(lookup_rows, DeviceBuffer) else np.ascontiguousarray(lookup_rows, np.uint64)
    # a table of no row has no leaf (instance.rs:228-231)
    leaves = 0 if rows == 0 else 1 << ((rows - 1).bit_length() if rows > 1 else 0)
    if out is None:
        out = np.zeros(leaves, dtype=np.uint64)
        dst = out.ctypes.data if leaves else None
    else:
        if not isinstance(out, DeviceBuffer) or out.count < leaves:
            raise ValueError(f"out must be a DeviceBuffer of at least {leaves} entries")
        dst = out.ptr
    check(lib().tsr_logup_multiplicities(device.handle, rows, ptr_of(lr), _count(lr), dst))
    return out


def _forest_array(forests, keep):
    arr = (_CForest * max(1, len(forests)))()
    for f, fo in enumerate(forests):
        if isinstance(fo.leaves, DeviceBuffer):
            arr[f].leaves, arr[f].len = fo.leaves.ptr, fo.leaves.count
        else:
            a = np.ascontiguousarray(fo.leaves, dtype=np.uint64)
            keep.append(a)
            arr[f].leaves, arr[f].len = a.ctypes.data, a.size
        arr[f].height = fo.height
    return arr


def _pairs(a, n):
    return [(int(a[2 * i]), int(a[2 * i + 1])) for i in range(n)]


def prove_logup(device, transcript, instances, schedule=None, forests=(), log_cap=1 << 22):
    """prove_logup (prover.rs:307-469): the instances' lookup and table trees beside the product
    forests, levels in `schedule` order (tree ids 2k lookup, 2k+1 table, 2 len(instances) + f
    forest f; None: paired, then the forests)."""
    keep = []
    arr = (_CInstance * max(1, len(instances)))()

    def p(x):
        if x is None:
            return None
        if isinstance(x, DeviceBuffer):
            return x.ptr
        a = np.ascontiguousarray(x, dtype=np.uint64)
        keep.append(a)
        return a.ctypes.data
    for k, ins in enumerate(instances):
        arr[k].rows = _count(ins.keys)
        arr[k].keys = p(ins.keys)
        arr[k].values = p(ins.values)
        arr[k].lookups = _count(ins.lookup_keys)
        arr[k].lookup_keys = p(ins.lookup_keys)
        arr[k].lookup_values = p(ins.lookup_values)
        arr[k].multiplicities = p(ins.multiplicities)
    forests = list(forests)
    farr = _forest_array(forests, keep)
    n, nf = len(instances), len(forests)
    sched = None if schedule is None else np.ascontiguousarray(schedule, dtype=np.uint32)
    alpha, beta = np.zeros(2, np.uint64), np.zeros(2, np.uint64)
    roots = np.zeros(max(1, 8 * n), np.uint64)
    values = np.zeros(max(1, 2 * n), np.uint64)
    claims = np.zeros(max(1, 8 * n), np.uint64)
    fclaims = np.zeros(max(1, 2 * nf), np.uint64)
    nv, ll = u64(), u64()
    log = np.zeros(log_cap, np.uint64)
    check(lib().tsr_logup_prove(device.handle, transcript.handle, arr, n, farr, nf,
                                sched.ctypes.data if sched is not None else None,
                                alpha.ctypes.data, beta.ctypes.data, roots.ctypes.data,
                                values.ctypes.data, ctypes.byref(nv), claims.ctypes.data,
                                fclaims.ctypes.data, log.ctypes.data, log_cap, ctypes.byref(ll)))
    return ProvenLogUp((int(alpha[0]), int(alpha[1])), (int(beta[0]), int(beta[1])),
                       roots[:8 * n].reshape(-1, 2), values[:2 * nv.value].reshape(-1, 2),
                       claims[:8 * n].reshape(-1, 2), parse_level_log(log[:ll.value]),
                       _pairs(fclaims, nf))


def prove_trees(device, transcript, forests, log_cap=1 << 22):
    """prove_trees (trees.rs:761-776) over product forests alone: (the forests' leaf claim
    values, the per-level records)."""
    keep = []
    forests = list(forests)
    farr = _forest_array(forests, keep)
    fclaims = np.zeros(max(1, 2 * len(forests)), np.uint64)
    ll = u64()
    log = np.zeros(log_cap, np.uint64)
    check(lib().tsr_prove_trees(device.handle, transcript.handle, farr, len(forests),
                                fclaims.ctypes.data, log.ctypes.data, log_cap, ctypes.byref(ll)))
    return _pairs(fclaims, len(forests)), parse_level_log(log[:ll.value])


def parse_level_log(log):
    """Per level: root_challenges, gamma, lambda (None without fraction trees), rounds,
    challenges, final_evals, rho, state, counter."""
    out, at = [], 0
    log = [int(x) for x in log]

    def f2(i):
        return (log[i], log[i + 1])
    while at < len(log):
        s_max = log[at]
        at += 1
        lv = {"root_challenges": [f2(at + 2 * c) for c in range(s_max)]}
        at += 2 * s_max
        has_lambda = log[at]
        n = log[at + 1]
        at += 2
        lv["gamma"] = f2(at)
        lv["lambda"] = f2(at + 2) if has_lambda else None
        at += 4
        lv["rounds"] = [[f2(at + 6 * j + 2 * z) for z in range(3)] for j in range(n)]
        at += 6 * n
        lv["challenges"] = [f2(at + 2 * j) for j in range(n)]
        at += 2 * n
        nf = log[at]
        at += 1
        lv["final_evals"] = [f2(at + 2 * i) for i in range(nf)]
        at += 2 * nf
        lv["rho"] = f2(at)
        at += 2
        lv["state"] = b"".join(w.to_bytes(8, "little") for w in log[at:at + 4]).hex()
        at += 4
        lv["counter"] = log[at]
        at += 1
        out.append(lv)
    return out
