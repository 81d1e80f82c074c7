"""LogUp-GKR (crates/tessera-logup: LogUpInstance, multiplicities, prove_logup) on the MI355X
back end: fraction trees are built and proven level by level on the device."""
import ctypes
from dataclasses import dataclass

import numpy as np

from .device import DeviceBuffer, ptr_of
from .lib import check, lib, u64


class _CInstance(ctypes.Structure):
    _fields_ = [("rows", u64), ("keys", ctypes.c_void_p), ("values", ctypes.c_void_p),
                ("lookups", u64), ("lookup_keys", ctypes.c_void_p),
                ("lookup_values", ctypes.c_void_p), ("multiplicities", ctypes.c_void_p)]


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


def _count(x):
    return x.count if isinstance(x, DeviceBuffer) else np.asarray(x).size


def multiplicities(device, rows: int, lookup_rows):
    """Exact counts over 2^ceil(log2 rows) entries, row 0 counting the lookup padding."""
    lr = lookup_rows if isinstance(lookup_rows, DeviceBuffer) else np.ascontiguousarray(lookup_rows, np.uint64)
    h = (rows - 1).bit_length() if rows > 1 else 0
    out = np.zeros(1 << h, dtype=np.uint64)
    check(lib().tsr_logup_multiplicities(device.handle, rows, ptr_of(lr), _count(lr), out.ctypes.data))
    return out


def prove_logup(device, transcript, instances, schedule=None, log_cap=1 << 22):
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
    n = len(instances)
    sched = None if schedule is None else np.ascontiguousarray(schedule, dtype=np.uint32)
    alpha, beta = np.zeros(2, np.uint64), np.zeros(2, np.uint64)
    roots = np.zeros(max(1, 8 * n), np.uint64)
    values = np.zeros(max(1, 2 * n), np.uint64)
    claims = np.zeros(max(1, 8 * n), np.uint64)
    nv, ll = u64(), u64()
    log = np.zeros(log_cap, np.uint64)
    check(lib().tsr_logup_prove(device.handle, transcript.handle, arr, n,
                                sched.ctypes.data if sched is not None else None,
                                alpha.ctypes.data, beta.ctypes.data, roots.ctypes.data,
                                values.ctypes.data, ctypes.byref(nv), claims.ctypes.data,
                                log.ctypes.data, log_cap, ctypes.byref(ll)))
    return ProvenLogUp((int(alpha[0]), int(alpha[1])), (int(beta[0]), int(beta[1])),
                       roots[:8 * n].reshape(-1, 2), values[:2 * nv.value].reshape(-1, 2),
                       claims[:8 * n].reshape(-1, 2), parse_level_log(log[:ll.value]))


def parse_level_log(log):
    """Per level: gamma, lambda, rounds, challenges, final_evals, rho, state, counter."""
    out, at = [], 0
    log = [int(x) for x in log]

    def f2(i):
        return (log[i], log[i + 1])
    while at < len(log):
        n = log[at]
        at += 1
        lv = {"gamma": f2(at), "lambda": f2(at + 2)}
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
