"""The batched product sumcheck (crates/tessera-sumcheck: Table, Instance, prove_batch) on the
MI355X back end. Tables are numpy uint64 arrays (Fp: shape (2^n,), Fp2: shape (2^n, 2)) or
DeviceBuffers; the round loop and its transcript run on the device."""
import ctypes
from dataclasses import dataclass

import numpy as np

from .device import DeviceBuffer
from .lib import check, lib, u32, u64, vp

_P = 0xFFFFFFFF00000001


class _CTable(ctypes.Structure):
    _fields_ = [("data", vp), ("ext", u32), ("is_public", u32)]


class _CTerm(ctypes.Structure):
    _fields_ = [("weight", u64 * 2), ("factors", ctypes.POINTER(u32)), ("nfactors", u32)]


class _CInstance(ctypes.Structure):
    _fields_ = [("vars", u32), ("factors", u32), ("tables", ctypes.POINTER(_CTable)),
                ("terms", ctypes.POINTER(_CTerm)), ("nterms", u32)]


@dataclass
class SumOfProducts:
    """Instance::sum_of_products(tables, terms) (prover.rs:300-312): the relation
    sum_t weight_t * prod_{k in factors_t} tables[k]; terms are (weight, [positions]) with the
    weight an Fp2 pair or an int (Term::new, terms.rs:24-28)."""
    tables: list
    terms: list


TABLE_EQ = 2  # tsr_table.ext of an eq table given by its point


@dataclass
class Table:
    """Table::base / Table::ext, optionally .public() (prover.rs:78-230). Table.eq(point) is the
    table eq_table(point) (tessera-poly eq.rs:23) given by its point: the device never builds it
    in a product of three tables at degree 3, and builds it from the point elsewhere."""
    values: object
    ext: bool = False
    is_public: bool = False
    vars: int = None
    eq_point: bool = False

    @staticmethod
    def base(values):
        return Table(values, ext=False)

    @staticmethod
    def ext_(values):
        return Table(values, ext=True)

    @staticmethod
    def eq(point):
        pt = np.ascontiguousarray(np.asarray(point, dtype=np.uint64).reshape(-1, 2))
        return Table(pt, ext=True, vars=pt.shape[0], eq_point=True)

    def public(self):
        return Table(self.values, self.ext, True, self.vars, self.eq_point)

    def n_vars(self):
        if self.vars is not None:
            return self.vars
        if isinstance(self.values, DeviceBuffer):
            n = self.values.count // (2 if self.ext else 1)
        else:
            n = np.asarray(self.values).size // (2 if self.ext else 1)
        return n.bit_length() - 1


@dataclass
class ProvenBatch:
    gamma: tuple
    rounds: np.ndarray      # n_max x degree x 2
    challenges: np.ndarray  # n_max x 2
    final_evals: np.ndarray  # n_finals x 2


_prove_fn = None


def _prove():
    """tsr_sumcheck_prove with the instance array typed (set once)."""
    global _prove_fn
    if _prove_fn is None:
        f = lib().tsr_sumcheck_prove
        f.restype = ctypes.c_int
        f.argtypes = [vp, vp, ctypes.c_char_p, u32, ctypes.POINTER(_CInstance), u32, vp, vp, vp, vp,
                      u64, ctypes.POINTER(u64)]
        _prove_fn = f
    return _prove_fn


def prove_batch(device, transcript, prefix: str, degree: int, instances) -> ProvenBatch:
    """prove_batch(transcript, prefix, degree, instances) (prover.rs:1122-1145); each instance is
    a list of Table (an Instance::new product) or a SumOfProducts."""
    keep = []
    cinst = (_CInstance * max(1, len(instances)))()
    n_max, n_final = 0, 0
    for i, tabs in enumerate(instances):
        terms = None
        if isinstance(tabs, SumOfProducts):
            tabs, terms = tabs.tables, tabs.terms
        arr = (_CTable * max(1, len(tabs)))()
        vars_ = tabs[0].n_vars() if tabs else 0
        for k, t in enumerate(tabs):
            if isinstance(t.values, DeviceBuffer):
                ptr = t.values.ptr
            else:
                v = np.ascontiguousarray(np.asarray(t.values, dtype=np.uint64).reshape(-1))
                keep.append(v)
                ptr = v.ctypes.data
            arr[k].data = ptr
            arr[k].ext = TABLE_EQ if t.eq_point else int(t.ext)
            arr[k].is_public = int(t.is_public)
            n_final += 0 if t.is_public else 1
        keep.append(arr)
        cinst[i].vars = vars_
        cinst[i].factors = len(tabs)
        cinst[i].tables = arr
        if terms is not None:
            tarr = (_CTerm * max(1, len(terms)))()
            for t, (w, factors) in enumerate(terms):
                w = (w, 0) if isinstance(w, int) else w
                f = (u32 * max(1, len(factors)))(*factors)
                keep.append(f)
                tarr[t].weight[0], tarr[t].weight[1] = int(w[0]) % _P, int(w[1]) % _P
                tarr[t].factors = f
                tarr[t].nfactors = len(factors)
            keep.append(tarr)
            cinst[i].terms = tarr
            cinst[i].nterms = len(terms)
        n_max = max(n_max, vars_)
    # one host block for gamma, the round messages, the challenges and the final evaluations
    n_round = max(1, n_max * degree) * 2
    n_chal = max(1, n_max) * 2
    out = np.empty(2 + n_round + n_chal + max(1, n_final) * 2, dtype=np.uint64)
    base = out.ctypes.data
    nf = u64()
    check(_prove()(device.handle, transcript.handle, prefix.encode(), degree, cinst, len(instances),
                   base, base + 16, base + 16 + 8 * n_round, base + 16 + 8 * (n_round + n_chal),
                   max(1, n_final), ctypes.byref(nf)))
    o_r, o_c, o_f = 2, 2 + n_round, 2 + n_round + n_chal
    return ProvenBatch((int(out[0]), int(out[1])),
                       out[o_r:o_r + n_max * degree * 2].reshape(n_max, degree, 2),
                       out[o_c:o_c + n_max * 2].reshape(n_max, 2),
                       out[o_f:o_f + nf.value * 2].reshape(nf.value, 2))


def eq_table(device, point, out=None):
    """eq(r, .) (tessera-poly eq.rs:23) for a point of n Fp2 (array (n, 2)); into `out` (a
    DeviceBuffer of 2^n * 2 words) or returned as a numpy (2^n, 2) array."""
    pt = np.ascontiguousarray(np.asarray(point, dtype=np.uint64).reshape(-1, 2))
    n = pt.shape[0]
    dst = out if out is not None else np.zeros((1 << n, 2), dtype=np.uint64)
    ptr = dst.ptr if isinstance(dst, DeviceBuffer) else dst.ctypes.data
    check(lib().tsr_eq_table(device.handle, pt.ctypes.data if n else None, n, ptr))
    return dst
