"""The level offload over the C ABI: LevelOffload::load and DeviceLevel::{sums, bind, firsts}
(crates/tessera-sumcheck/src/offload.rs:17-79). Fp2 values are (c0, c1) pairs."""
import ctypes

import numpy as np

from .device import DeviceBuffer
from .lib import check, lib, vp

FRACTION, PRODUCT = 0, 1


def _fp2s(values):
    return np.ascontiguousarray(np.asarray(values, dtype=np.uint64).reshape(-1, 2))


class DeviceLevel:
    def __init__(self, handle, tables):
        self.handle, self.tables = handle, tables

    def sums(self, low, high, low_bits):
        lo, hi = _fp2s(low), _fp2s(high)
        out = np.zeros(6, dtype=np.uint64)
        check(lib().tsr_level_sums(self.handle, lo.ctypes.data, low_bits, hi.ctypes.data,
                                   hi.shape[0], out.ctypes.data))
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(3)]

    def bind(self, r):
        rr = np.array([int(r[0]), int(r[1])], dtype=np.uint64)
        check(lib().tsr_level_bind(self.handle, rr.ctypes.data))

    def firsts(self):
        out = np.zeros(2 * self.tables, dtype=np.uint64)
        check(lib().tsr_level_firsts(self.handle, out.ctypes.data))
        return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(self.tables)]

    def free(self):
        if self.handle:
            lib().tsr_level_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class LevelOffload:
    """load(shape, tables, lambda_): tables of 2^m Fp2 entries each (numpy (2^m, 2) or
    DeviceBuffers of 2^(m+1) words); 4 tables [N0, N1, D0, D1] for FRACTION, 2 for PRODUCT."""

    def __init__(self, device):
        self.device = device

    def load(self, shape, tables, lambda_=(0, 0)):
        keep, ptrs = [], []
        m = None
        for t in tables:
            if isinstance(t, DeviceBuffer):
                ptrs.append(t.ptr)
                entries = t.count // 2
            else:
                a = _fp2s(t)
                keep.append(a)
                ptrs.append(a.ctypes.data)
                entries = a.shape[0]
            m = entries.bit_length() - 1
        arr = (ctypes.c_void_p * len(ptrs))(*ptrs)
        lam = np.array([int(lambda_[0]), int(lambda_[1])], dtype=np.uint64)
        h = vp()
        check(lib().tsr_level_load(self.device.handle, shape, lam.ctypes.data,
                                   ctypes.cast(arr, ctypes.c_void_p), len(ptrs), m or 0, ctypes.byref(h)))
        return DeviceLevel(h.value, len(ptrs))
