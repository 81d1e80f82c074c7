"""The GPU context and device buffers (mirrors tessera-gpu's Device, crates/tessera-gpu/src/device.rs:113-279).

A Device owns one tsr_ctx: one MI355X, one HIP stream, an optional memory budget. DeviceBuffer is
HBM-resident u64 storage that any entry point accepts in place of host memory.
"""
import ctypes

import numpy as np

from .lib import check, lib, vp


class DeviceBuffer:
    def __init__(self, device, count, dtype=np.uint64):
        self.device = device
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = vp()
        check(lib().tsr_alloc(device.handle, max(self.nbytes, 1), ctypes.byref(p)))
        self.ptr = p.value

    def upload(self, array):
        a = np.ascontiguousarray(array, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        check(lib().tsr_copy(self.device.handle, self.ptr, a.ctypes.data, a.nbytes))
        return self

    def download(self, count=None):
        count = self.count if count is None else count
        out = np.empty(count, dtype=self.dtype)
        check(lib().tsr_copy(self.device.handle, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def fill_uniform(self, seed):
        """SplitMix64 values below p, exactly uniform, generated on the device."""
        check(lib().tsr_fill_uniform(self.device.handle, self.ptr, self.count, seed))
        return self

    def free(self):
        if self.ptr:
            check(lib().tsr_free(self.device.handle, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.device.handle:
                lib().tsr_free(self.device.handle, self.ptr)
        except Exception:
            pass


class Device:
    """Device::new(ordinal) (device.rs:125)."""

    def __init__(self, ordinal=0):
        h = vp()
        check(lib().tsr_ctx_create(ordinal, ctypes.byref(h)))
        self.handle = h.value
        self.ordinal = ordinal

    def buffer(self, count, dtype=np.uint64):
        return DeviceBuffer(self, count, dtype)

    def upload(self, array):
        a = np.ascontiguousarray(array)
        return DeviceBuffer(self, a.size, a.dtype).upload(a)

    def synchronize(self):
        check(lib().tsr_ctx_synchronize(self.handle))

    def set_budget(self, nbytes):
        check(lib().tsr_ctx_set_budget(self.handle, nbytes))

    def set_round_blocks(self, blocks):
        """The most workgroups a sumcheck round launch uses (1..1024, 0 = default 1024). Proofs
        do not depend on it; fewer workgroups give each thread more pairs of a round."""
        check(lib().tsr_ctx_set_round_blocks(self.handle, blocks))

    def memory(self):
        return int(lib().tsr_ctx_memory(self.handle))

    def close(self):
        if self.handle:
            lib().tsr_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def ptr_of(x):
    """The address an entry point takes: a DeviceBuffer's device pointer or a numpy array's."""
    if isinstance(x, DeviceBuffer):
        return x.ptr
    if x is None:
        return None
    return x.ctypes.data
