"""Batched commitments: many unmasked matrices encoded and committed by one planned, autotuned,
hipGraph-captured launch sequence (each root equal to ReferenceCommitment::commit_unmasked's)."""
import ctypes

import numpy as np

from .device import DeviceBuffer
from .lib import check, lib, u32, u64, vp

AUTOTUNE, GRAPH, SERIAL, PLAN = 1, 2, 4, 8


class CommitDesc(ctypes.Structure):
    _fields_ = [("log2_k", u32), ("log2_inv_rate", u32), ("row_block", u64),
                ("commitment_id", ctypes.c_uint8 * 32), ("rows", vp), ("count", u64)]


class CommitBatch:
    """matrices: list of (rows, log2_k, log2_inv_rate, row_block, commitment_id); rows a numpy
    array or a DeviceBuffer of count x 2^log2_k values."""

    def __init__(self, device, matrices, autotune=True, graph=True, serial=False, plan=None):
        """plan: (tile_log, most stage bits per strided pass) for every group, instead of tuning."""
        L = lib()
        L.tsr_commit_batch_create.restype = ctypes.c_int
        L.tsr_commit_batch_create.argtypes = [vp, ctypes.POINTER(CommitDesc), u32, u32, ctypes.POINTER(vp)]
        L.tsr_commit_batch_run.restype = ctypes.c_int
        L.tsr_commit_batch_run.argtypes = [vp]
        L.tsr_commit_batch_roots.restype = ctypes.c_int
        L.tsr_commit_batch_roots.argtypes = [vp, vp]
        L.tsr_commit_batch_tuning.restype = ctypes.c_int
        L.tsr_commit_batch_tuning.argtypes = [vp, vp, u32, ctypes.POINTER(u32)]
        L.tsr_commit_batch_free.restype = None
        L.tsr_commit_batch_free.argtypes = [vp]
        self.device = device
        self.count = len(matrices)
        descs = (CommitDesc * max(1, self.count))()
        self._keep = []
        for i, (rows, log_k, log_r, block, cid) in enumerate(matrices):
            if isinstance(rows, DeviceBuffer):
                ptr, n_vals = rows.ptr, rows.count
            else:
                a = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
                self._keep.append(a)
                ptr, n_vals = a.ctypes.data, a.size
            descs[i].log2_k = log_k
            descs[i].log2_inv_rate = log_r
            descs[i].row_block = block
            descs[i].commitment_id[:] = list(cid)
            descs[i].rows = ptr
            descs[i].count = n_vals >> log_k
        h = vp()
        flags = (AUTOTUNE if autotune else 0) | (GRAPH if graph else 0) | (SERIAL if serial else 0)
        if plan is not None:
            flags |= PLAN | (plan[0] << 8) | (plan[1] << 16)
        check(L.tsr_commit_batch_create(device.handle, descs, self.count, flags, ctypes.byref(h)))
        self.handle = h.value

    def run(self):
        check(lib().tsr_commit_batch_run(self.handle))

    def roots(self):
        out = ctypes.create_string_buffer(32 * self.count)
        check(lib().tsr_commit_batch_roots(self.handle, out))
        raw = out.raw
        return [raw[32 * i:32 * i + 32] for i in range(self.count)]

    def tuning(self):
        """Per shape group: (log2_k, r, tile_log, most stage bits per strided pass (0: default),
        tuned encode time in us (0: not timed by this batch))."""
        buf = np.zeros(5 * 64, dtype=np.uint32)
        n = u32()
        check(lib().tsr_commit_batch_tuning(self.handle, buf.ctypes.data, 64, ctypes.byref(n)))
        return [tuple(int(x) for x in buf[5 * g:5 * g + 5]) for g in range(min(n.value, 64))]

    def free(self):
        if self.handle:
            lib().tsr_commit_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
