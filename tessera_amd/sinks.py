"""The reference's device sinks over the C ABI: BlockSink (crates/tessera-pcs/src/streaming.rs:24-38)
and WeightEncoding::matrix / MatrixSink (crates/tessera-pcs/src/weights/regenerated.rs:24-62)."""
import ctypes

import numpy as np

from .device import DeviceBuffer
from .lib import Declined, check, lib, vp


def _id(commitment_id, keep):
    buf = ctypes.create_string_buffer(bytes(commitment_id), 32)
    keep.append(buf)
    return ctypes.cast(buf, ctypes.c_void_p)


def _ptr(x, keep):
    if isinstance(x, DeviceBuffer):
        return x.ptr
    a = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1)
    keep.append(a)
    return a.ctypes.data


class BlockSink:
    """Encodes blocks of rows with masks and absorbs the codewords into every column chain."""

    def __init__(self, device, log_k, log2_inv_rate, mask_coeffs, block, commitment_id):
        self.k, self.n, self.P = 1 << log_k, 1 << (log_k + log2_inv_rate), mask_coeffs
        h, keep = vp(), []
        check(lib().tsr_block_sink_create(device.handle, log_k, log2_inv_rate, mask_coeffs, block,
                                          _id(commitment_id, keep), ctypes.byref(h)))
        self.handle = h.value

    def absorb(self, rows, masks=None):
        keep = []
        count = (rows.count if isinstance(rows, DeviceBuffer) else np.asarray(rows).size) // self.k
        m = _ptr(masks, keep) if self.P else None
        check(lib().tsr_block_sink_absorb(self.handle, _ptr(rows, keep), count, m))

    def finish(self):
        out = ctypes.create_string_buffer(32 * self.n)
        check(lib().tsr_block_sink_finish(self.handle, out))
        return [out.raw[32 * j:32 * j + 32] for j in range(self.n)]

    def free(self):
        if self.handle:
            lib().tsr_block_sink_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class MatrixSink:
    """A weight matrix's rows encoded and absorbed on the device, answering the queried columns."""

    def __init__(self, handle, k, n, q):
        self.handle, self.k, self.n, self.q = handle, k, n, q

    def absorb(self, rows):
        keep = []
        count = (rows.count if isinstance(rows, DeviceBuffer) else np.asarray(rows).size) // self.k
        out = np.zeros(max(1, count * self.q), dtype=np.uint64)
        check(lib().tsr_matrix_sink_absorb(self.handle, _ptr(rows, keep), count, out.ctypes.data))
        return out[:count * self.q].reshape(count, self.q)

    def finish(self):
        out = ctypes.create_string_buffer(32 * self.n)
        check(lib().tsr_matrix_sink_finish(self.handle, out))
        return [out.raw[32 * j:32 * j + 32] for j in range(self.n)]

    def free(self):
        if self.handle:
            lib().tsr_matrix_sink_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class WeightEncoding:
    """WeightEncoding::matrix on the device: a MatrixSink, or None when the device declines."""

    def __init__(self, device):
        self.device = device

    def matrix(self, log_k, log2_inv_rate, block, commitment_id, columns, queries):
        q = np.ascontiguousarray(np.asarray(queries, dtype=np.uint64))
        h, keep = vp(), []
        try:
            check(lib().tsr_matrix_sink_create(self.device.handle, log_k, log2_inv_rate, block,
                                               _id(commitment_id, keep), columns,
                                               q.ctypes.data if q.size else None, q.size,
                                               ctypes.byref(h)))
        except Declined:
            return None
        return MatrixSink(h.value, 1 << log_k, columns, q.size)


class ColumnSink:
    """ColumnSink (crates/tessera-pcs/src/streaming/opening.rs:845-890): P5's symbols at the sink's
    columns from re-encoded blocks of rows, and P4's combined rows summed on the device."""

    def __init__(self, device, log_k, log2_inv_rate, mask_coeffs, block, columns):
        self.k, self.P = 1 << log_k, mask_coeffs
        cols = np.ascontiguousarray(columns, dtype=np.uint64)
        self.m = cols.size
        h = vp()
        check(lib().tsr_column_sink_create(device.handle, log_k, log2_inv_rate, mask_coeffs, block,
                                           cols.ctypes.data if cols.size else None, cols.size,
                                           ctypes.byref(h)))
        self.handle = h.value

    def combines(self):
        return bool(lib().tsr_column_sink_combines(self.handle))

    def _args(self, rows, masks, keep):
        count = (rows.count if isinstance(rows, DeviceBuffer) else np.asarray(rows).size) // self.k
        m = _ptr(masks, keep) if self.P else None
        out = np.zeros(count * self.m, dtype=np.uint64)
        return count, m, out

    def select(self, rows, masks=None):
        """Each row's symbols at the columns: an array of count x m."""
        keep = []
        count, m, out = self._args(rows, masks, keep)
        check(lib().tsr_column_sink_select(self.handle, _ptr(rows, keep), m, count, out.ctypes.data))
        return out.reshape(count, self.m)

    def select_combining(self, rows, masks, coefficients):
        """coefficients: groups x count Fp2 (array (groups, count, 2)), each group's coefficient of
        each row of the block."""
        keep = []
        count, m, out = self._args(rows, masks, keep)
        c = np.ascontiguousarray(coefficients, dtype=np.uint64).reshape(-1, count, 2)
        check(lib().tsr_column_sink_select_combining(self.handle, _ptr(rows, keep), m, count,
                                                     c.ctypes.data, c.shape[0], out.ctypes.data))
        return out.reshape(count, self.m)

    def combined(self, out):
        """Adds the summed combined rows into out (groups x k Fp2, a uint64 array (groups, k, 2),
        updated in place) and clears them."""
        assert out.dtype == np.uint64 and out.flags.c_contiguous
        check(lib().tsr_column_sink_combined(self.handle, out.ctypes.data, out.size // (2 * self.k)))
        return out

    def free(self):
        if self.handle:
            lib().tsr_column_sink_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
