"""The weight commitment (RFC-0040) on the MI355X back end, with the reference's API shape:
WeightLayoutParams, WeightShapes and ModelWeights in, WeightCommitment out (public artifact, R_W,
the .twr envelope, each matrix's commitment). Mirrors crates/tessera-pcs/src/weights/{layout.rs,
commit.rs, public.rs}; the rows are laid out and committed on the device."""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from .device import DeviceBuffer
from .lib import check, lib, u32, u64, vp

WEIGHT_MATRICES = ("Embed", "Q", "K", "V", "O", "Gate", "Up", "Down", "LmHead", "ChannelParams")


class _Params(ctypes.Structure):
    _fields_ = [("log2_k", ctypes.c_uint8), ("log2_inv_rate", ctypes.c_uint8),
                ("mask_coeffs", u32), ("blinding_rows", u32)]


class _Shape(ctypes.Structure):
    _fields_ = [("out", u32), ("inputs", u32)]


class _Tensor(ctypes.Structure):
    _fields_ = [("matrix", u32), ("layer", u32), ("data", vp), ("len", u64)]


class _Channel(ctypes.Structure):
    _fields_ = [("matrix", u32), ("layer", u32), ("kind", u32), ("count", u64), ("values", vp)]


@dataclass
class WeightShapes:
    """n_layers, the nine (out, inputs) shapes in WeightMatrix order, and the (matrix, layer)
    weights that have a ChannelParams (layout.rs:66-77)."""
    layers: int
    shapes: list
    channels: list = field(default_factory=list)


def _index(matrix):
    return WEIGHT_MATRICES.index(matrix) if isinstance(matrix, str) else int(matrix)


class WeightCommitment:
    """WeightCommitment::commit(layout, row_block, epoch, weights) (commit.rs:61-135)."""

    def __init__(self, device, params: dict, shapes: WeightShapes, row_block: int, epoch: int,
                 tensors, channels=()):
        """tensors: (matrix, layer, INT8 out x inputs array or DeviceBuffer of int8);
        channels: (matrix, layer, uint8 class indices or uint16 multipliers)."""
        keep = []
        p = _Params(params["log2_k"], params["log2_inv_rate"], params["mask_coeffs"], params["blinding_rows"])
        sh = (_Shape * 9)(*[_Shape(o, i) for o, i in shapes.shapes])
        cw = np.array([x for m, l in shapes.channels for x in (_index(m), l)] or [0], dtype=np.uint32)
        tarr = (_Tensor * max(1, len(tensors)))()
        for k, (m, l, data) in enumerate(tensors):
            if isinstance(data, DeviceBuffer):
                ptr, n = data.ptr, data.count
            else:
                a = np.ascontiguousarray(data, dtype=np.int8).reshape(-1)
                keep.append(a)
                ptr, n = a.ctypes.data, a.size
            tarr[k] = _Tensor(_index(m), l, ptr, n)
        carr = (_Channel * max(1, len(channels)))()
        for k, (m, l, values) in enumerate(channels):
            v = np.ascontiguousarray(values)
            if v.dtype not in (np.uint8, np.uint16):
                raise ValueError("channel values are u8 class indices or u16 multipliers")
            keep.append(v)
            carr[k] = _Channel(_index(m), l, 0 if v.dtype == np.uint8 else 1, v.size, v.ctypes.data)
        h = vp()
        check(lib().tsr_weight_commit(device.handle, ctypes.byref(p), shapes.layers, sh,
                                      cw.ctypes.data, len(shapes.channels), row_block, epoch, tarr,
                                      len(tensors), carr, len(channels), ctypes.byref(h)))
        self.handle = h.value

    def public(self) -> bytes:
        out = ctypes.create_string_buffer(334)
        check(lib().tsr_weight_commitment_public(self.handle, out))
        return out.raw

    def weights_root(self) -> bytes:
        out = ctypes.create_string_buffer(32)
        check(lib().tsr_weight_commitment_weights_root(self.handle, out))
        return out.raw

    def artifact(self) -> bytes:
        """The .twr file's bytes."""
        out = ctypes.create_string_buffer(386)
        check(lib().tsr_weight_commitment_artifact(self.handle, out))
        return out.raw

    def matrix(self, matrix):
        """(root, log2_k, row_vars, committed rows) of one matrix."""
        c, lk, rv, rows = vp(), u32(), u32(), u64()
        check(lib().tsr_weight_commitment_matrix(self.handle, _index(matrix), ctypes.byref(c),
                                                 ctypes.byref(lk), ctypes.byref(rv), ctypes.byref(rows)))
        root = ctypes.create_string_buffer(32)
        check(lib().tsr_commitment_root(c.value, root))
        return root.raw, lk.value, rv.value, rows.value

    def free(self):
        if self.handle:
            lib().tsr_weight_commitment_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def envelope_seal(device, schema_id: int, schema_version: int, payload: bytes) -> bytes:
    """Envelope::seal (RFC-0003): the payload in its envelope, digest last."""
    out = ctypes.create_string_buffer(len(payload) + 52)
    n = u64()
    check(lib().tsr_envelope_seal(device.handle, schema_id, schema_version, payload, len(payload),
                                  out, len(payload) + 52, ctypes.byref(n)))
    return out.raw[:n.value]
