"""The commitment side of Tessera's prover on the MI355X back end, with the reference's API shape.

Mirrors crates/tessera-pcs (RowEncoder, CommitmentParams, ReferenceCommitment, ColumnAnswer) and
crates/tessera-hash (ColumnChains, MerkleTree::root) over the C ABI. Field elements are numpy
uint64 arrays of canonical Goldilocks values; digests are 32-byte `bytes`. Errors are raised as
the reference's PcsError/ColumnError variants would be (tessera_amd.lib.BadShape with the
variant's name in the message), before any device work.
"""
import ctypes
import struct
from dataclasses import dataclass, field

import numpy as np

from .device import Device, DeviceBuffer, ptr_of
from .lib import CommitmentParams as _CParams
from .lib import check, lib, u64, vp


class NttPlan:
    """NttPlan (crates/tessera-pcs/src/ntt.rs:397-516) on the device: transforms of `size` points
    in place, natural order on both sides. `values` is a numpy uint64 array of n values or of
    rows x n (each row one transform; numpy arrays are transformed in place) or a DeviceBuffer of
    a multiple of n values."""

    def __init__(self, device: Device, size: int):
        h = vp()
        check(lib().tsr_ntt_plan_create(device.handle, size, ctypes.byref(h)))
        self.handle = h.value
        self.device = device

    def size(self):
        return int(lib().tsr_ntt_plan_size(self.handle))

    def _run(self, fn, values, *shift):
        if isinstance(values, DeviceBuffer):
            check(fn(self.handle, values.ptr, values.count, *shift))
            return values
        if not (isinstance(values, np.ndarray) and values.dtype == np.uint64 and values.flags.c_contiguous):
            raise TypeError("values: a C-contiguous numpy uint64 array or a DeviceBuffer")
        check(fn(self.handle, values.ctypes.data if values.size else None, values.size, *shift))
        return values

    def ntt(self, values):
        return self._run(lib().tsr_ntt, values)

    def intt(self, values):
        return self._run(lib().tsr_intt, values)

    def coset_ntt(self, values, shift):
        return self._run(lib().tsr_coset_ntt, values, shift)

    def coset_intt(self, values, shift):
        return self._run(lib().tsr_coset_intt, values, shift)

    def free(self):
        if self.handle:
            lib().tsr_ntt_plan_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class RowEncoder:
    """RowEncoder::new(message_len, log2_inv_rate, mask_coeffs) (encoding.rs:126-170)."""

    def __init__(self, device: Device, message_len: int, log2_inv_rate: int, mask_coeffs: int = 0):
        if message_len <= 0 or message_len & (message_len - 1):
            from .lib import BadShape
            raise BadShape(2, f"NotPowerOfTwo: {message_len}")
        self.device = device
        self.log_k = message_len.bit_length() - 1
        self.log_r = log2_inv_rate
        self.mask_coeffs = mask_coeffs

    def message_len(self):
        return 1 << self.log_k

    def codeword_len(self):
        return 1 << (self.log_k + self.log_r)

    def encode_rows(self, rows, masks=None, out=None):
        """encode_rows (encoding.rs:257-290): rows (count x k) -> codewords (count x n)."""
        k, n = self.message_len(), self.codeword_len()
        if isinstance(rows, DeviceBuffer):
            count = rows.count // k
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
            count = rows.size // k
        if masks is not None and not isinstance(masks, DeviceBuffer):
            masks = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1)
        if out is None:
            out = np.zeros(count * n, dtype=np.uint64)
        check(lib().tsr_encode_rows(self.device.handle, self.log_k, self.log_r, self.mask_coeffs,
                                    ptr_of(rows), ptr_of(masks), count, ptr_of(out)))
        return out.reshape(count, n) if isinstance(out, np.ndarray) else out

    def encode_row(self, row, masks=None):
        return self.encode_rows(row, masks)[0]


class ColumnChains:
    """ColumnChains::new / absorb_block / finish (crates/tessera-hash/src/columns.rs:72-148)."""

    def __init__(self, device: Device, commitment_id: bytes, columns: int, block: int):
        h = vp()
        check(lib().tsr_chains_create(device.handle, bytes(commitment_id), columns, block,
                                      ctypes.byref(h)))
        self.handle = h.value
        self.columns = columns

    def absorb_block(self, rows):
        if isinstance(rows, DeviceBuffer):
            count = rows.count // self.columns
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
            count = rows.size // self.columns if self.columns else 0
            if rows.size % self.columns:
                from .lib import BadShape
                raise BadShape(2, "PartialRow: symbols do not fill whole rows")
        check(lib().tsr_chains_absorb(self.handle, ptr_of(rows), count))

    def finish(self):
        out = ctypes.create_string_buffer(32 * self.columns)
        check(lib().tsr_chains_read(self.handle, out))
        raw = out.raw
        lib().tsr_chains_free(self.handle)
        self.handle = None
        return [raw[32 * j:32 * j + 32] for j in range(self.columns)]


class MaskStream:
    """The per-proof advice mask source (Masks as a MaskSource, crates/tessera-prover/src/cpu/
    random.rs:123-229,275-320), generated on the device: BLAKE3 derive-key output of the 32-byte
    seed under tessera/v1/advice-masks, u64 LE words, words >= p discarded whole."""

    def __init__(self, device: Device, seed: bytes):
        if len(seed) != 32:
            raise ValueError("the mask seed is 32 bytes")
        h = vp()
        check(lib().tsr_mask_stream_create(device.handle, bytes(seed), ctypes.byref(h)))
        self.device, self.handle = device, h.value

    def draw(self, count: int, out: DeviceBuffer = None):
        """The next `count` masks: into `out` (device) when given, else a numpy array."""
        if out is not None:
            if out.count < count:
                raise ValueError("the output buffer is shorter than the draw")
            check(lib().tsr_mask_stream_draw(self.handle, count, out.ptr))
            return out
        host = np.zeros(max(1, count), dtype=np.uint64)
        check(lib().tsr_mask_stream_draw(self.handle, count, host.ctypes.data))
        return host[:count]

    @property
    def position(self) -> int:
        """Stream words consumed so far (draws and discarded words)."""
        return int(lib().tsr_mask_stream_position(self.handle))

    def close(self):
        if self.handle:
            lib().tsr_mask_stream_free(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def merkle_root(device: Device, leaves):
    """MerkleTree::new(leaves).root() (crates/tessera-hash/src/merkle.rs:93-119)."""
    data = b"".join(leaves)
    out = ctypes.create_string_buffer(32)
    check(lib().tsr_merkle_root(device.handle, data, len(leaves), out))
    return out.raw


@dataclass
class CommitmentParamsPy:
    """CommitmentParams (crates/tessera-pcs/src/commitment.rs:79-103)."""
    log2_k: int
    log2_inv_rate: int
    mask_coeffs: int
    row_block: int
    commitment_id: bytes = field(default=bytes(32))

    def c(self):
        p = _CParams()
        p.log2_k = self.log2_k
        p.log2_inv_rate = self.log2_inv_rate
        p.mask_coeffs = self.mask_coeffs
        p.row_block = self.row_block
        p.commitment_id[:] = list(self.commitment_id)
        return p


_DIGEST_SPLITS = {}


def _split_digests(buf, count):
    """The first count 32-byte digests of a ctypes buffer, as a list of bytes."""
    if count == 0:
        return []
    st = _DIGEST_SPLITS.get(count)
    if st is None:
        if len(_DIGEST_SPLITS) >= 64:  # a small cache of the formats last used
            _DIGEST_SPLITS.clear()
        st = _DIGEST_SPLITS[count] = struct.Struct("32s" * count)
    return list(st.unpack_from(buf))


@dataclass
class ColumnAnswer:
    """ColumnAnswer (commitment.rs:578-598): columns of the distinct queries + multiproof."""
    indices: list
    columns: list
    siblings: list


class ReferenceCommitment:
    """ReferenceCommitment::commit_unmasked / root / leaves / answer (commitment.rs:343-576),
    with the codewords, leaves and tree resident in HBM."""

    def __init__(self, device, handle, params, count):
        self.device = device
        self.handle = handle
        self.params = params
        self.count = count
        self.n = 1 << (params.log2_k + params.log2_inv_rate)

    @classmethod
    def commit_unmasked(cls, device: Device, params: CommitmentParamsPy, rows):
        k = 1 << params.log2_k
        if isinstance(rows, DeviceBuffer):
            count = rows.count // k
        else:
            rows = np.ascontiguousarray(rows, dtype=np.uint64).reshape(-1)
            if rows.size % k:
                from .lib import BadShape
                raise BadShape(2, "MatrixLength: rows do not hold N*k values")
            count = rows.size // k
        h = vp()
        cp = params.c()
        check(lib().tsr_commit_unmasked(device.handle, ctypes.byref(cp), ptr_of(rows), count,
                                        ctypes.byref(h)))
        return cls(device, h.value, params, count)

    @classmethod
    def commit(cls, device: Device, params: CommitmentParamsPy, segment_rows, rows, masks):
        """ReferenceCommitment::commit: the masked advice commitment of a layout whose segments
        hold segment_rows[s] rows; `rows` (N x k) and `masks` (N x P, m_0 first) in layout order."""
        seg = np.ascontiguousarray(np.asarray(segment_rows, dtype=np.uint64))
        keep = []

        def p(x):
            if isinstance(x, DeviceBuffer):
                return x.ptr
            a = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1)
            keep.append(a)
            return a.ctypes.data if a.size else None
        h = vp()
        cp = params.c()
        check(lib().tsr_commit(device.handle, ctypes.byref(cp), seg.ctypes.data if seg.size else None,
                               seg.size, p(rows), p(masks), ctypes.byref(h)))
        return cls(device, h.value, params, int(seg.sum()))

    @classmethod
    def from_encoded(cls, device: Device, params: CommitmentParamsPy, segment_rows, codewords,
                     merged=True):
        """ReferenceCommitment::from_encoded: steps 2 to 4 from N x n codewords."""
        seg = np.ascontiguousarray(np.asarray(segment_rows, dtype=np.uint64))
        cw = codewords if isinstance(codewords, DeviceBuffer) else \
            np.ascontiguousarray(codewords, dtype=np.uint64).reshape(-1)
        h = vp()
        cp = params.c()
        ptr = cw.ptr if isinstance(cw, DeviceBuffer) else (cw.ctypes.data if cw.size else None)
        check(lib().tsr_commit_from_encoded(device.handle, ctypes.byref(cp),
                                            seg.ctypes.data if seg.size else None, seg.size,
                                            int(merged), ptr, ctypes.byref(h)))
        return cls(device, h.value, params, int(seg.sum()))

    def open(self, groups, xi_point, row_indices, row_vars):
        """ReferenceCommitment::open (commitment.rs:489-524) of a masked commitment: groups =
        [(column_point, [(row_point, gamma), ...]), ...] (Fp2 values as (c0, c1)); the layout's
        row indices in layout order and row_vars. Returns [(u (k x 2), mask (P x 2))], the groups'
        combined rows then the proximity row's."""
        from .lib import BadShape
        log_k = self.params.log2_k
        for g, (cp, claims) in enumerate(groups):
            if len(cp) != log_k:
                raise BadShape(2, f"ColumnPoint: group {g} has {len(cp)} coordinates, not {log_k}")
            for c, (rp, _) in enumerate(claims):
                if len(rp) != row_vars:
                    raise BadShape(2, f"RowPoint: group {g} claim {c} has {len(rp)} coordinates")
        if len(xi_point) != row_vars:
            raise BadShape(2, f"XiPoint: {len(xi_point)} coordinates, not {row_vars}")
        rows = np.ascontiguousarray(np.asarray(row_indices, dtype=np.uint64))
        pts = [np.asarray(rp, dtype=np.uint64).reshape(-1, 2) for _, cl in groups for rp, _ in cl]
        gam = [np.asarray(gm, dtype=np.uint64).reshape(2) for _, cl in groups for _, gm in cl]
        points = np.ascontiguousarray(np.concatenate(pts) if pts else np.zeros((1, 2), np.uint64))
        gammas = np.ascontiguousarray(np.stack(gam) if gam else np.zeros((1, 2), np.uint64))
        sizes = np.ascontiguousarray(np.array([len(cl) for _, cl in groups] or [0], dtype=np.uint32))
        xi = np.ascontiguousarray(np.asarray(xi_point, dtype=np.uint64).reshape(-1, 2)
                                  if row_vars else np.zeros((1, 2), np.uint64))
        count = len(groups) + 1
        coef = np.zeros((count * max(1, rows.size), 2), dtype=np.uint64)
        check(lib().tsr_row_coefficients(self.device.handle, row_vars, rows.ctypes.data if rows.size else None,
                                         rows.size, points.ctypes.data, gammas.ctypes.data,
                                         sizes.ctypes.data, len(groups), xi.ctypes.data, coef.ctypes.data))
        k, P = 1 << log_k, self.params.mask_coeffs
        u = np.zeros((count * k, 2), dtype=np.uint64)
        m = np.zeros((count * max(1, P), 2), dtype=np.uint64)
        check(lib().tsr_commitment_open(self.handle, coef.ctypes.data, count, u.ctypes.data, m.ctypes.data))
        return [(u[g * k:(g + 1) * k], m[g * P:(g + 1) * P]) for g in range(count)]

    def root(self) -> bytes:
        out = ctypes.create_string_buffer(32)
        check(lib().tsr_commitment_root(self.handle, out))
        return out.raw

    def leaves(self):
        out = ctypes.create_string_buffer(32 * self.n)
        check(lib().tsr_commitment_leaves(self.handle, out))
        raw = out.raw
        return [raw[32 * j:32 * j + 32] for j in range(self.n)]

    def codewords_device_ptr(self):
        p = vp()
        cnt = u64()
        n = u64()
        check(lib().tsr_commitment_codewords(self.handle, ctypes.byref(p), ctypes.byref(cnt),
                                             ctypes.byref(n)))
        return p.value, cnt.value, n.value

    def answer(self, queries) -> ColumnAnswer:
        q = np.ascontiguousarray(np.array(queries, dtype=np.uint64))
        cap_d = max(1, int(np.unique(q).size))
        cap_s = 64 * cap_d + 64
        distinct = np.zeros(cap_d, dtype=np.uint64)
        cols = np.zeros(cap_d * max(self.count, 1), dtype=np.uint64)
        sib = ctypes.create_string_buffer(32 * cap_s)
        nd, ns = u64(), u64()
        check(lib().tsr_commitment_answer(self.handle, q.ctypes.data if q.size else None, q.size,
                                          distinct.ctypes.data, cap_d, cols.ctypes.data, sib, cap_s,
                                          ctypes.byref(nd), ctypes.byref(ns)))
        d = nd.value
        # rows of one fresh array (views), and the 32-byte digests split in C
        columns = list(cols[:d * self.count].reshape(d, self.count)) if d else []
        return ColumnAnswer(distinct[:d].tolist(), columns, _split_digests(sib, ns.value))

    def free(self):
        if self.handle:
            lib().tsr_commitment_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
