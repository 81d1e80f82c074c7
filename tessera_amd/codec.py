"""Wire formats of the proving path (RFC-0003 TCE): SumcheckProof (crates/tessera-sumcheck/src/
proof.rs:9-24) encode and decode with the reference decoder's checks, and the commitment
descriptors' commitment_id (crates/tessera-pcs/src/descriptor.rs:10-160). Host code of the
library: no device is needed."""
import ctypes

import numpy as np

from .lib import BadShape, check, lib, u32, u64

CODEC_ERRORS = {0: "OK", 1: "ERR-CODEC-MAGIC", 2: "ERR-CODEC-VERSION", 3: "ERR-CODEC-NONCANONICAL",
                4: "ERR-CODEC-TRUNCATED", 5: "ERR-CODEC-DIGEST", 6: "ERR-CODEC-LIMIT"}


class DecodeLimits(ctypes.Structure):
    """DecodeLimits (RFC-0003 'Limits'); the defaults are a verifier message's."""
    _fields_ = [("max_total_bytes", u64), ("max_vec_len", u64), ("max_depth", u32)]


DEFAULT_LIMITS = DecodeLimits(8 << 30, 1 << 34, 32)


class CodecError(BadShape):
    """A decoding failure: `code` (ERR-CODEC-...) at `offset`."""

    def __init__(self, code, offset, message):
        super().__init__(2, message)
        self.code, self.offset = code, offset


def encode_sumcheck_proof(rounds, final_evals) -> bytes:
    """TCE(SumcheckProof): rounds (n x d x 2 u64, each message p(0), p(2), ..., p(d)) and the
    final evaluations (m x 2)."""
    r = np.ascontiguousarray(np.asarray(rounds, dtype=np.uint64).reshape(-1))
    f = np.ascontiguousarray(np.asarray(final_evals, dtype=np.uint64).reshape(-1))
    shape = np.asarray(rounds).shape
    n, d = (shape[0], shape[1]) if len(shape) == 3 and shape[0] else (0, 0)
    need = u64()
    lib().tsr_sumcheck_proof_encode(r.ctypes.data, n, d, f.ctypes.data, f.size // 2, None, 0, ctypes.byref(need))
    out = ctypes.create_string_buffer(max(1, need.value))
    check(lib().tsr_sumcheck_proof_encode(r.ctypes.data, n, d, f.ctypes.data, f.size // 2, out,
                                          need.value, ctypes.byref(need)))
    return out.raw[:need.value]


def decode_sumcheck_proof(data: bytes, limits: DecodeLimits = DEFAULT_LIMITS):
    """(rounds: list of lists of (c0, c1), final_evals: list of (c0, c1)); a malformed input
    raises CodecError with the RFC's code and offset."""
    cap = len(data) // 16 + 1
    lens = (u32 * (len(data) // 8 + 1))()
    vals = np.zeros(2 * cap, np.uint64)
    fins = np.zeros(2 * cap, np.uint64)
    nr, nv, nf, err, off = u32(), u64(), u64(), ctypes.c_int(), u64()
    rc = lib().tsr_sumcheck_proof_decode(data, len(data), ctypes.byref(limits), ctypes.byref(nr), lens,
                                         len(lens), vals.ctypes.data, cap, ctypes.byref(nv),
                                         fins.ctypes.data, cap, ctypes.byref(nf), ctypes.byref(err),
                                         ctypes.byref(off))
    if rc and err.value:
        raise CodecError(CODEC_ERRORS[err.value], off.value, lib().tsr_last_error().decode())
    check(rc)
    rounds, at = [], 0
    for j in range(nr.value):
        rounds.append([(int(vals[2 * (at + z)]), int(vals[2 * (at + z) + 1])) for z in range(lens[j])])
        at += lens[j]
    return rounds, [(int(fins[2 * i]), int(fins[2 * i + 1])) for i in range(nf.value)]


class AdviceDescriptor(ctypes.Structure):
    """CommitmentDescriptor::Advice (statement digest, unit, AdviceLayoutParams, rows)."""
    _fields_ = [("statement_digest", ctypes.c_uint8 * 32), ("unit", u32), ("log2_k", ctypes.c_uint8),
                ("log2_inv_rate", ctypes.c_uint8), ("mask_coeffs", u32), ("row_block", u32),
                ("segments", u32), ("rows", u64)]


def advice_commitment_id(statement_digest: bytes, unit, log2_k, log2_inv_rate, mask_coeffs,
                         row_block, segments, rows):
    """(TCE(CommitmentDescriptor::Advice), commitment_id)."""
    d = AdviceDescriptor((ctypes.c_uint8 * 32)(*statement_digest), unit, log2_k, log2_inv_rate,
                         mask_coeffs, row_block, segments, rows)
    tce, cid = ctypes.create_string_buffer(60), ctypes.create_string_buffer(32)
    check(lib().tsr_advice_commitment_id(ctypes.byref(d), tce, cid))
    return tce.raw, cid.raw


def weight_matrix_commitment_id(layout: dict, matrix: int, log2_k, row_block, rows, epoch):
    """(TCE(CommitmentDescriptor::WeightMatrix), commitment_id)."""
    from .weights import _Params
    p = _Params(layout["log2_k"], layout["log2_inv_rate"], layout["mask_coeffs"], layout["blinding_rows"])
    tce, cid = ctypes.create_string_buffer(31), ctypes.create_string_buffer(32)
    check(lib().tsr_weight_matrix_commitment_id(ctypes.byref(p), matrix, log2_k, row_block, rows,
                                                epoch, tce, cid))
    return tce.raw, cid.raw
