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


class MatrixColumns(ctypes.Structure):
    """MatrixColumns (crates/tessera-protocol/src/round3.rs:150-175) as tsr_round3_encode takes it."""
    _fields_ = [("matrix", ctypes.c_uint8), ("columns", u64), ("symbols", u64), ("values", ctypes.c_void_p),
                ("siblings", u64), ("proof", ctypes.c_void_p)]


def _columns_array(columns):
    cols = [list(c) for c in columns]
    syms = len(cols[0]) if cols else 0
    if any(len(c) != syms for c in cols):
        raise ValueError("every answered column holds the same number of symbols")
    return np.ascontiguousarray(np.array(cols, dtype=np.uint64).reshape(-1)), len(cols), syms


def encode_round3(advice_columns, advice_siblings, weights=()) -> bytes:
    """TCE(Round3Response): the advice's answered columns (each a list of symbols, one per
    committed row) and its multiproof siblings (32-byte digests), then (matrix id, columns,
    siblings) per opened weight matrix in increasing id order."""
    keep = []
    av, ac, asym = _columns_array(advice_columns)
    ap = b"".join(advice_siblings)
    ws = (MatrixColumns * max(1, len(weights)))()
    for i, (matrix, cols, sib) in enumerate(weights):
        v, c, sy = _columns_array(cols)
        pr = ctypes.create_string_buffer(b"".join(sib), max(1, 32 * len(sib)))
        keep += [v, pr]
        ws[i] = MatrixColumns(matrix, c, sy, v.ctypes.data if v.size else None, len(sib), ctypes.addressof(pr))
    need = u64()
    args = (av.ctypes.data if av.size else None, ac, asym, ap or None, len(advice_siblings), ws, len(weights))
    lib().tsr_round3_encode(*args, None, 0, ctypes.byref(need))
    out = ctypes.create_string_buffer(max(1, need.value))
    check(lib().tsr_round3_encode(*args, out, need.value, ctypes.byref(need)))
    return out.raw[:need.value]


def round3_from_answers(advice, weights=()) -> bytes:
    """Round3Response bytes from ColumnAnswers: the advice commitment's, then (matrix id,
    answer) per opened weight matrix."""
    return encode_round3(advice.columns, advice.siblings,
                         [(m, a.columns, a.siblings) for m, a in weights])


def decode_round3(data: bytes, limits: DecodeLimits = DEFAULT_LIMITS):
    """Checks a Round3Response: (advice columns, opened matrices), or CodecError."""
    cols, mats, err, off = u64(), u32(), ctypes.c_int(), u64()
    rc = lib().tsr_round3_decode(data, len(data), ctypes.byref(limits), ctypes.byref(cols), ctypes.byref(mats),
                                 ctypes.byref(err), ctypes.byref(off))
    if rc and err.value:
        raise CodecError(CODEC_ERRORS[err.value], off.value, lib().tsr_last_error().decode())
    check(rc)
    return cols.value, mats.value


def encode_column_queries(advice, weights=()) -> bytes:
    """TCE(ColumnQueries): the advice indices, then (matrix id, indices) per opened matrix."""
    a = np.ascontiguousarray(np.asarray(advice, dtype=np.uint32))
    ids = np.array([m for m, _ in weights], dtype=np.uint8)
    counts = np.array([len(q) for _, q in weights], dtype=np.uint64)
    idx = np.ascontiguousarray(np.concatenate([np.asarray(q, dtype=np.uint32) for _, q in weights])
                               if weights else np.zeros(0, np.uint32))
    args = (a.ctypes.data if a.size else None, a.size, ids.ctypes.data if ids.size else None,
            counts.ctypes.data if counts.size else None, idx.ctypes.data if idx.size else None, len(weights))
    need = u64()
    lib().tsr_column_queries_encode(*args, None, 0, ctypes.byref(need))
    out = ctypes.create_string_buffer(max(1, need.value))
    check(lib().tsr_column_queries_encode(*args, out, need.value, ctypes.byref(need)))
    return out.raw[:need.value]


def decode_column_queries(data: bytes, limits: DecodeLimits = DEFAULT_LIMITS):
    """(advice indices, [(matrix id, indices)]) of a ColumnQueries, or CodecError."""
    cap = len(data) // 4 + 1
    adv = np.zeros(cap, np.uint32)
    ids = np.zeros(cap, np.uint8)
    cnt = np.zeros(cap, np.uint64)
    idx = np.zeros(cap, np.uint32)
    na, nm, ni, err, off = u64(), u32(), u64(), ctypes.c_int(), u64()
    rc = lib().tsr_column_queries_decode(data, len(data), ctypes.byref(limits), adv.ctypes.data, cap,
                                         ctypes.byref(na), ids.ctypes.data, cnt.ctypes.data, cap,
                                         ctypes.byref(nm), idx.ctypes.data, cap, ctypes.byref(ni),
                                         ctypes.byref(err), ctypes.byref(off))
    if rc and err.value:
        raise CodecError(CODEC_ERRORS[err.value], off.value, lib().tsr_last_error().decode())
    check(rc)
    out, at = [], 0
    for m in range(nm.value):
        c = int(cnt[m])
        out.append((int(ids[m]), [int(x) for x in idx[at:at + c]]))
        at += c
    return [int(x) for x in adv[:na.value]], out
