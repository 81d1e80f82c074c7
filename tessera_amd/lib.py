"""ctypes binding of libtessera_mi355x.so (the C ABI in include/tessera_mi355x.h).

The library is the product: there is no CPU fallback. Loading fails loudly when the built
library is missing, and contexts fail when no GPU is present.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
# TSR_LIB: another build of the same sources (tests/sanitize.sh loads the one whose host code is
# built with ASan and UBSan)
LIB_PATH = os.environ.get("TSR_LIB") or os.path.join(HERE, "libtessera_mi355x.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "tessera_mi355x.h")

OK, DECLINED, BAD_SHAPE, DEVICE, OOM, BAD_ARG, CANCELLED = range(7)
STATUS_NAMES = {0: "OK", 1: "DECLINED", 2: "BAD_SHAPE", 3: "DEVICE", 4: "OOM", 5: "BAD_ARG", 6: "CANCELLED"}


class TsrError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


class Declined(TsrError):
    """The back end declined the call; the caller computes it on the CPU (no side effects)."""


class BadShape(TsrError):
    pass


u8p = ctypes.POINTER(ctypes.c_uint8)
u64 = ctypes.c_uint64
u32 = ctypes.c_uint32
vp = ctypes.c_void_p
i32 = ctypes.c_int


class CommitmentParams(ctypes.Structure):
    _fields_ = [("log2_k", u32), ("log2_inv_rate", u32), ("mask_coeffs", u64),
                ("row_block", u64), ("commitment_id", ctypes.c_uint8 * 32)]


_SIGS = {
    "tsr_last_error": (ctypes.c_char_p, []),
    "tsr_version": (ctypes.c_char_p, []),
    "tsr_ctx_create": (i32, [i32, ctypes.POINTER(vp)]),
    "tsr_ctx_destroy": (None, [vp]),
    "tsr_ctx_synchronize": (i32, [vp]),
    "tsr_ctx_set_budget": (i32, [vp, u64]),
    "tsr_ctx_memory": (u64, [vp]),
    "tsr_alloc": (i32, [vp, u64, ctypes.POINTER(vp)]),
    "tsr_free": (i32, [vp, vp]),
    "tsr_copy": (i32, [vp, vp, vp, u64]),
    "tsr_fill_uniform": (i32, [vp, vp, u64, u64]),
    "tsr_encode_rows": (i32, [vp, u32, u32, u64, vp, vp, u64, vp]),
    "tsr_ntt_plan_create": (i32, [vp, u64, ctypes.POINTER(vp)]),
    "tsr_ntt_plan_size": (u64, [vp]),
    "tsr_ntt": (i32, [vp, vp, u64]),
    "tsr_intt": (i32, [vp, vp, u64]),
    "tsr_coset_ntt": (i32, [vp, vp, u64, u64]),
    "tsr_coset_intt": (i32, [vp, vp, u64, u64]),
    "tsr_ntt_plan_free": (None, [vp]),
    "tsr_chains_create": (i32, [vp, vp, u64, u64, ctypes.POINTER(vp)]),
    "tsr_chains_absorb": (i32, [vp, vp, u64]),
    "tsr_chains_read": (i32, [vp, vp]),
    "tsr_chains_free": (None, [vp]),
    "tsr_column_sink_create": (i32, [vp, u32, u32, u64, u64, vp, u64, ctypes.POINTER(vp)]),
    "tsr_column_sink_select": (i32, [vp, vp, vp, u64, vp]),
    "tsr_column_sink_combines": (i32, [vp]),
    "tsr_column_sink_select_combining": (i32, [vp, vp, vp, u64, vp, u32, vp]),
    "tsr_column_sink_combined": (i32, [vp, vp, u32]),
    "tsr_column_sink_free": (None, [vp]),
    "tsr_merkle_root": (i32, [vp, vp, u64, vp]),
    "tsr_merkle_root_noalloc": (i32, [vp, vp, u64, vp]),
    "tsr_commit_unmasked": (i32, [vp, ctypes.POINTER(CommitmentParams), vp, u64,
                                  ctypes.POINTER(vp)]),
    "tsr_commitment_root": (i32, [vp, vp]),
    "tsr_commitment_leaves": (i32, [vp, vp]),
    "tsr_commitment_codewords": (i32, [vp, ctypes.POINTER(vp), ctypes.POINTER(u64),
                                       ctypes.POINTER(u64)]),
    "tsr_commitment_answer": (i32, [vp, vp, u64, vp, u64, vp, vp, u64, ctypes.POINTER(u64),
                                    ctypes.POINTER(u64)]),
    "tsr_commitment_free": (None, [vp]),
    "tsr_hash": (i32, [vp, i32, vp, u64, vp]),
    "tsr_transcript_create": (i32, [vp, vp, u64, ctypes.POINTER(vp)]),
    "tsr_transcript_free": (None, [vp]),
    "tsr_transcript_state": (i32, [vp, vp, ctypes.POINTER(u64)]),
    "tsr_transcript_absorb": (i32, [vp, ctypes.c_char_p, vp, u64]),
    "tsr_transcript_squeeze": (i32, [vp, ctypes.c_char_p, vp]),
    "tsr_transcript_challenge_fp": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(u64)]),
    "tsr_transcript_challenge_fp2": (i32, [vp, ctypes.c_char_p, vp]),
    "tsr_transcript_challenge_indices": (i32, [vp, ctypes.c_char_p, u32, u32, vp]),
    "tsr_transcript_grind": (i32, [vp, ctypes.c_char_p, ctypes.c_uint8, ctypes.POINTER(u64)]),
    "tsr_transcript_check_grind": (i32, [vp, ctypes.c_char_p, ctypes.c_uint8, u64,
                                         ctypes.POINTER(i32)]),
    "tsr_sumcheck_prove": (i32, [vp, vp, ctypes.c_char_p, u32, vp, u32, vp, vp, vp, vp, u64,
                                 ctypes.POINTER(u64)]),
    "tsr_logup_prove": (i32, [vp, vp, vp, u32, vp, u32, vp, vp, vp, vp, vp, ctypes.POINTER(u64),
                              vp, vp, vp, u64, ctypes.POINTER(u64)]),
    "tsr_prove_trees": (i32, [vp, vp, vp, u32, vp, vp, u64, ctypes.POINTER(u64)]),
    "tsr_logup_multiplicities": (i32, [vp, u64, vp, u64, vp]),
    "tsr_commit_batch_create": (i32, [vp, vp, u32, u32, ctypes.POINTER(vp)]),
    "tsr_commit_batch_run": (i32, [vp]),
    "tsr_commit_batch_roots": (i32, [vp, vp]),
    "tsr_commit_batch_tuning": (i32, [vp, vp, u32, ctypes.POINTER(u32)]),
    "tsr_commit_batch_free": (None, [vp]),
    "tsr_eq_table": (i32, [vp, vp, u32, vp]),
    "tsr_commit": (i32, [vp, vp, vp, u32, vp, vp, ctypes.POINTER(vp)]),
    "tsr_row_coefficients": (i32, [vp, u32, vp, u64, vp, vp, vp, u32, vp, vp]),
    "tsr_commitment_open": (i32, [vp, vp, u32, vp, vp]),
    "tsr_commit_from_encoded": (i32, [vp, vp, vp, u32, i32, vp, ctypes.POINTER(vp)]),
    "tsr_transcript_from_state": (i32, [vp, vp, u64, ctypes.POINTER(vp)]),
    "tsr_block_sink_create": (i32, [vp, u32, u32, u64, u64, vp, ctypes.POINTER(vp)]),
    "tsr_block_sink_absorb": (i32, [vp, vp, u64, vp]),
    "tsr_block_sink_finish": (i32, [vp, vp]),
    "tsr_block_sink_free": (None, [vp]),
    "tsr_weight_commit": (i32, [vp, vp, u32, vp, vp, u32, u32, u32, vp, u32, vp, u32,
                                ctypes.POINTER(vp)]),
    "tsr_weight_commitment_public": (i32, [vp, vp]),
    "tsr_weight_commitment_weights_root": (i32, [vp, vp]),
    "tsr_weight_commitment_artifact": (i32, [vp, vp]),
    "tsr_weight_commitment_matrix": (i32, [vp, u32, ctypes.POINTER(vp), ctypes.POINTER(u32),
                                           ctypes.POINTER(u32), ctypes.POINTER(u64)]),
    "tsr_weight_commitment_free": (None, [vp]),
    "tsr_envelope_seal": (i32, [vp, ctypes.c_uint16, ctypes.c_uint16, vp, u64, vp, u64,
                                ctypes.POINTER(u64)]),
    "tsr_round3_encode": (i32, [vp, u64, u64, vp, u64, vp, u32, vp, u64, ctypes.POINTER(u64)]),
    "tsr_round3_decode": (i32, [ctypes.c_char_p, u64, vp, ctypes.POINTER(u64), ctypes.POINTER(u32),
                                ctypes.POINTER(i32), ctypes.POINTER(u64)]),
    "tsr_column_queries_encode": (i32, [vp, u64, vp, vp, vp, u32, vp, u64, ctypes.POINTER(u64)]),
    "tsr_column_queries_decode": (i32, [ctypes.c_char_p, u64, vp, vp, u64, ctypes.POINTER(u64), vp, vp, u32,
                                        ctypes.POINTER(u32), vp, u64, ctypes.POINTER(u64), ctypes.POINTER(i32),
                                        ctypes.POINTER(u64)]),
    "tsr_sumcheck_proof_encode": (i32, [vp, u32, u32, vp, u64, vp, u64, ctypes.POINTER(u64)]),
    "tsr_sumcheck_proof_decode": (i32, [vp, u64, vp, ctypes.POINTER(u32), vp, u32, vp, u64,
                                        ctypes.POINTER(u64), vp, u64, ctypes.POINTER(u64),
                                        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(u64)]),
    "tsr_advice_commitment_id": (i32, [vp, vp, vp]),
    "tsr_weight_matrix_commitment_id": (i32, [vp, u32, ctypes.c_uint8, u32, u64, u32, vp, vp]),
    "tsr_mask_stream_create": (i32, [vp, vp, ctypes.POINTER(vp)]),
    "tsr_mask_stream_draw": (i32, [vp, u64, vp]),
    "tsr_mask_stream_position": (u64, [vp]),
    "tsr_mask_stream_free": (None, [vp]),
    "tsr_matrix_sink_create": (i32, [vp, u32, u32, u64, vp, u64, vp, u64, ctypes.POINTER(vp)]),
    "tsr_matrix_sink_absorb": (i32, [vp, vp, u64, vp]),
    "tsr_matrix_sink_finish": (i32, [vp, vp]),
    "tsr_matrix_sink_free": (None, [vp]),
    "tsr_level_load": (i32, [vp, i32, vp, vp, u32, u32, ctypes.POINTER(vp)]),
    "tsr_level_sums": (i32, [vp, vp, u32, vp, u64, vp]),
    "tsr_level_bind": (i32, [vp, vp]),
    "tsr_level_firsts": (i32, [vp, vp]),
    "tsr_level_free": (None, [vp]),
    "tsr_butterfly_peak": (i32, [vp, ctypes.POINTER(ctypes.c_double)]),
    "tsr_server_create": (i32, [i32, vp, ctypes.POINTER(vp)]),
    "tsr_server_submit": (i32, [vp, vp]),
    "tsr_server_cancel": (i32, [vp, u64]),
    "tsr_server_wait": (i32, [vp, vp, u32]),
    "tsr_server_stats": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u64)]),
    "tsr_server_prove": (i32, [vp, vp, vp]),
    "tsr_server_open": (i32, [vp, u64, vp, u64, vp, u64, vp, vp, u64, ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    "tsr_server_release": (i32, [vp, u64]),
    "tsr_server_resident": (i32, [vp, ctypes.POINTER(u32), ctypes.POINTER(u64)]),
    "tsr_server_footprint": (u64, [u32]),
    "tsr_server_clock": (ctypes.c_double, [vp]),
    "tsr_server_destroy": (None, [vp]),
    "tsr_ctx_profile": (i32, [vp, i32]),
    "tsr_ctx_set_round_blocks": (i32, [vp, u32]),
    "tsr_ctx_kernel_time": (i32, [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(u64)]),
}

_lib = None


def declared_symbols():
    """Every function the public header declares."""
    with open(HEADER) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(tsr_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built: run `python -m tessera_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status == OK:
        return
    msg = lib().tsr_last_error().decode(errors="replace")
    if status == DECLINED:
        raise Declined(status, msg)
    if status == BAD_SHAPE:
        raise BadShape(status, msg)
    raise TsrError(status, msg)
