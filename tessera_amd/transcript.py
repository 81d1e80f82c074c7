"""The Fiat-Shamir transcript (crates/tessera-hash/src/transcript.rs) over the C ABI, and the TCE
encodings the hot path absorbs (crates/tessera-codec/src/tce.rs:306-320,419-441)."""
import ctypes
import struct

import numpy as np

from .lib import check, lib, u64, vp

DOMAINS = ["envelope", "commitment-id", "column-init", "column-step", "merkle-node",
           "column-merge", "weights-root", "transcript-init", "transcript-absorb",
           "transcript-squeeze", "transcript-ratchet", "pow", "coin-commit", "coin-toss", "table",
           "audit-sample", "audit-select", "claim-list", "statement"]


def hash_domain(domain: str, data: bytes, device=None) -> bytes:
    """H(domain, bytes) (crates/tessera-hash/src/domain.rs:177-181)."""
    out = ctypes.create_string_buffer(32)
    check(lib().tsr_hash(device.handle if device else None, DOMAINS.index(domain), data or None,
                         len(data), out))
    return out.raw


def tce_ascii(s: str) -> bytes:
    b = s.encode("ascii")
    return struct.pack("<Q", len(b)) + b


def tce_fp2s(values) -> bytes:
    """TCE of Vec<Fp2>: u64LE length, then c0 || c1 of each element."""
    v = np.ascontiguousarray(np.array(values, dtype=np.uint64).reshape(-1, 2))
    return struct.pack("<Q", v.shape[0]) + v.astype("<u8").tobytes()


def session_context(protocol="tessera-protocol/1", mode="NonInteractive", params="test-small/1",
                    statement_digest=bytes(32), session_id=bytes(16)) -> bytes:
    """TCE(SessionContext) (transcript.rs:40-51): Ascii, u16 mode tag, Ascii, [u8;32], [u8;16]."""
    tag = {"Live": 0, "NonInteractive": 1}[mode]
    return (tce_ascii(protocol) + struct.pack("<H", tag) + tce_ascii(params) + bytes(statement_digest)
            + bytes(session_id))


class Transcript:
    def __init__(self, context_tce: bytes, device=None):
        h = vp()
        check(lib().tsr_transcript_create(device.handle if device else None, context_tce or None,
                                          len(context_tce), ctypes.byref(h)))
        self.handle = h.value

    @classmethod
    def from_state(cls, state: bytes, counter: int, device=None):
        """A transcript continuing another's session from its state and squeeze counter."""
        t = cls.__new__(cls)
        h = vp()
        check(lib().tsr_transcript_from_state(device.handle if device else None, bytes(state),
                                              counter, ctypes.byref(h)))
        t.handle = h.value
        return t

    @property
    def state(self) -> bytes:
        out = ctypes.create_string_buffer(32)
        check(lib().tsr_transcript_state(self.handle, out, None))
        return out.raw

    @property
    def counter(self) -> int:
        c = u64()
        check(lib().tsr_transcript_state(self.handle, None, ctypes.byref(c)))
        return c.value

    def absorb(self, label: str, tce: bytes):
        check(lib().tsr_transcript_absorb(self.handle, label.encode(), tce or None, len(tce)))

    def squeeze(self, label: str) -> bytes:
        out = ctypes.create_string_buffer(32)
        check(lib().tsr_transcript_squeeze(self.handle, label.encode(), out))
        return out.raw

    def challenge_fp(self, label: str) -> int:
        v = u64()
        check(lib().tsr_transcript_challenge_fp(self.handle, label.encode(), ctypes.byref(v)))
        return v.value

    def challenge_fp2(self, label: str):
        # one reusable output pair: an ndarray's .ctypes costs more than the squeeze itself
        out = self.__dict__.get("_fp2_out")
        if out is None:
            out = self._fp2_out = (ctypes.c_uint64 * 2)()
        check(lib().tsr_transcript_challenge_fp2(self.handle, label.encode(), out))
        return (out[0], out[1])

    def challenge_indices(self, label: str, log_n: int, count: int):
        out = np.zeros(max(count, 1), dtype=np.uint64)
        check(lib().tsr_transcript_challenge_indices(self.handle, label.encode(), log_n, count,
                                                     out.ctypes.data))
        return [int(x) for x in out[:count]]

    def grind(self, label: str, bits: int) -> int:
        v = u64()
        check(lib().tsr_transcript_grind(self.handle, label.encode(), bits, ctypes.byref(v)))
        return v.value

    def check_grind(self, label: str, bits: int, nonce: int) -> bool:
        ok = ctypes.c_int(0)
        check(lib().tsr_transcript_check_grind(self.handle, label.encode(), bits, nonce,
                                               ctypes.byref(ok)))
        return bool(ok.value)

    def free(self):
        if self.handle:
            lib().tsr_transcript_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
