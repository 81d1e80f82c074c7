"""The library's host-side transcript and hash (no GPU needed) against the reference's vectors:
every step of the 6 transcript sessions, the 19 domains, and long absorbs (multi-chunk BLAKE3)
against the oracle."""
import numpy as np

from oracle import oracle as O
from tests.helpers import load


def test_transcript_sessions():
    from tessera_amd.transcript import Transcript
    for case in load("transcript.json")["cases"]:
        t = Transcript(bytes.fromhex(case["context_tce"]))
        assert t.state.hex() == case["state"]
        for s in case["steps"]:
            op, label = s["op"], s["label"]
            if op == "absorb":
                t.absorb(label, bytes.fromhex(s["tce"]))
            elif op == "squeeze":
                assert t.squeeze(label).hex() == s["out"]
            elif op == "challenge_fp":
                assert str(t.challenge_fp(label)) == s["value"]
            elif op == "challenge_fp2":
                assert [str(x) for x in t.challenge_fp2(label)] == s["value"]
            elif op == "challenge_indices":
                assert [str(x) for x in t.challenge_indices(label, s["log_n"], s["count"])] == s["value"]
            elif op == "grind":
                assert str(t.grind(label, s["bits"])) == s["nonce"]
            elif op == "check_grind":
                assert t.check_grind(label, s["bits"], int(s["nonce"])) == s["valid"]
            assert t.state.hex() == s["state"], (case["name"], op, label)
            assert t.counter == s["counter"]


def test_domain_digests():
    from tessera_amd.transcript import DOMAINS, hash_domain
    d = load("hash.json")
    for dom in d["domains"]:
        for length, want in zip(d["input_lengths"], dom["digests"]):
            data = bytes(j % 251 for j in range(length))
            assert hash_domain(dom["name"], data).hex() == want


def test_long_absorbs_match_oracle():
    from tessera_amd.transcript import Transcript
    rng = np.random.default_rng(1)
    ctx = b"\x01" * 90
    a, b = Transcript(ctx), O.OracleTranscript(ctx)
    for length in (0, 1, 63, 64, 65, 1023, 1024, 1025, 2048, 3000, 5000, 70000):
        data = rng.integers(0, 256, length, dtype=np.uint8).tobytes()
        a.absorb("x", data)
        b.absorb("x", data)
        assert a.state == b.state, length
        assert a.squeeze("y") == b.squeeze("y")


def test_transcript_from_state_continues_the_session():
    """A transcript handed over as (state, counter) continues exactly as the original."""
    from tessera_amd.transcript import Transcript
    a = Transcript(b"hand-over")
    a.absorb("x", b"\x01\x02")
    a.challenge_fp2("c")
    b = Transcript.from_state(a.state, a.counter)
    for t in (a, b):
        t.absorb("y", b"\x03")
    assert a.challenge_fp2("d") == b.challenge_fp2("d")
    assert a.state == b.state and a.counter == b.counter
