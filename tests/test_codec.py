"""The proving path's wire formats in the library's host code (no device): TCE(SumcheckProof)
against a small TCE writer that the reference's codec goldens pin (Vec<u32>, Vec<Vec<i16>>, Fp,
Fp2, Vec<Fp>), over every recorded stage of the sumcheck sessions; the decoder's round trips
(INV-TCE-1/2), its codes and offsets (every proper prefix truncated, the non-canonical Fp2
cases of the noncanonical corpus inside a proof, limits, trailing bytes); and the commitment
descriptors' TCE and commitment_id against spec/vectors/hash commitment_ids, the streaming
commitments and the weight matrices."""
import numpy as np
import pytest

from tests.helpers import fp2, h, load

P = 0xFFFFFFFF00000001


def u64(v):
    return int(v).to_bytes(8, "little")


def tce(value, typ):
    """An independent TCE writer for the types the goldens below pin."""
    if typ in ("u32",):
        return int(value).to_bytes(4, "little")
    if typ == "i16":
        return (int(value) & 0xFFFF).to_bytes(2, "little")
    if typ in ("u64", "Fp"):
        return u64(value)
    if typ == "Fp2":
        return u64(value[0]) + u64(value[1])
    if typ.startswith("Vec<") and typ.endswith(">"):
        inner = typ[4:-1]
        return u64(len(value)) + b"".join(tce(v, inner) for v in value)
    raise ValueError(typ)


def test_writer_against_codec_goldens():
    seen = set()
    for case in load("codec_golden.json")["cases"]:
        typ = case["type"]
        if typ not in ("Vec<u32>", "Vec<Vec<i16>>", "Fp", "Fp2", "Vec<Fp>"):
            continue
        assert tce(case["value"], typ).hex() == case["hex"], typ
        seen.add(typ)
    assert len(seen) == 5


def session_proofs():
    for sess in load("sumcheck.json")["sessions"]:
        for st in sess["stages"]:
            rounds = [[fp2(v) for v in m] for m in st["rounds"]]
            yield st["degree"], rounds, [fp2(f) for f in st["final_evals"]]


def test_sumcheck_proofs_encode_and_round_trip():
    from tessera_amd.codec import decode_sumcheck_proof, encode_sumcheck_proof
    n = 0
    for degree, rounds, finals in session_proofs():
        want = tce(rounds, "Vec<Vec<Fp2>>") + tce(finals, "Vec<Fp2>")
        arr = np.array(rounds, dtype=np.uint64).reshape(len(rounds), degree, 2) if rounds else np.zeros((0, degree, 2), np.uint64)
        got = encode_sumcheck_proof(arr, np.array(finals, dtype=np.uint64).reshape(-1, 2))
        assert got == want
        r, f = decode_sumcheck_proof(got)
        assert r == rounds and f == finals
        n += 1
    assert n >= 40


def test_every_proper_prefix_is_truncated():
    from tessera_amd.codec import CodecError, decode_sumcheck_proof
    data = tce([[(1, 2), (3, 4)], [(5, 6), (7, 8)]], "Vec<Vec<Fp2>>") + tce([(9, 10)], "Vec<Fp2>")
    assert decode_sumcheck_proof(data)[1] == [(9, 10)]
    for cut in range(len(data)):
        with pytest.raises(CodecError) as e:
            decode_sumcheck_proof(data[:cut])
        assert e.value.code == "ERR-CODEC-TRUNCATED"


def test_noncanonical_fp2_cases_inside_a_proof():
    """The corpus' non-canonical Fp2 inputs as the one value of one round: the same code, at the
    offset moved by the two lengths before it."""
    from tessera_amd.codec import CodecError, decode_sumcheck_proof
    n = 0
    for case in load("codec_noncanonical.json")["cases"]:
        if case["type"] != "Fp2" or case["code"] != "ERR-CODEC-NONCANONICAL":
            continue
        data = u64(1) + u64(1) + bytes.fromhex(case["hex"]) + u64(0)
        with pytest.raises(CodecError) as e:
            decode_sumcheck_proof(data)
        assert (e.value.code, e.value.offset) == (case["code"], int(case["offset"]) + 16)
        n += 1
    assert n == 2


def test_limits_and_trailing_bytes():
    from tessera_amd.codec import CodecError, DecodeLimits, decode_sumcheck_proof
    data = tce([[(1, 2)], [(3, 4)], [(5, 6)]], "Vec<Vec<Fp2>>") + tce([], "Vec<Fp2>")

    def err(d, limits=DecodeLimits(1 << 20, 1 << 20, 32)):
        with pytest.raises(CodecError) as e:
            decode_sumcheck_proof(d, limits)
        return e.value.code, e.value.offset
    assert err(data, DecodeLimits(len(data) - 1, 100, 32)) == ("ERR-CODEC-LIMIT", 0)
    assert err(data, DecodeLimits(1 << 20, 2, 32)) == ("ERR-CODEC-LIMIT", 0)  # three rounds
    # SumcheckProof, the rounds Vec, a round's Vec, then its first Fp2 at depth 4
    assert err(data, DecodeLimits(1 << 20, 100, 3)) == ("ERR-CODEC-LIMIT", 16)
    assert decode_sumcheck_proof(data, DecodeLimits(len(data), 3, 4))[0] == [[(1, 2)], [(3, 4)], [(5, 6)]]
    assert err(data + b"\x00") == ("ERR-CODEC-NONCANONICAL", len(data))
    # a length whose items cannot fit: truncated at the length's first byte
    assert err(u64(5) + u64(1) + tce((1, 1), "Fp2") + u64(0)) == ("ERR-CODEC-TRUNCATED", 0)
    # two rounds fit the length check; the second round's length then takes the finals' bytes
    assert err(u64(2) + u64(1) + tce((1, 1), "Fp2") + u64(0)) == ("ERR-CODEC-TRUNCATED", 40)
    # a non-canonical final evaluation's c1
    bad = tce([], "Vec<Vec<Fp2>>") + u64(1) + u64(5) + u64(P)
    assert err(bad) == ("ERR-CODEC-NONCANONICAL", 24)


def test_advice_commitment_ids():
    from tessera_amd.codec import advice_commitment_id
    n = 0
    for c in load("hash.json")["commitment_ids"]:
        d = c["descriptor"]
        if d["kind"] != "Advice":
            continue
        L = d["layout"]
        tce_, cid = advice_commitment_id(bytes.fromhex(d["statement_digest"]), d["unit"], L["log2_k"],
                                         L["log2_inv_rate"], L["mask_coeffs"], L["row_block"],
                                         L["segments"], d["rows"])
        assert tce_.hex() == c["tce"] and cid.hex() == c["commitment_id"]
        n += 1
    for c in load("pcs_streaming.json")["cases"]:
        log_k = c["layout"]["log2_k"]
        log_r = len(c["leaves"]).bit_length() - 1 - log_k
        tce_, cid = advice_commitment_id(bytes.fromhex(c["statement_digest"]), c["unit"], log_k, log_r,
                                         c["mask_coeffs"], c["row_block"], len(c["segments"]),
                                         sum(len(s) for s in c["segments"]))
        assert tce_.hex() == c["descriptor_tce"] and cid.hex() == c["commitment_id"], c["name"]
        n += 1
    assert n >= 10


def test_weight_matrix_commitment_ids():
    from tessera_amd.codec import weight_matrix_commitment_id
    names = ["Embed", "Q", "K", "V", "O", "Gate", "Up", "Down", "LmHead", "ChannelParams"]
    n = 0
    for c in load("hash.json")["commitment_ids"]:
        d = c["descriptor"]
        if d["kind"] != "WeightMatrix":
            continue
        tce_, cid = weight_matrix_commitment_id(d["layout"], names.index(d["matrix"]), d["log2_k"],
                                                d["row_block"], d["rows"], d["epoch"])
        assert tce_.hex() == c["tce"] and cid.hex() == c["commitment_id"]
        n += 1
    w = load("weights.json")
    for m in w["models"]:
        for t, mat in enumerate(m["matrices"]):
            tce_, cid = weight_matrix_commitment_id(w["layout"], t, mat["log2_k"], w["row_block"],
                                                    mat["rows"], w["epoch"])
            assert tce_.hex() == mat["descriptor"] and cid.hex() == mat["commitment_id"]
            n += 1
    assert n == 52


# ---- Round3Response and ColumnQueries (crates/tessera-protocol/src/round3.rs:19-175) against
# spec/vectors/protocol/protocol.json: the payload bytes of the recorded messages, every
# truncation's code and offset, every refusal, and the accepted limits.

PROTO = load("protocol.json")


def _limits(d):
    from tessera_amd.codec import DEFAULT_LIMITS, DecodeLimits
    if not d:
        return DEFAULT_LIMITS
    return DecodeLimits(int(d["max_total_bytes"]), int(d["max_vec_len"]), int(d["max_depth"]))


def _messages(typ):
    return [m for m in PROTO["messages"] if m["type"] == typ]


def test_round3_payloads_from_values():
    from tessera_amd.codec import decode_round3, encode_round3
    msgs = _messages("Round3Response")
    assert msgs
    for m in msgs:
        v = m["value"]
        adv = [[int(x) for x in col] for col in v["advice_columns"]]
        sib = [bytes.fromhex(s) for s in v["advice_proof"]["siblings"]]
        ws = [(int(w["matrix"]), [[int(x) for x in col] for col in w["columns"]],
               [bytes.fromhex(s) for s in w["proof"]["siblings"]]) for w in v["weight_columns"]]
        got = encode_round3(adv, sib, ws)
        assert got.hex() == m["payload"], m["name"]
        assert decode_round3(got) == (len(adv), len(ws))


def test_column_queries_payloads_from_values():
    from tessera_amd.codec import decode_column_queries, encode_column_queries
    msgs = _messages("ColumnQueries")
    assert len(msgs) == 2
    for m in msgs:
        v = m["value"]
        adv = [int(x) for x in v["advice"]]
        ws = [(int(w["matrix"]), [int(x) for x in w["indices"]]) for w in v["weights"]]
        got = encode_column_queries(adv, ws)
        assert got.hex() == m["payload"], m["name"]
        assert decode_column_queries(got) == (adv, ws)


def _decode(typ, data, limits):
    from tessera_amd.codec import decode_column_queries, decode_round3
    return (decode_round3 if typ == "Round3Response" else decode_column_queries)(data, limits)


@pytest.mark.parametrize("typ", ["Round3Response", "ColumnQueries"])
def test_truncations(typ):
    from tessera_amd.codec import CodecError
    for m in _messages(typ):
        payload = bytes.fromhex(m["payload"])
        for cut, offset in m["truncated"]:
            with pytest.raises(CodecError) as e:
                _decode(typ, payload[:cut], _limits(None))
            assert (e.value.code, e.value.offset) == ("ERR-CODEC-TRUNCATED", offset), (m["name"], cut)


@pytest.mark.parametrize("typ", ["Round3Response", "ColumnQueries"])
def test_refused_and_accepted(typ):
    from tessera_amd.codec import CodecError
    refused = [r for r in PROTO["refused"] if r["type"] == typ]
    assert len(refused) >= 9
    for r in refused:
        with pytest.raises(CodecError) as e:
            _decode(typ, bytes.fromhex(r["hex"]), _limits(r.get("limits")))
        assert (e.value.code, e.value.offset) == (r["code"], r["offset"]), r["why"]
    accepted = [a for a in PROTO["accepted"] if a["type"] == typ]
    assert accepted
    for a in accepted:
        _decode(typ, bytes.fromhex(a["hex"]), _limits(a.get("limits")))


def test_round3_encoder_refuses_bad_input():
    from tessera_amd.codec import encode_round3
    from tessera_amd.lib import TsrError
    with pytest.raises(TsrError):
        encode_round3([[P]], [])  # a symbol that is not canonical
    with pytest.raises(TsrError):
        encode_round3([[1]], [], [(3, [[1]], []), (3, [[2]], [])])  # repeated matrix id
    with pytest.raises(TsrError):
        encode_round3([[1]], [], [(10, [[1]], [])])  # no such matrix
