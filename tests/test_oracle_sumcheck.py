"""Pins the oracle's multilinear helpers and batched sumcheck against the reference's vectors:
spec/vectors/poly (MLE values, eq tables, round polynomials) and spec/vectors/sumcheck sessions
(every stage that records its tables is proven and compared message by message, challenge by
challenge, final evaluation and transcript state; the others are replayed verifier-side, which
pins labels, encodings and claims). Base-field tables are checked against their Fp2 embedding."""
import numpy as np

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, load


def f2s(xs):
    return np.array([[h(a), h(b)] for a, b in xs], dtype=np.uint64)


def test_poly_vectors():
    d = load("poly.json")
    for case in d["multilinear"]:
        ext = case["field"] != "Fp"
        tab = f2s(case["table"]) if ext else np.array([h(x) for x in case["table"]], dtype=np.uint64)
        assert O.mle_eval(tab, ext, f2s(case["point"])) == fp2(case["value"])
    for case in d["eq_tables"]:
        assert [tuple(int(x) for x in r) for r in O.eq_table(f2s(case["point"]))] == \
            [fp2(x) for x in case["table"]]
    for case in d["round_polys"]:
        ev = f2s(case["evaluations"])
        assert O.round_eval(ev, case["degree"], fp2(case["claim"]), fp2(case["r"])) == fp2(case["value"])


def replay_stage(t, st):
    """The verifier's transcript moves for a recorded stage (gamma, rounds, challenges, evals),
    checking every claim p_j(0) + p_j(1) and the final relation."""
    prefix, d = st["prefix"], st["degree"]
    g, _ = t.challenge_fp2(prefix + "/gamma")
    assert g == fp2(st["gamma"])
    claim = fp2(st["claims"][0])
    for j, msg in enumerate(st["rounds"]):
        t.absorb_fp2s(f"{prefix}/round/{j}", f2s(msg))
        r, _ = t.challenge_fp2(f"{prefix}/r/{j}")
        assert r == fp2(st["challenges"][j])
        claim = O.round_eval(f2s(msg), d, claim, r)
        assert claim == fp2(st["claims"][j + 1])
    t.absorb_fp2s(f"{prefix}/evals", f2s(st["final_evals"]))


def test_sumcheck_sessions():
    proven = 0
    for sess in load("sumcheck.json")["sessions"]:
        t = O.OracleTranscript(context_from_json(sess["context"]))
        for st in sess["stages"]:
            if all("tables" in inst for inst in st["instances"]):
                insts = [[(f2s(tab), True, False) for tab in inst["tables"]] for inst in st["instances"]]
                out = O.prove_batch(t, st["prefix"], st["degree"], insts)
                assert out["gamma"] == fp2(st["gamma"])
                assert out["claim"] == fp2(st["claims"][0])
                assert [[tuple(int(x) for x in v) for v in m] for m in out["rounds"]] == \
                    [[fp2(v) for v in m] for m in st["rounds"]]
                assert [tuple(int(x) for x in c) for c in out["challenges"]] == [fp2(c) for c in st["challenges"]]
                assert [tuple(int(x) for x in f) for f in out["finals"]] == [fp2(f) for f in st["final_evals"]]
                proven += 1
            else:
                replay_stage(t, st)
            assert t.state.hex() == st["state"], (sess["name"], st["prefix"])
            assert t.counter == st["counter"]
    assert proven >= 31


def test_base_tables_equal_their_embedding_and_public_factors_are_omitted():
    rng = np.random.default_rng(9)
    n = 7
    a = rng.integers(0, P, 1 << n, dtype=np.uint64)
    b = rng.integers(0, P, 1 << n, dtype=np.uint64)
    emb = lambda v: np.stack([v, np.zeros_like(v)], axis=1)
    t1 = O.OracleTranscript(b"ctx")
    t2 = O.OracleTranscript(b"ctx")
    x = O.prove_batch(t1, "stage/0", 3, [[(a, False, False), (b, False, True), (a, False, False)]])
    y = O.prove_batch(t2, "stage/0", 3, [[(emb(a), True, False), (emb(b), True, True), (emb(a), True, False)]])
    assert np.array_equal(x["rounds"], y["rounds"]) and np.array_equal(x["finals"], y["finals"])
    assert x["finals"].shape[0] == 2 and t1.state == t2.state
    # the final evaluations are the tables' values at the point
    assert tuple(int(v) for v in x["finals"][0]) == O.mle_eval(a, False, x["challenges"])


def test_sumcheck_round_sums_brute_force():
    """p_j(0) + p_j(1) equals the running claim and the last claim equals the product of the
    final evaluations (the verifier's checks, tv-core sumcheck.rs:306-398)."""
    rng = np.random.default_rng(3)
    for n, k, d in [(1, 1, 1), (3, 2, 2), (6, 3, 3), (9, 2, 4)]:
        tabs = [(rng.integers(0, P, 1 << n, dtype=np.uint64), False, False) for _ in range(k)]
        t = O.OracleTranscript(b"x")
        out = O.prove_batch(t, "stage/1", d, [tabs])
        claim = out["claim"]
        direct = 0
        prod = np.ones(1 << n, dtype=object)
        for v, _, _ in tabs:
            prod = prod * np.array([int(x) for x in v], dtype=object) % P
        assert claim == (int(sum(prod) % P), 0)
        for j in range(n):
            claim = O.round_eval(out["rounds"][j], d, claim, out["challenges"][j])
        final = (1, 0)
        for f in out["finals"]:
            final = O.fp2_mul(final, tuple(int(x) for x in f))
        assert claim == final
