"""Pins the oracle's LogUp (fraction trees, tree-level sumchecks, values) against the reference's
logup batches of spec/vectors/sumcheck/sumcheck.json: alpha, beta, the roots, every level's gamma,
lambda, messages, challenges, final evaluations, rho and transcript state, the values and the
final state, with product forests in the schedule where a batch has them; cheating batches whose
recorded messages are not an honest prover's over the given data (a forged roots message) are
checked only up to where the honest prover's output must differ. The tree and forest batches
pin prove_trees over product forests alone (root challenges, gamma, messages, challenges,
evaluations, rho, next claims, states)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, load


def logup_inputs(batch):
    insts = []
    for ins in batch["instances"]:
        rows = [[h(k), h(v)] for k, v in ins["rows"]]
        keys = [r[0] for r in rows]
        vals = [r[1] for r in rows] if ins["has_values"] else None
        lk, lv = [], []
        for look in ins["lookups"]:
            if isinstance(look, int):
                lk.append(keys[look])
                lv.append(rows[look][1])
            else:
                lk.append(h(look[0]))
                lv.append(h(look[1]))
        insts.append({"keys": keys, "values": vals, "lookup_keys": lk,
                      "lookup_values": lv if ins["has_values"] else None,
                      "multiplicities": [h(m) for m in ins["multiplicities"]]})
    n = len(insts)
    sched = [2 * n + t["index"] if t["kind"] == "forest" else 2 * t["index"] + (1 if t["kind"] == "table" else 0)
             for t in batch["trees"]]
    return insts, sched


def batch_forests(batch):
    return [([h(x) for x in f["leaves"]], f["height"]) for f in batch.get("forests", [])]


def honest_batches():
    yield from load("sumcheck.json")["logup_batches"]


def test_logup_batches():
    checked = 0
    for b in honest_batches():
        insts, sched = logup_inputs(b)
        t = O.OracleTranscript(context_from_json(b["context"]))
        out = O.prove_logup(t, insts, sched, batch_forests(b))
        assert out["alpha"] == fp2(b["alpha"]) and out["beta"] == fp2(b["beta"]), b["name"]
        roots_ok = [tuple(int(x) for x in r) for r in out["roots"]] == [fp2(x) for x in b["roots"]]
        if not roots_ok:
            # a cheating prover whose trees or roots message are not an honest prover's over the
            # recorded data (forged roots, or trees over a changed table)
            assert b["expected"] != "accept", b["name"]
            continue
        assert len(out["levels"]) == len(b["levels"])
        for got, want in zip(out["levels"], b["levels"]):
            assert got["root_challenges"] == [fp2(c) for c in want.get("root_challenges", [])]
            assert got["gamma"] == fp2(want["gamma"])
            assert got["lambda"] == (fp2(want["lambda"]) if want["lambda"] is not None else None)
            assert got["rounds"] == [[fp2(v) for v in m] for m in want["rounds"]], (b["name"], want["prefix"])
            assert got["challenges"] == [fp2(c) for c in want["challenges"]]
            assert got["final_evals"] == [fp2(f) for f in want["final_evals"]]
            assert got["rho"] == fp2(want["rho"])
            assert got["state"] == want["state"] and got["counter"] == want["counter"]
        assert [tuple(int(x) for x in v) for v in out["values"]] == [fp2(v) for v in b["values"]]
        assert t.state.hex() == b["state"] and t.counter == b["counter"], b["name"]
        checked += 1
    assert checked >= 13


def test_multiplicities_exact():
    rng = np.random.default_rng(4)
    T, N = 13, 37
    rows = rng.integers(0, T, N)
    m = O.multiplicities(T, rows)
    assert m.size == 16
    want = np.bincount(rows, minlength=16)
    want[0] += 64 - N
    assert np.array_equal(m, want.astype(np.uint64))


def test_root_balance_and_leaf_checks():
    """Honest instances: n_a d_t + n_t d_a = 0, the lookup leaf numerator is 1, and the table's
    leaf denominator is alpha minus the table polynomial at rho_t (the verifier's checks)."""
    rng = np.random.default_rng(8)
    T, N = 11, 29
    keys = rng.integers(0, P, T, dtype=np.uint64)
    vals = rng.integers(0, P, T, dtype=np.uint64)
    rows = rng.integers(0, T, N)
    inst = {"keys": keys, "values": vals, "lookup_keys": keys[rows], "lookup_values": vals[rows],
            "multiplicities": O.multiplicities(T, rows)}
    t = O.OracleTranscript(b"logup")
    out = O.prove_logup(t, [inst], [0, 1])
    na, da, nt, dt = [tuple(int(x) for x in r) for r in out["roots"]]
    s1, s2 = O.fp2_mul(na, dt), O.fp2_mul(nt, da)
    assert ((s1[0] + s2[0]) % P, (s1[1] + s2[1]) % P) == (0, 0)
    cl = [tuple(int(x) for x in r) for r in out["claims"]]
    assert cl[0] == (1, 0)


def check_tree_levels(got_levels, want_levels):
    assert len(got_levels) == len(want_levels)
    for got, want in zip(got_levels, want_levels):
        assert got["root_challenges"] == [fp2(c) for c in want.get("root_challenges", [])], want["prefix"]
        assert got["gamma"] == fp2(want["gamma"]) and got["lambda"] is None
        assert got["rounds"] == [[fp2(v) for v in m] for m in want["rounds"]], want["prefix"]
        assert got["challenges"] == [fp2(c) for c in want["challenges"]]
        assert got["final_evals"] == [fp2(f) for f in want["final_evals"]]
        assert got["rho"] == fp2(want["rho"])
        assert got["state"] == want["state"] and got["counter"] == want["counter"]


TREE_BATCHES = load("sumcheck.json")["tree_batches"]
FOREST_BATCHES = load("sumcheck.json")["forest_batches"]


@pytest.mark.parametrize("batch", TREE_BATCHES, ids=[b["name"][:40] for b in TREE_BATCHES])
def test_tree_batches(batch):
    """prove_trees over single product trees (forests of one tree)."""
    t = O.OracleTranscript(context_from_json(batch["context"]))
    forests = [([h(x) for x in tr["leaves"]], tr["height"]) for tr in batch["trees"]]
    claims, levels = O.prove_trees(t, forests)
    check_tree_levels(levels, batch["levels"])
    # every tree ends at the last level: its next claim is the tree's leaf claim
    last = batch["levels"][-1]["next"]
    assert len(last) == len(claims)
    for n in last:
        assert claims[n["tree"]] == fp2(n["value"])


@pytest.mark.parametrize("batch", FOREST_BATCHES, ids=[b["name"][:40] for b in FOREST_BATCHES])
def test_forest_batches(batch):
    """prove_trees over product forests of 2^s trees (points from the root challenges)."""
    t = O.OracleTranscript(context_from_json(batch["context"]))
    forests = [([h(x) for x in f["leaves"]], f["height"]) for f in batch["forests"]]
    claims, levels = O.prove_trees(t, forests)
    check_tree_levels(levels, batch["levels"])
    last = batch["levels"][-1]["next"]
    assert len(last) == len(claims)
    for n in last:
        assert claims[n["forest"]] == fp2(n["value"])
