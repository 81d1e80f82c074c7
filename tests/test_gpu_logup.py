"""Parity of the gfx950 LogUp (device-built fraction trees and product forests, level sumchecks
with the device transcript) with the reference's logup, tree and forest batches and with the
oracle: alpha, beta, roots, every level's root challenges, gamma, lambda, messages, challenges,
evaluations, rho and transcript state, values, leaf and forest claims; the multiplicity counts;
and the 2^22-lookup / 2^16-row Zipf(1.1) workload whole against the oracle."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, load, logup_config3
from tests.test_oracle_logup import (FOREST_BATCHES, TREE_BATCHES, batch_forests,
                                     check_tree_levels, honest_batches, logup_inputs)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def to_gpu_instances(insts):
    from tessera_amd.logup import LogUpInstance
    return [LogUpInstance(i["keys"], i["values"], i["lookup_keys"], i["lookup_values"],
                          i["multiplicities"]) for i in insts]


def to_gpu_forests(forests):
    from tessera_amd.logup import Forest
    return [Forest(np.asarray(leaves, dtype=np.uint64), height) for leaves, height in forests]


def as_tuples(a):
    return [tuple(int(x) for x in v) for v in np.asarray(a).reshape(-1, 2)]


def test_reference_logup_batches(dev):
    from tessera_amd.logup import prove_logup
    from tessera_amd.transcript import Transcript
    checked = 0
    for b in honest_batches():
        insts, sched = logup_inputs(b)
        # the oracle decides which batches an honest prover reproduces (forged roots aside)
        forests = batch_forests(b)
        ot = O.OracleTranscript(context_from_json(b["context"]))
        want = O.prove_logup(ot, insts, sched, forests)
        if as_tuples(want["roots"]) != [fp2(x) for x in b["roots"]]:
            continue
        t = Transcript(context_from_json(b["context"]))
        got = prove_logup(dev, t, to_gpu_instances(insts), sched, to_gpu_forests(forests))
        assert got.alpha == fp2(b["alpha"]) and got.beta == fp2(b["beta"])
        assert as_tuples(got.roots) == [fp2(x) for x in b["roots"]]
        assert len(got.levels) == len(b["levels"])
        for g, w in zip(got.levels, b["levels"]):
            assert g["root_challenges"] == [fp2(c) for c in w.get("root_challenges", [])]
            assert g["gamma"] == fp2(w["gamma"])
            assert g["lambda"] == (fp2(w["lambda"]) if w["lambda"] is not None else None)
            assert g["rounds"] == [[fp2(v) for v in m] for m in w["rounds"]], (b["name"], w["prefix"])
            assert g["challenges"] == [fp2(c) for c in w["challenges"]]
            assert g["final_evals"] == [fp2(f) for f in w["final_evals"]]
            assert g["rho"] == fp2(w["rho"]) and g["state"] == w["state"]
        assert as_tuples(got.values) == [fp2(v) for v in b["values"]]
        assert t.state.hex() == b["state"] and t.counter == b["counter"]
        assert got.forest_claims == want["forest_claims"]
        checked += 1
    assert checked >= 13


@pytest.mark.parametrize("batch", TREE_BATCHES, ids=[b["name"][:40] for b in TREE_BATCHES])
def test_tree_batches(dev, batch):
    """prove_trees over single product trees (forests of one tree), against the reference's
    recorded levels and leaf claims."""
    from tessera_amd.logup import prove_trees
    from tessera_amd.transcript import Transcript
    t = Transcript(context_from_json(batch["context"]))
    claims, levels = prove_trees(dev, t, to_gpu_forests(
        [([h(x) for x in tr["leaves"]], tr["height"]) for tr in batch["trees"]]))
    check_tree_levels(levels, batch["levels"])
    for n in batch["levels"][-1]["next"]:
        assert claims[n["tree"]] == fp2(n["value"])


@pytest.mark.parametrize("batch", FOREST_BATCHES, ids=[b["name"][:40] for b in FOREST_BATCHES])
def test_forest_batches(dev, batch):
    """prove_trees over product forests of 2^s trees (points from the root challenges)."""
    from tessera_amd.logup import prove_trees
    from tessera_amd.transcript import Transcript
    t = Transcript(context_from_json(batch["context"]))
    claims, levels = prove_trees(dev, t, to_gpu_forests(
        [([h(x) for x in f["leaves"]], f["height"]) for f in batch["forests"]]))
    check_tree_levels(levels, batch["levels"])
    for n in batch["levels"][-1]["next"]:
        assert claims[n["forest"]] == fp2(n["value"])


@pytest.mark.parametrize("shapes", [[(12, 3)], [(10, 10), (3, 1)], [(16, 2), (16, 16), (5, 4)],
                                    [(20, 5)]])
def test_random_forests(dev, shapes):
    """prove_trees over random forests of 2^(s+h) leaves, against the oracle; the larger ones
    run the level kernels over many blocks."""
    from tessera_amd.logup import prove_trees
    from tessera_amd.transcript import Transcript
    rng = np.random.default_rng(sum(v for v, _ in shapes))
    forests = [(rng.integers(0, P, 1 << v, dtype=np.uint64), ht) for v, ht in shapes]
    t = Transcript(b"forests")
    claims, levels = prove_trees(dev, t, to_gpu_forests(forests))
    ot = O.OracleTranscript(b"forests")
    want_claims, want_levels = O.prove_trees(ot, forests)
    assert claims == want_claims and levels == want_levels
    assert t.state == ot.state and t.counter == ot.counter


def test_forest_shape_errors(dev):
    import tessera_amd as T
    from tessera_amd.logup import prove_trees
    from tessera_amd.transcript import Transcript
    t = Transcript(b"forests")
    with pytest.raises(T.BadShape, match="LeafCount: forest 1"):
        prove_trees(dev, t, to_gpu_forests([(np.ones(4, np.uint64), 1), (np.ones(6, np.uint64), 1)]))
    with pytest.raises(T.BadShape, match="Height: forest 0"):
        prove_trees(dev, t, to_gpu_forests([(np.ones(4, np.uint64), 3)]))
    with pytest.raises(T.BadShape, match="Height: forest 0"):
        prove_trees(dev, t, to_gpu_forests([(np.ones(4, np.uint64), 0)]))
    assert t.state == Transcript(b"forests").state  # before the transcript moves


def compare_with_oracle(dev, insts, sched, ctx=b"logup", forests=()):
    from tessera_amd.logup import prove_logup
    from tessera_amd.transcript import Transcript
    t = Transcript(ctx)
    got = prove_logup(dev, t, to_gpu_instances(insts), sched, to_gpu_forests(forests))
    ot = O.OracleTranscript(ctx)
    want = O.prove_logup(ot, insts, sched, forests)
    assert got.alpha == want["alpha"] and got.beta == want["beta"]
    assert as_tuples(got.roots) == as_tuples(want["roots"])
    assert got.levels == want["levels"]
    assert as_tuples(got.values) == as_tuples(want["values"])
    assert as_tuples(got.claims) == as_tuples(want["claims"])
    assert got.forest_claims == want["forest_claims"]
    assert t.state == ot.state and t.counter == ot.counter
    return got


def random_instance(rng, T, N, values=True):
    keys = rng.integers(0, P, T, dtype=np.uint64)
    vals = rng.integers(0, P, T, dtype=np.uint64) if values else None
    rows = rng.integers(0, T, N)
    return {"keys": keys, "values": vals, "lookup_keys": keys[rows],
            "lookup_values": vals[rows] if values else None,
            "multiplicities": O.multiplicities(T, rows)}


@pytest.mark.parametrize("shape", [[(7, 6)], [(1, 1)], [(300, 1000), (17, 5), (2, 2)],
                                   [(1 << 10, 1 << 13)]])
def test_random_batches(dev, shape):
    rng = np.random.default_rng(len(shape) * 100 + shape[0][1])
    insts = [random_instance(rng, T, N, values=(i % 2 == 0)) for i, (T, N) in enumerate(shape)]
    sched = list(range(2 * len(insts)))
    sched = sched[::-1] if len(insts) > 1 else sched
    compare_with_oracle(dev, insts, sched)


@pytest.mark.parametrize("case", ["taller_forest", "shorter_forest", "two_forests_interleaved"])
def test_random_batches_with_forests(dev, case):
    """Fraction trees beside product forests in one level loop: a forest taller than every
    tree (levels without lambda first), one joining late, and two interleaved in the schedule."""
    rng = np.random.default_rng({"taller_forest": 1, "shorter_forest": 2, "two_forests_interleaved": 3}[case])
    if case == "taller_forest":
        insts = [random_instance(rng, 9, 20)]
        forests = [(rng.integers(0, P, 1 << 9, dtype=np.uint64), 7)]
        sched = [2, 0, 1]
    elif case == "shorter_forest":
        insts = [random_instance(rng, 300, 700, values=False)]
        forests = [(rng.integers(0, P, 1 << 6, dtype=np.uint64), 2)]
        sched = [0, 2, 1]
    else:
        insts = [random_instance(rng, 40, 100), random_instance(rng, 5, 3, values=False)]
        forests = [(rng.integers(0, P, 1 << 8, dtype=np.uint64), 8),
                   (rng.integers(0, P, 1 << 12, dtype=np.uint64), 4)]
        sched = [4, 0, 3, 5, 2, 1]
    compare_with_oracle(dev, insts, sched, forests=forests)


def test_multiplicities(dev):
    from tessera_amd.logup import multiplicities
    rng = np.random.default_rng(2)
    rows = rng.integers(0, 1000, 5000).astype(np.uint64)
    assert np.array_equal(multiplicities(dev, 1000, rows), O.multiplicities(1000, rows))
    import tessera_amd as T
    with pytest.raises(T.BadShape, match="LookupRow"):
        multiplicities(dev, 10, np.array([3, 10], dtype=np.uint64))
    # into a device buffer: the same counts; a bad row leaves the buffer as it was
    out = dev.buffer(1024)
    d_rows = dev.upload(rows)
    assert multiplicities(dev, 1000, d_rows, out=out) is out
    assert np.array_equal(out.download(), O.multiplicities(1000, rows))
    before = out.download()
    with pytest.raises(T.BadShape, match="LookupRow"):
        multiplicities(dev, 1000, dev.upload(np.array([3, 1000], dtype=np.uint64)), out=out)
    assert np.array_equal(out.download(), before)
    with pytest.raises(ValueError):
        multiplicities(dev, 1000, d_rows, out=dev.buffer(1000))
    # a table of no row: no leaf, so no lookup is an empty result and any lookup names no row
    assert multiplicities(dev, 0, np.zeros(0, dtype=np.uint64)).size == 0
    with pytest.raises(T.BadShape, match="LookupRow"):
        multiplicities(dev, 0, np.array([0], dtype=np.uint64))


@pytest.mark.parametrize("rows,lookups", [(1, 70000), (1 << 14, 1), ((1 << 14) + 1, 300000),
                                          (1 << 17, 200000), ((1 << 17) + 1, 200000)])
def test_multiplicities_histogram_and_global_paths(dev, rows, lookups):
    """Tables up to 2^17 rows count in LDS slices of 2^14 (chunks of 2^16 lookups), larger ones
    with global atomics; both against exact counting, with Zipf-hot rows."""
    from tessera_amd.logup import multiplicities
    from tests.helpers import zipf_rows
    idx = zipf_rows(lookups, rows, 1.1, rows + lookups)
    assert np.array_equal(multiplicities(dev, rows, idx), O.multiplicities(rows, idx))


def test_config3_logup_against_oracle(dev):
    """2^22 lookups Zipf(1.1) into 2^16 rows with values: the whole proof against the oracle."""
    from tessera_amd.logup import multiplicities
    keys, vals, idx = logup_config3()
    m = multiplicities(dev, keys.size, idx)
    assert np.array_equal(m, O.multiplicities(keys.size, idx))
    inst = {"keys": keys, "values": vals, "lookup_keys": keys[idx], "lookup_values": vals[idx],
            "multiplicities": m}
    got = compare_with_oracle(dev, [inst], [0, 1], ctx=b"config-3-logup")
    # the verifier's root check: n_a d_t + n_t d_a = 0
    na, da, nt, dt = as_tuples(got.roots)
    s1, s2 = O.fp2_mul(na, dt), O.fp2_mul(nt, da)
    assert ((s1[0] + s2[0]) % P, (s1[1] + s2[1]) % P) == (0, 0)


@pytest.mark.parametrize("blocks", [16, 7, 1])
def test_narrow_round_launches(dev, blocks):
    """With at most `blocks` workgroups per round launch, a thread takes 8 or more pairs of a
    round from 2^15 pairs on (2^14 with 7, every round with 1), so the level rounds weight each
    pair by eq_lo alone and scale each run of pairs by its eq_hi entry, with runs that cross
    block chunks (7) and several runs per block (1): LogUp of 2^19 lookups with values and a
    2^18-leaf forest, against the oracle. The launch width must not change a proof."""
    from tessera_amd.lib import TsrError
    from tessera_amd.logup import prove_trees
    from tessera_amd.transcript import Transcript
    with pytest.raises(TsrError):
        dev.set_round_blocks(1025)
    dev.set_round_blocks(blocks)
    try:
        rng = np.random.default_rng(900 + blocks)
        compare_with_oracle(dev, [random_instance(rng, 1 << 10, 1 << 19)], [0, 1],
                            ctx=b"narrow-rounds")
        forests = [(rng.integers(0, P, 1 << 18, dtype=np.uint64), 4)]
        t = Transcript(b"narrow-forest")
        claims, levels = prove_trees(dev, t, to_gpu_forests(forests))
        ot = O.OracleTranscript(b"narrow-forest")
        want_claims, want_levels = O.prove_trees(ot, forests)
        assert claims == want_claims and levels == want_levels
        assert t.state == ot.state and t.counter == ot.counter
    finally:
        dev.set_round_blocks(0)
