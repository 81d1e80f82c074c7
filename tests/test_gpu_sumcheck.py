"""Parity of the gfx950 batched sumcheck (device-side round loop and transcript) with the oracle
and the reference's sumcheck vectors: every message, challenge, final evaluation and transcript
state byte for byte; front padding, public factors, base-field tables, degrees 1..10; config 1's
Sum a*b (n = 14) and config 3's Sum eq(r,x) a b (n = 24, checked by the verifier's invariants)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, load, splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def f2s(xs):
    return np.array([[h(a), h(b)] for a, b in xs], dtype=np.uint64)


def as_tuples(a):
    return [tuple(int(x) for x in v) for v in np.asarray(a).reshape(-1, 2)]


def test_sessions_match_reference_vectors(dev):
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript, tce_fp2s
    proven = 0
    for sess in load("sumcheck.json")["sessions"]:
        t = Transcript(context_from_json(sess["context"]))
        for st in sess["stages"]:
            prefix, d = st["prefix"], st["degree"]
            if all("tables" in inst for inst in st["instances"]):
                insts = [[Table(f2s(tab), ext=True, vars=inst["vars"]) for tab in inst["tables"]]
                         for inst in st["instances"]]
                out = prove_batch(dev, t, prefix, d, insts)
                assert out.gamma == fp2(st["gamma"])
                assert as_tuples(out.rounds) == [fp2(v) for m in st["rounds"] for v in m]
                assert as_tuples(out.challenges) == [fp2(c) for c in st["challenges"]]
                assert as_tuples(out.final_evals) == [fp2(f) for f in st["final_evals"]]
                proven += 1
            else:  # replay the recorded messages to keep the transcript in step
                t.challenge_fp2(prefix + "/gamma")
                for j, msg in enumerate(st["rounds"]):
                    t.absorb(f"{prefix}/round/{j}", tce_fp2s(f2s(msg)))
                    t.challenge_fp2(f"{prefix}/r/{j}")
                t.absorb(prefix + "/evals", tce_fp2s(f2s(st["final_evals"])))
            assert t.state.hex() == st["state"], (sess["name"], prefix)
            assert t.counter == st["counter"]
    assert proven >= 31


def run_both(dev, prefix, degree, spec, ctx=b"ctx"):
    """spec: list of instances, each a list of (values, ext, public)."""
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    t = Transcript(ctx)
    g = prove_batch(dev, t, prefix, degree, [[Table(v, e, p) for v, e, p in inst] for inst in spec])
    ot = O.OracleTranscript(ctx)
    want = O.prove_batch(ot, prefix, degree, spec, threads=8)
    assert g.gamma == want["gamma"]
    assert np.array_equal(g.rounds, want["rounds"])
    assert np.array_equal(g.challenges, want["challenges"])
    assert np.array_equal(g.final_evals, want["finals"])
    assert t.state == ot.state and t.counter == ot.counter
    return g


@pytest.mark.parametrize("degree", list(range(1, 11)))
def test_degrees_and_padding(dev, degree):
    rng = np.random.default_rng(degree)
    spec = []
    for n, k in [(9, degree), (3, max(1, degree - 1)), (0, 1), (12, min(degree, 2))]:
        inst = []
        for f in range(k):
            ext = (f + n) % 2 == 1
            v = rng.integers(0, P, (1 << n) * (2 if ext else 1), dtype=np.uint64)
            inst.append((v.reshape(-1, 2) if ext else v, ext, f == 1))
        spec.append(inst)
    run_both(dev, f"stage/{degree}", degree, spec)


@pytest.mark.parametrize("plen", [1, 2, 3, 4, 22, 37, 48, 55])
def test_label_lengths(dev, plen):
    """Stage prefixes whose round labels ("<prefix>/round/<j>", "<prefix>/r/<j>") take every
    length mod 4 up to the 64-byte label slot, so the device transcript's hash inputs put the
    message at every byte alignment and span one to three BLAKE3 blocks; 12 variables, so both
    the transcript launches and the tail's in-block rounds run."""
    rng = np.random.default_rng(plen)
    prefix = "".join(chr(ord("a") + (i % 26)) for i in range(plen))
    spec = [[(rng.integers(0, P, 1 << 12, dtype=np.uint64), False, False) for _ in range(3)]]
    run_both(dev, prefix, 3, spec)


def test_all_instances_without_variables(dev):
    spec = [[(np.array([5], dtype=np.uint64), False, False)], [(np.array([[7, 9]], dtype=np.uint64), True, False)]]
    run_both(dev, "stage/0", 2, spec)


def test_config1_sum_a_b(dev):
    """14-variable Sum a(x) b(x) over two uniform base tables (config 1's sumcheck)."""
    a = splitmix_uniform(1 << 14, 11)
    b = splitmix_uniform(1 << 14, 12)
    run_both(dev, "stage/0", 2, [[(a, False, False), (b, False, False)]])


def test_shape_errors_leave_transcript(dev):
    import tessera_amd as T
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    t = Transcript(b"x")
    s0 = t.state
    with pytest.raises(T.BadShape, match="Degree"):
        prove_batch(dev, t, "stage/0", 11, [[Table(np.ones(4, np.uint64))]])
    with pytest.raises(T.BadShape, match="TooManyTables"):
        prove_batch(dev, t, "stage/0", 1, [[Table(np.ones(4, np.uint64)), Table(np.ones(4, np.uint64))]])
    with pytest.raises(T.BadShape, match="NoInstances"):
        prove_batch(dev, t, "stage/0", 1, [])
    assert t.state == s0 and t.counter == 0


def test_config3_eq_a_b_invariants(dev):
    """24 variables, degree 3, Sum eq(r,x) a(x) b(x), eq public, a and b base and resident on the
    device: every round satisfies p_j(0) + p_j(1) = claim_j from the claim the oracle's brute-force
    sum gives, and the last claim equals eq(r, point) a~(point) b~(point) with a~, b~ the final
    evaluations checked against the oracle's MLE evaluation."""
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    n = 24
    t = Transcript(b"config-3")
    r = [t.challenge_fp2(f"bench/r/{i}") for i in range(n)]
    eq = O.eq_table(np.array(r, dtype=np.uint64))
    a = splitmix_uniform(1 << n, 31)
    b = splitmix_uniform(1 << n, 32)
    da, db, de = dev.upload(a), dev.upload(b), dev.upload(eq.reshape(-1))
    out = prove_batch(dev, t, "stage/0", 3, [[Table(de, ext=True).public(), Table(da), Table(db)]])
    # the claim: Sum_x eq(r,x) a(x) b(x) = (a*b)~(r), the MLE of the entrywise product at r
    claim = O.mle_eval(O.fp_mul_vec(a, b), False, np.array(r, dtype=np.uint64))
    for j in range(n):
        claim = O.round_eval(out.rounds[j], 3, claim, out.challenges[j])
    pt = out.challenges
    ea = O.mle_eval(a, False, pt)
    eb = O.mle_eval(b, False, pt)
    assert as_tuples(out.final_evals) == [ea, eb]
    e = (1, 0)
    for ri, xi in zip(r, as_tuples(pt)):
        # eq(r_i, x_i) = r x + (1 - r)(1 - x)
        rx = O.fp2_mul(ri, xi)
        one_r = ((1 - ri[0]) % P, (-ri[1]) % P)
        one_x = ((1 - xi[0]) % P, (-xi[1]) % P)
        t2 = O.fp2_mul(one_r, one_x)
        e = O.fp2_mul(e, ((rx[0] + t2[0]) % P, (rx[1] + t2[1]) % P))
    assert claim == O.fp2_mul(O.fp2_mul(e, ea), eb)


def oracle_eq_proof(ctx, prefix, degree, specs, threads=16):
    """The oracle's proof over materialized eq tables: specs are lists of ("eq", point, public) or
    (values, ext, public)."""
    t = O.OracleTranscript(ctx)
    insts = []
    for spec in specs:
        tabs = []
        for x in spec:
            if isinstance(x[0], str):
                tabs.append((O.eq_table(np.asarray(x[1], dtype=np.uint64).reshape(-1)), True, x[2]))
            else:
                tabs.append(x)
        insts.append(tabs)
    return O.prove_batch(t, prefix, degree, insts, threads=threads), t


def device_eq_proof(dev, ctx, prefix, degree, specs):
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    t = Transcript(ctx)
    insts = []
    for spec in specs:
        tabs = []
        for x in spec:
            if isinstance(x[0], str):
                tab = Table.eq(x[1])
            else:
                tab = Table(x[0] if not hasattr(x[0], "ptr") else x[0], ext=bool(x[1]))
            tabs.append(tab.public() if x[2] else tab)
        insts.append(tabs)
    return prove_batch(dev, t, prefix, degree, insts), t


def assert_same_proof(want, wt, got, gt):
    assert got.gamma == want["gamma"]
    assert np.array_equal(got.rounds.reshape(-1), want["rounds"].reshape(-1))
    assert np.array_equal(got.challenges.reshape(-1), want["challenges"].reshape(-1))
    assert np.array_equal(got.final_evals.reshape(-1), want["finals"].reshape(-1))
    assert gt.state == wt.state and gt.counter == wt.counter


@pytest.mark.parametrize("n", [1, 2, 9, 12, 16, 20, 24])
def test_eq_point_sum_eq_a_b_matches_oracle(dev, n):
    """Sum eq(r, x) a(x) b(x), degree 3, eq public and given by its point: no eq table on the
    device, the same proof bytes as the oracle's prove_batch over the materialized table. n = 24
    is the sumcheck/lookup configuration's sumcheck at its full size, with r squeezed from the
    transcript as the configuration draws it."""
    from tessera_amd.transcript import Transcript
    ctx = b"config-3"
    t0 = Transcript(ctx)
    r = np.array([t0.challenge_fp2(f"bench/r/{i}") for i in range(n)], dtype=np.uint64)
    a, b = splitmix_uniform(1 << n, 40 + n), splitmix_uniform(1 << n, 80 + n)
    print(f"eq point sumcheck n = {n}", flush=True)
    want, wt = oracle_eq_proof(ctx, "stage/0", 3, [[("eq", r, True), (a, False, False), (b, False, False)]])
    src = [dev.upload(a), dev.upload(b)] if n >= 16 else [a, b]
    got, gt = device_eq_proof(dev, ctx, "stage/0", 3, [[("eq", r, True), (src[0], False, False),
                                                        (src[1], False, False)]])
    assert_same_proof(want, wt, got, gt)


def test_eq_point_positions_padding_and_fallbacks(dev):
    """The eq point anywhere in the table list, beside Fp2 tables, in a batch where it is padded
    (its S from its first sums with the eq line), and the uses the round kernels do not take as
    a point (a non-public eq, a product of two, degree 4: the table is built on the device) -- all equal to the oracle over materialized tables."""
    rng = np.random.default_rng(5)

    def pt(n):
        return rng.integers(0, P, (n, 2), dtype=np.uint64)

    def fp(n):
        return rng.integers(0, P, 1 << n, dtype=np.uint64)

    def f2(n):
        return rng.integers(0, P, (1 << n, 2), dtype=np.uint64)

    cases = [
        (3, [[(fp(10), False, False), ("eq", pt(10), True), (f2(10), True, False)]]),
        (3, [[(f2(9), True, False), (fp(9), False, False), ("eq", pt(9), True)],
             [("eq", pt(13), True), (fp(13), False, False), (fp(13), False, False)],
             [("eq", pt(6), True), (fp(6), False, False), (fp(6), False, False)]]),
        (3, [[("eq", pt(8), False), (fp(8), False, False), (fp(8), False, False)]]),
        (3, [[("eq", pt(8), True), (fp(8), False, False)]]),
        (4, [[("eq", pt(7), True), (fp(7), False, False), (fp(7), False, False)]]),
    ]
    for degree, specs in cases:
        want, wt = oracle_eq_proof(b"eq-cases", "stage/3", degree, specs)
        got, gt = device_eq_proof(dev, b"eq-cases", "stage/3", degree, specs)
        assert_same_proof(want, wt, got, gt)


@pytest.mark.parametrize("n", [0, 1, 5, 12, 13, 20])
def test_eq_table_matches_oracle(dev, n):
    """Direct products up to 12 variables, the outer product of two half tables above."""
    from tessera_amd.sumcheck import eq_table
    rng = np.random.default_rng(n)
    r = rng.integers(0, P, (n, 2), dtype=np.uint64)
    got = eq_table(dev, r)
    want = O.eq_table(r.reshape(-1)).reshape(-1, 2)
    assert np.array_equal(got, want)
    d = dev.buffer(2 << n)
    eq_table(dev, r, out=d)
    assert np.array_equal(d.download().reshape(-1, 2), want)


def run_terms(dev, prefix, degree, spec, ctx=b"terms"):
    """spec: list of (tables [(values, ext, public)], terms [(weight, [factors])] or None)."""
    from tessera_amd.sumcheck import SumOfProducts, Table, prove_batch
    from tessera_amd.transcript import Transcript
    dev_insts, orc_insts = [], []
    for tabs, terms in spec:
        dt = [Table(v, e, p) for v, e, p in tabs]
        dev_insts.append(dt if terms is None else SumOfProducts(dt, terms))
        orc_insts.append(tabs if terms is None else O.SumOfProducts(tabs, terms))
    t = Transcript(ctx)
    g = prove_batch(dev, t, prefix, degree, dev_insts)
    ot = O.OracleTranscript(ctx)
    want = O.prove_batch(ot, prefix, degree, orc_insts, threads=8)
    assert g.gamma == want["gamma"]
    assert np.array_equal(g.rounds, want["rounds"])
    assert np.array_equal(g.challenges, want["challenges"])
    assert np.array_equal(g.final_evals, want["finals"])
    assert t.state == ot.state and t.counter == ot.counter


def rand_tables(rng, n, kinds):
    return [(rng.integers(0, P, ((1 << n), 2) if ext else (1 << n,), dtype=np.uint64), ext, public)
            for ext, public in kinds]


@pytest.mark.parametrize("degree", [1, 2, 3, 5, 7, 10])
def test_sums_of_products_against_oracle(dev, degree):
    """Sum-of-products instances beside products: a tree-level shape (eq public, four tables,
    three terms, a non-unit weight), repeated factors, more tables than the degree, padded
    instances down to no variable, base and Fp2 tables, weights with both coordinates."""
    rng = np.random.default_rng(100 + degree)
    d = degree
    spec = []
    kinds5 = [(True, True), (False, False), (True, False), (False, False), (True, False)]
    terms5 = [((1, 0), [0, 1, 4][:d]), ((1, 0), [0, 2, 3][:d]), ((12345, 678), [0, 3, 4][:d])]
    spec.append((rand_tables(rng, 11, kinds5), terms5))
    spec.append((rand_tables(rng, 6, [(False, False), (True, False)]),
                 [((2, 0), ([0, 1] * d)[:d]), ((7, 9), [1]), ((P - 1, P - 2), [0] * d)]))
    spec.append((rand_tables(rng, 0, [(True, False), (False, False)]),
                 [((5, 5), [0, 1][:d]), ((1, 0), [1] * d)]))
    spec.append((rand_tables(rng, 9, [(False, False)] * min(d, 3)), None))
    kinds10 = [(i % 2 == 0, i == 3) for i in range(10)]
    spec.append((rand_tables(rng, 13, kinds10),
                 [((i + 1, i), [(i + m) % 10 for m in range(1 + i % d)]) for i in range(16)]))
    run_terms(dev, "stage/3", degree, spec)


def test_sum_of_products_shape_errors(dev):
    import tessera_amd as T
    from tessera_amd.sumcheck import SumOfProducts, Table, prove_batch
    from tessera_amd.transcript import Transcript
    a = Table(np.ones(4, np.uint64), False, False)
    t = Transcript(b"x")
    for terms, name in [([], "NoTerms: instance 1"), ([((1, 0), [])], "TermFactors: instance 1 term 0"),
                        ([((1, 0), [0]), ((1, 0), [0, 0, 0, 1])], "TermFactors: instance 1 term 1"),
                        ([((1, 0), [2])], "TermFactor: instance 1 term 0")]:
        with pytest.raises(T.BadShape, match=name):
            prove_batch(dev, t, "stage/0", 3, [[a], SumOfProducts([a, a], terms)])
    assert t.state == Transcript(b"x").state
    # more tables than the degree is a product's refusal, not a sum's
    with pytest.raises(T.BadShape, match="TooManyTables"):
        prove_batch(dev, t, "stage/0", 1, [[a, a]])
    prove_batch(dev, t, "stage/0", 1, [SumOfProducts([a, a, a], [((1, 0), [2])])])


@pytest.mark.parametrize("count", [65, 256])
def test_many_instances(dev, count):
    """Batches past the transcript kernel's staged instance facts (64) up to the largest batch
    (256): front padding from 0 to 6 variables, base and Fp2 tables, public factors, products of
    two and three factors at degree 3."""
    rng = np.random.default_rng(count)
    spec = []
    for i in range(count):
        n = int(rng.integers(0, 7))
        k = 2 + (i % 2)
        inst = []
        for f in range(k):
            ext = (i + f) % 3 == 0
            v = rng.integers(0, P, (1 << n) * (2 if ext else 1), dtype=np.uint64)
            inst.append((v.reshape(-1, 2) if ext else v, ext, f == 2))
        spec.append(inst)
    run_both(dev, "stage/many", 3, spec)


def test_more_than_256_instances_are_declined(dev):
    import tessera_amd as T
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    t = Transcript(b"x")
    s0 = t.state
    with pytest.raises(T.Declined, match="256"):
        prove_batch(dev, t, "stage/0", 2, [[Table(np.ones(2, np.uint64))] for _ in range(257)])
    assert t.state == s0 and t.counter == 0


@pytest.mark.parametrize("blocks", [1, 7, 64])
def test_narrow_round_launches(dev, blocks):
    """At most `blocks` workgroups per round launch (each thread then takes many pairs of a round,
    and a round's partials come from a block count that is not a multiple of the wave): products
    of degree 3 and 2 over 16 and 13 variables, base and Fp2 tables, against the oracle."""
    rng = np.random.default_rng(700 + blocks)
    spec = [[(v, e, False) for v, e, _ in rand_tables(rng, 16, [(False, False), (True, False), (False, False)])],
            [(v, e, False) for v, e, _ in rand_tables(rng, 13, [(True, False), (False, False), (False, False)])]]
    dev.set_round_blocks(blocks)
    try:
        run_both(dev, "stage/narrow", 3, spec)
    finally:
        dev.set_round_blocks(0)
