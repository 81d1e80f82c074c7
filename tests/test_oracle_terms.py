"""Pins the oracle's sum-of-products instances (Instance::sum_of_products, prover.rs:243-312; the
terms' kernel, kernel.rs:502-633). The reference's vectors hold product instances only, so the
pins are: a one-term sum over every table in order is the product instance the session vectors
pin (message by message, challenge, final evaluation and state); the weights act linearly; a
repeated factor is a squared table; and the sumcheck verifier's checks (tv-core sumcheck.rs:
306-398) on brute-force sums and on each table's MLE at the instance's point, for batches of
padded instances with public tables, repeated factors and more tables than the degree."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, load


def f2s(xs):
    return np.array([[h(a), h(b)] for a, b in xs], dtype=np.uint64)


def tuples(a):
    return [tuple(int(x) for x in v) for v in np.asarray(a).reshape(-1, 2)]


def add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def test_one_term_is_the_product_on_the_session_vectors():
    proven = 0
    for sess in load("sumcheck.json")["sessions"]:
        t = O.OracleTranscript(context_from_json(sess["context"]))
        for st in sess["stages"]:
            if not all("tables" in inst for inst in st["instances"]):
                break
            insts = [O.SumOfProducts([(f2s(tab), True, False) for tab in inst["tables"]],
                                     [((1, 0), list(range(len(inst["tables"]))))])
                     for inst in st["instances"]]
            out = O.prove_batch(t, st["prefix"], st["degree"], insts)
            assert out["claim"] == fp2(st["claims"][0])
            assert [tuples(m) for m in out["rounds"]] == [[fp2(v) for v in m] for m in st["rounds"]]
            assert tuples(out["challenges"]) == [fp2(c) for c in st["challenges"]]
            assert tuples(out["finals"]) == [fp2(f) for f in st["final_evals"]]
            assert t.state.hex() == st["state"]
            proven += 1
    assert proven >= 10


def run(instances, degree, ctx=b"terms"):
    t = O.OracleTranscript(ctx)
    return O.prove_batch(t, "stage/2", degree, instances), t


def test_weights_act_linearly_and_repeated_factors_square():
    rng = np.random.default_rng(5)
    n = 6
    a, b = (rng.integers(0, P, (1 << n, 2), dtype=np.uint64) for _ in range(2))
    tabs = [(a, True, False), (b, True, False)]
    w1, w2 = (3, 5), (P - 1, 11)
    split, _ = run([O.SumOfProducts(tabs, [(w1, [0, 1]), (w2, [0, 1])])], 2)
    joined, _ = run([O.SumOfProducts(tabs, [(add(w1, w2), [0, 1])])], 2)
    assert np.array_equal(split["rounds"], joined["rounds"]) and split["claim"] == joined["claim"]
    # a * a * b as a term with a repeated factor, and as the product of [a, a, b]
    sq, _ = run([O.SumOfProducts(tabs, [((1, 0), [0, 0, 1])])], 3)
    prod, _ = run([[tabs[0], tabs[0], tabs[1]]], 3)
    assert np.array_equal(sq["rounds"], prod["rounds"]) and sq["claim"] == prod["claim"]
    assert np.array_equal(sq["challenges"], prod["challenges"])
    # the finals are the table list's (a, b), the product's (a, a, b)
    assert tuples(sq["finals"]) == [tuples(prod["finals"])[0], tuples(prod["finals"])[2]]


def brute_sum(tables, terms):
    """Sum over the hypercube of sum_t w_t prod_f T_f(x), in Python integers."""
    n = len(tables[0])
    total = (0, 0)
    for x in range(n):
        for w, factors in terms:
            p = (w, 0) if isinstance(w, int) else w
            for f in factors:
                p = O.fp2_mul(p, tables[f][x])
            total = add(total, p)
    return total


def as_pairs(vals, ext):
    return [tuple(int(y) for y in v) for v in vals] if ext else [(int(v), 0) for v in vals]


CASES = [
    # (degree, [(vars, [(ext, public)...], terms)])
    (3, [(5, [(True, True), (False, False), (True, False), (False, False), (True, False)],
          [((1, 0), [0, 1, 4]), ((1, 0), [0, 2, 3]), ((12345, 678), [0, 3, 4])])]),
    (4, [(4, [(False, False), (True, False)], [((2, 0), [0, 0, 1, 1]), ((7, 9), [1])]),
         (7, [(True, False), (True, True), (False, False)], [((1, 0), [2]), ((P - 3, 1), [0, 1, 2, 2])]),
         (0, [(True, False), (False, False)], [((5, 5), [0, 1]), ((1, 0), [1, 1, 1])])]),
    (2, [(3, [(False, False)] * 6, [((i + 1, i), [i, (i + 1) % 6]) for i in range(6)])]),
]


@pytest.mark.parametrize("degree,spec", CASES)
def test_verifier_checks_on_sums_of_products(degree, spec):
    rng = np.random.default_rng(degree * 31 + len(spec))
    insts, raw = [], []
    for vars_, kinds, terms in spec:
        tabs, vals = [], []
        for ext, public in kinds:
            v = rng.integers(0, P, ((1 << vars_), 2) if ext else (1 << vars_,), dtype=np.uint64)
            tabs.append((v, ext, public))
            vals.append(as_pairs(v, ext))
        insts.append(O.SumOfProducts(tabs, terms))
        raw.append((vars_, tabs, vals, terms))
    out, _ = run(insts, degree)
    gamma = out["gamma"]
    n_max = max(v for v, *_ in raw)
    claim, w = (0, 0), (1, 0)
    for vars_, _, vals, terms in raw:
        s = brute_sum(vals, terms)
        claim = add(claim, O.fp2_mul(O.fp2_mul(w, s), (pow(2, n_max - vars_, P), 0)))
        w = O.fp2_mul(w, gamma)
    assert out["claim"] == claim
    chal = tuples(out["challenges"])
    for j in range(n_max):
        claim = O.round_eval(out["rounds"][j], degree, claim, chal[j])
    finals = tuples(out["finals"])
    at, final, w = 0, (0, 0), (1, 0)
    for vars_, tabs, _, terms in raw:
        point = np.array(chal[n_max - vars_:], dtype=np.uint64).reshape(-1, 2)
        values = [O.mle_eval(v, ext, point) for v, ext, _ in tabs]
        for (v, ext, public), val in zip(tabs, values):
            if not public:
                assert finals[at] == val
                at += 1
        rel = (0, 0)
        for tw, factors in terms:
            p = tw
            for f in factors:
                p = O.fp2_mul(p, values[f])
            rel = add(rel, p)
        final = add(final, O.fp2_mul(w, rel))
        w = O.fp2_mul(w, gamma)
    assert at == len(finals)
    assert claim == final


def test_term_shape_errors():
    a = np.ones(4, np.uint64)
    tabs = [(a, False, False), (a, False, False)]
    for terms in ([], [((1, 0), [])], [((1, 0), [0, 0, 0, 1])], [((1, 0), [2])]):
        t = O.OracleTranscript(b"x")
        with pytest.raises(ValueError):
            O.prove_batch(t, "stage/0", 3, [O.SumOfProducts(tabs, terms)])
        assert t.state == O.OracleTranscript(b"x").state
