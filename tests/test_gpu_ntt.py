"""The device NttPlan (crates/tessera-pcs/src/ntt.rs:397-516): ntt, intt, coset_ntt and
coset_intt, natural order on both sides, against the reference's NTT vectors (n = 2^1..2^8,
subgroup and coset 7, spec/vectors/encoding) and the oracle's transforms at 2^10..2^22, as single
rows and batches of rows, host and device memory; inverses round-trip; errors in the reference's
order before any value moves."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, hexes, load, splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def test_reference_ntt_vectors(dev):
    import tessera_amd as T
    vec = load("encoding.json")
    shift = int(vec["coset_shift"], 0) if isinstance(vec["coset_shift"], str) else vec["coset_shift"]
    for case in vec["ntt"]:
        n = case["n"]
        coeffs, values, coset = hexes(case["coefficients"]), hexes(case["values"]), hexes(case["coset_values"])
        plan = T.NttPlan(dev, n)
        assert plan.size() == n
        assert np.array_equal(plan.ntt(coeffs.copy()), values), n
        assert np.array_equal(plan.intt(values.copy()), coeffs), n
        assert np.array_equal(plan.coset_ntt(coeffs.copy(), shift), coset), n
        assert np.array_equal(plan.coset_intt(coset.copy(), shift), coeffs), n
        plan.free()


@pytest.mark.parametrize("log_n", [0, 1, 3, 9, 10, 11, 12, 13, 14, 17, 20, 22])
def test_against_oracle(dev, log_n):
    import tessera_amd as T
    n = 1 << log_n
    rows = 1 if log_n >= 20 else 3
    x = splitmix_uniform(rows * n, 100 + log_n).reshape(rows, n)
    shift = 7 if log_n % 2 else 0x1234567890ABCDEF % P
    plan = T.NttPlan(dev, n)
    print(f"ntt plan 2^{log_n} x {rows} rows", flush=True)
    d = dev.upload(x)
    for kind in ("ntt", "intt", "coset_ntt", "coset_intt"):
        want = np.stack([O.ntt(r.copy(), log_n, inverse=kind.endswith("intt"),
                               shift=shift if kind.startswith("coset") else None) for r in x])
        args = (shift,) if kind.startswith("coset") else ()
        got = getattr(plan, kind)(x.copy(), *args)
        assert np.array_equal(got, want), kind
        d.upload(x)
        getattr(plan, kind)(d, *args)
        assert np.array_equal(d.download().reshape(rows, n), want), kind
    # round trips
    y = x.copy()
    plan.intt(plan.ntt(y))
    assert np.array_equal(y, x)
    plan.coset_intt(plan.coset_ntt(y, shift), shift)
    assert np.array_equal(y, x)
    d.free()
    plan.free()


def test_direct_evaluation(dev):
    """At 2^6: values[j] = sum_i c_i (s w^j)^i by Python integers, against every transform."""
    import tessera_amd as T
    n, s = 64, 0xDEADBEEF
    w = O.fp_root(6)
    c = splitmix_uniform(n, 9)
    plan = T.NttPlan(dev, n)
    for shift in (1, s):
        want = np.array([sum(int(ci) * pow(shift * pow(w, j, P) % P, i, P) for i, ci in enumerate(c)) % P
                         for j in range(n)], dtype=np.uint64)
        got = plan.coset_ntt(c.copy(), shift) if shift != 1 else plan.ntt(c.copy())
        assert np.array_equal(got, want)


def test_plan_errors(dev):
    import tessera_amd as T
    from tessera_amd.lib import TsrError
    for bad in (0, 3, 12, (1 << 32) + 1):
        with pytest.raises(T.BadShape, match="NotPowerOfTwo|TooLong"):
            T.NttPlan(dev, bad)
    with pytest.raises(T.BadShape, match="TooLong"):
        T.NttPlan(dev, 1 << 33)
    plan = T.NttPlan(dev, 16)
    x = np.arange(24, dtype=np.uint64)
    with pytest.raises(T.BadShape, match="TransformLength"):
        plan.ntt(x)
    assert np.array_equal(x, np.arange(24, dtype=np.uint64))
    y = np.arange(16, dtype=np.uint64)
    with pytest.raises(T.BadShape, match="ZeroShift"):
        plan.coset_ntt(y, 0)
    with pytest.raises(T.BadShape, match="TransformLength"):
        plan.coset_intt(np.zeros(0, dtype=np.uint64), 0)  # length is checked before the shift
    with pytest.raises(TsrError):
        plan.coset_intt(y, P)
    assert np.array_equal(y, np.arange(16, dtype=np.uint64))
