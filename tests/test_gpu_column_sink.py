"""ColumnSink on the device (crates/tessera-pcs/src/streaming/opening.rs:845-890): P5's answers
from re-encoded blocks of rows and P4's combined rows summed as the blocks pass. Pinned by the
reference's opening vectors (the answered columns and the combined rows u of every group and of
the proximity point, from the witness streamed block by block with the vectors' row
coefficients) and by the oracle's encoder and combination at larger sizes with a short last
block; its shape errors come before any work."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, h, load, splitmix_uniform

pytestmark = pytest.mark.gpu

CASES = load("pcs_opening.json")["openings"]


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_opening_vectors_through_the_sink(dev, case):
    from tessera_amd.sinks import ColumnSink
    log_k, log_r, Pm, B = (int(case[x]) for x in ("log2_k", "log2_inv_rate", "mask_coeffs", "row_block"))
    k = 1 << log_k
    rows = np.array([[h(x) for x in r] for r in case["witness"]["rows"]], dtype=np.uint64).reshape(-1, k)
    masks = np.array([[h(x) for x in m] for m in case["witness"]["masks"]], dtype=np.uint64).reshape(len(rows), Pm)
    coef = np.array([[[h(a), h(b)] for a, b in g] for g in case["intermediate"]["row_coefficients"]],
                    dtype=np.uint64)  # (groups + 1) x N x 2
    distinct = sorted(set(int(q) for q in case["queries"]))
    sink = ColumnSink(dev, log_k, log_r, Pm, B, distinct)
    assert sink.combines()
    got = []
    for r0 in range(0, len(rows), B):
        r1 = min(r0 + B, len(rows))
        got.append(sink.select_combining(rows[r0:r1], masks[r0:r1], coef[:, r0:r1]))
    symbols = np.concatenate(got) if got else np.zeros((0, len(distinct)), np.uint64)
    # the answer's columns: every row's symbol in each distinct column, in row order
    assert [[int(x) for x in symbols[:, c]] for c in range(len(distinct))] == \
        [[h(x) for x in col] for col in case["columns"]]
    out = np.zeros((coef.shape[0], k, 2), dtype=np.uint64)
    sink.combined(out)
    want = np.array([[[h(a), h(b)] for a, b in g["u"]] for g in case["combined"]], dtype=np.uint64)
    assert np.array_equal(out, want)
    sink.free()


@pytest.mark.parametrize("log_k,log_r,Pm,B,N,groups", [(10, 1, 37, 32, 70, 3), (8, 2, 0, 5, 23, 1),
                                                     (12, 1, 200, 124, 130, 17), (3, 1, 8, 1, 4, 2)])
def test_against_oracle(dev, log_k, log_r, Pm, B, N, groups):
    from tessera_amd.sinks import ColumnSink
    k, n = 1 << log_k, 1 << (log_k + log_r)
    rows = splitmix_uniform(N * k, log_k).reshape(N, k)
    masks = splitmix_uniform(N * Pm, 5 + log_k).reshape(N, Pm)
    rng = np.random.default_rng(log_k)
    cols = np.unique(rng.integers(0, n, 40)).astype(np.uint64)
    coef = rng.integers(0, P, (groups, N, 2), dtype=np.uint64)
    cw = O.encode_rows(rows, log_k, log_r, masks=masks, mask_coeffs=Pm, threads=8)
    sink = ColumnSink(dev, log_k, log_r, Pm, B, cols)
    out0 = rng.integers(0, P, (groups, k, 2), dtype=np.uint64)
    out = out0.copy()
    parts = []
    for r0 in range(0, N, B):
        r1 = min(r0 + B, N)
        if r0 // B % 2:  # select alone gives the same symbols and adds nothing
            parts.append(sink.select(rows[r0:r1], masks[r0:r1]))
            co = coef.copy()
            co[:, r0:r1] = 0
            coef = co
        else:
            parts.append(sink.select_combining(rows[r0:r1], masks[r0:r1], coef[:, r0:r1]))
    assert np.array_equal(np.concatenate(parts), cw[:, cols.astype(np.int64)])
    sink.combined(out)
    want = O.combine(coef, rows, k)
    for g in range(groups):
        for i in range(k):
            assert tuple(out[g, i]) == ((int(out0[g, i, 0]) + int(want[g, i, 0])) % P,
                                        (int(out0[g, i, 1]) + int(want[g, i, 1])) % P)
    # the sums were cleared: another combined call adds nothing
    again = out.copy()
    sink.combined(again)
    assert np.array_equal(again, out)
    sink.free()


def test_shape_errors(dev):
    import tessera_amd as T
    from tessera_amd.sinks import ColumnSink
    with pytest.raises(T.BadShape, match="NoQueries"):
        ColumnSink(dev, 4, 1, 0, 8, [])
    with pytest.raises(T.BadShape, match="QueryRange"):
        ColumnSink(dev, 4, 1, 0, 8, [3, 32])
    sink = ColumnSink(dev, 4, 1, 2, 8, [0, 5, 31])
    rows, masks = splitmix_uniform(16 * 9, 1), splitmix_uniform(2 * 9, 2)
    with pytest.raises(T.BadShape, match="a block"):
        sink.select(rows, masks)  # nine rows in a sink of blocks of eight
    c3 = np.zeros((3, 8, 2), np.uint64)
    sink.select_combining(rows[:128], masks[:16], c3)
    with pytest.raises(T.BadShape, match="group count"):
        sink.select_combining(rows[:128], masks[:16], np.zeros((2, 8, 2), np.uint64))
    with pytest.raises(T.BadShape, match="one combined row per group"):
        sink.combined(np.zeros((2, 16, 2), np.uint64))
    sink.free()
