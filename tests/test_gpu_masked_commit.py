"""The masked advice commitment on the device (ReferenceCommitment::commit and from_encoded)
against the reference's opening vectors (root, leaves at the queries, answered columns and
multiproof) and against the oracle at larger sizes, including layouts of more than 32 segments
(the column-merge input spans several BLAKE3 chunks) and empty segments."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import h, load, splitmix_uniform

pytestmark = pytest.mark.gpu

CASES = load("pcs_opening.json")["openings"]


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_opening_vectors(dev, case):
    import tessera_amd as T
    rows = np.array([h(x) for r in case["witness"]["rows"] for x in r], dtype=np.uint64)
    masks = np.array([h(x) for m in case["witness"]["masks"] for x in m], dtype=np.uint64)
    seg = [len(s) for s in case["segments"]]
    params = T.CommitmentParamsPy(case["log2_k"], case["log2_inv_rate"], case["mask_coeffs"],
                                  case["row_block"], bytes.fromhex(case["commitment_id"]))
    c = T.ReferenceCommitment.commit(dev, params, seg, rows, masks)
    assert c.root().hex() == case["root"]
    leaves = c.leaves()
    distinct = sorted(set(case["queries"]))
    assert [leaves[j].hex() for j in distinct] == case["intermediate"]["leaves"]
    ans = c.answer(case["queries"])
    assert ans.indices == distinct
    assert [[int(x) for x in col] for col in ans.columns] == [[h(x) for x in col] for col in case["columns"]]
    assert [s.hex() for s in ans.siblings] == case["siblings"]
    c.free()


@pytest.mark.parametrize("log_k,log_r,P_,B,seg", [
    (10, 1, 37, 32, [70, 0, 130, 5]),
    (12, 1, 4096, 124, [1, 300]),
    (4, 1, 3, 2, [i % 3 for i in range(40)]),      # 40 segments: two merge chunks
    (3, 1, 8, 1, [1] * 70 + [0, 2]),                 # 72 segments, three chunks
    (6, 2, 5, 4, [0]),                               # one empty segment: the leaves merge d_j[0]
])
def test_against_oracle(dev, log_k, log_r, P_, B, seg):
    import tessera_amd as T
    k, N = 1 << log_k, sum(seg)
    rows = splitmix_uniform(N * k, 7 + log_k)
    masks = splitmix_uniform(N * P_, 11 + log_k)
    cid = bytes(range(32, 64))
    cw, leaves, nodes = O.commit_masked(rows, masks, log_k, log_r, P_, B, cid, seg, threads=8)
    params = T.CommitmentParamsPy(log_k, log_r, P_, B, cid)
    c = T.ReferenceCommitment.commit(dev, params, seg, rows, masks)
    assert c.leaves() == leaves
    assert c.root() == nodes[-32:]
    c.free()
    # from_encoded with the oracle's codewords gives the same commitment
    e = T.ReferenceCommitment.from_encoded(dev, params, seg, cw.reshape(-1))
    assert e.root() == nodes[-32:]
    e.free()


def test_from_encoded_chained_equals_unmasked(dev):
    import tessera_amd as T
    log_k, log_r, B, N = 9, 2, 32, 45
    rows = splitmix_uniform(N << log_k, 3)
    cid = bytes([5] * 32)
    cw, nodes = O.commit_unmasked(rows, log_k, log_r, B, cid, threads=8)
    params = T.CommitmentParamsPy(log_k, log_r, 0, B, cid)
    e = T.ReferenceCommitment.from_encoded(dev, params, [N], cw.reshape(-1), merged=False)
    assert e.root() == nodes[-32:]
    e.free()


def test_shape_errors(dev):
    import tessera_amd as T
    rows = splitmix_uniform(16, 1)
    with pytest.raises(T.BadShape, match="NoMask"):
        T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(3, 1, 0, 4), [2], rows, [])
    with pytest.raises(T.BadShape, match="MaskTooLarge"):
        T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(3, 1, 9, 4), [2], rows, splitmix_uniform(18, 2))
    with pytest.raises(T.BadShape, match="NoSegments"):
        T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(3, 1, 1, 4), [], rows, splitmix_uniform(2, 2))
    with pytest.raises(T.BadShape):
        T.ReferenceCommitment.from_encoded(dev, T.CommitmentParamsPy(3, 1, 0, 4), [1, 1],
                                           np.zeros(32, dtype=np.uint64), merged=False)


STREAMING = load("pcs_streaming.json")["cases"]


@pytest.mark.parametrize("case", STREAMING, ids=[c["name"] for c in STREAMING])
def test_streaming_vectors(dev, case):
    """The reference's streaming advice commitments: the whole commitment at once (leaves, root),
    and the streaming path, one BlockSink per segment fed block by block (each segment's digests)."""
    import tessera_amd as T
    from tessera_amd.sinks import BlockSink
    log_k = case["layout"]["log2_k"]
    k, n = 1 << log_k, len(case["leaves"])
    log_r = n.bit_length() - 1 - log_k
    P_, B = case["mask_coeffs"], case["row_block"]
    rows = np.array([h(x) for r in case["rows"] for x in r], dtype=np.uint64)
    masks = np.array([h(x) for m in case["masks"] for x in m], dtype=np.uint64)
    seg = [len(s) for s in case["segments"]]
    cid = bytes.fromhex(case["commitment_id"])
    c = T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(log_k, log_r, P_, B, cid), seg, rows, masks)
    assert [x.hex() for x in c.leaves()] == case["leaves"]
    assert c.root().hex() == case["root"]
    c.free()
    off = 0
    for s, count in enumerate(seg):
        sink = BlockSink(dev, log_k, log_r, P_, B, cid)
        for r0 in range(off, off + count, B):
            r1 = min(off + count, r0 + B)
            sink.absorb(rows[r0 * k:r1 * k], masks[r0 * P_:r1 * P_])
        assert [d.hex() for d in sink.finish()] == case["segment_digests"][s]
        sink.free()
        off += count


def fp2s(xs):
    from tests.helpers import fp2
    return [fp2(x) for x in xs]


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_open_matches_vectors(dev, case):
    """ReferenceCommitment::open on the device: the combined rows u, M of every group and of the
    proximity row (RFC-0011 steps 2 to 5)."""
    import tessera_amd as T
    rows = np.array([h(x) for r in case["witness"]["rows"] for x in r], dtype=np.uint64)
    masks = np.array([h(x) for m in case["witness"]["masks"] for x in m], dtype=np.uint64)
    seg = [len(s) for s in case["segments"]]
    params = T.CommitmentParamsPy(case["log2_k"], case["log2_inv_rate"], case["mask_coeffs"],
                                  case["row_block"], bytes.fromhex(case["commitment_id"]))
    c = T.ReferenceCommitment.commit(dev, params, seg, rows, masks)
    groups = [(fp2s(g["column_point"]), [(fp2s(cl["row_point"]), fp2s([cl["gamma"]])[0]) for cl in g["claims"]])
              for g in case["groups"]]
    got = c.open(groups, fp2s(case["xi_point"]), [r for s in case["segments"] for r in s], case["row_vars"])
    for (u, m), want in zip(got, case["combined"]):
        assert [tuple(int(v) for v in x) for x in u] == fp2s(want["u"])
        assert [tuple(int(v) for v in x) for x in m] == fp2s(want["mask"])
    c.free()


def test_open_against_oracle(dev):
    """A larger opening: 300 rows of k = 2^10 over 10 row variables, 3 groups, P = 37."""
    import tessera_amd as T
    rng = np.random.default_rng(4)
    log_k, P_, row_vars = 10, 37, 10
    idx = np.sort(rng.choice(1 << row_vars, 300, replace=False)).astype(np.uint64)
    seg = [120, 0, 180]
    rows = splitmix_uniform(300 << log_k, 41)
    masks = splitmix_uniform(300 * P_, 42)
    rnd = lambda *s: rng.integers(0, (1 << 64) - (1 << 32), s, dtype=np.uint64)
    groups = [(rnd(log_k, 2), [(rnd(row_vars, 2), rnd(2)) for _ in range(nc)]) for nc in (2, 0, 3)]
    xi = rnd(row_vars, 2)
    c = T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(log_k, 1, P_, 32, bytes(32)), seg, rows, masks)
    got = c.open(groups, xi, idx, row_vars)
    co = O.row_coefficients(row_vars, idx, [cl for _, cl in groups], xi)
    u = O.combine(co, rows, 1 << log_k)
    m = O.combine(co, masks, P_)
    for g, (gu, gm) in enumerate(got):
        assert np.array_equal(gu, u[g]) and np.array_equal(gm, m[g])
    with pytest.raises(T.BadShape, match="XiPoint"):
        c.open(groups, xi[:3], idx, row_vars)
    c.free()


def test_mask_stream_vectors(dev):
    """The device mask stream against spec/vectors/hash prover_streams: the proof seed's 24 masks,
    and a seed whose seventh word is >= p, drawn whole and in pieces across the discarded word
    (host and device destinations)."""
    import tessera_amd as T
    from tessera_amd.pcs import MaskStream
    v = load("hash.json")["prover_streams"]
    seed = bytes.fromhex(v["proof_seed"])
    ms = MaskStream(dev, seed)
    assert [int(x) for x in ms.draw(len(v["advice_masks"]))] == [h(x) for x in v["advice_masks"]]
    assert ms.position == O.advice_masks(seed, len(v["advice_masks"]))[1]
    ms.close()
    rej = v["rejection"]
    want = [h(x) for x in rej["values"]]
    ms = MaskStream(dev, bytes.fromhex(rej["seed"]))
    assert [int(x) for x in ms.draw(len(want))] == want
    assert ms.position == len(want) + 1
    ms.close()
    ms = MaskStream(dev, bytes.fromhex(rej["seed"]))
    first = ms.draw(5)
    assert ms.position == 5
    buf = T.DeviceBuffer(dev, 10)
    ms.draw(10, out=buf)
    assert [int(x) for x in first] + [int(x) for x in buf.download()] == want
    assert ms.position == 16
    assert ms.draw(0).size == 0 and ms.position == 16
    ms.close()


def test_mask_stream_large_draws(dev):
    """Draws of 1, 7, 2^20 + 3 and 2^22 - 5 masks in a row, starting mid-block, against the
    oracle's single draw of their sum."""
    import tessera_amd as T
    from tessera_amd.pcs import MaskStream
    seed = bytes(range(100, 132))
    sizes = [1, 7, (1 << 20) + 3, (1 << 22) - 5]
    want, end = O.advice_masks(seed, sum(sizes))
    ms = MaskStream(dev, seed)
    got = []
    for s in sizes:
        buf = T.DeviceBuffer(dev, s)
        ms.draw(s, out=buf)
        got.append(buf.download())
    assert np.array_equal(np.concatenate(got), want)
    assert ms.position == end
    ms.close()


def test_commit_with_stream_masks(dev):
    """A masked commitment whose masks the device stream draws (N x P, m_0 first, row after row)
    equals the oracle's commitment with the oracle's draws of the same seed."""
    import tessera_amd as T
    from tessera_amd.pcs import MaskStream
    log_k, log_r, P_, B, seg = 9, 1, 17, 32, [40, 0, 25]
    k, N = 1 << log_k, sum(seg)
    seed = bytes([9] * 32)
    rows = splitmix_uniform(N * k, 77)
    masks, _ = O.advice_masks(seed, N * P_)
    cid = bytes(range(32))
    _, leaves, nodes = O.commit_masked(rows, masks, log_k, log_r, P_, B, cid, seg, threads=8)
    ms = MaskStream(dev, seed)
    dmasks = T.DeviceBuffer(dev, N * P_)
    ms.draw(N * P_, out=dmasks)
    c = T.ReferenceCommitment.commit(dev, T.CommitmentParamsPy(log_k, log_r, P_, B, cid), seg, rows, dmasks)
    assert c.leaves() == leaves and c.root() == nodes[-32:]
    c.free()
    ms.close()
