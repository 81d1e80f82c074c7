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
