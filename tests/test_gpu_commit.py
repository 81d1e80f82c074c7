"""Parity of the gfx950 commitment path (encode, column chains, Merkle, answer) with the oracle.

Bit-exact equality everywhere: codewords, chain digests, leaves, roots, opened columns and
multiproof siblings. Small shapes are compared whole against the oracle and the reference's
golden vectors; the full 256 x 2^20 rate-1/4 commitment is checked through sampled rows,
sampled columns' chains, the Merkle root over all leaves and verified openings.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, h, hexes, load, splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def test_encoding_golden_rate_half(dev):
    import tessera_amd as T
    for case in load("encoding.json")["encoding"]:
        k, P_ = case["k"], case["mask_coeffs"]
        enc = T.RowEncoder(dev, k, 1, P_)
        cw = enc.encode_rows(hexes(case["row"]), hexes(case["mask"]) if P_ else None)
        assert list(cw[0]) == list(hexes(case["codeword"])), (k, P_)


def test_encoding_golden_rate_quarter(dev):
    import tessera_amd as T
    for case in load("weights_encoding.json")["encoding"]:
        k, P_ = case["k"], case["mask_coeffs"]
        enc = T.RowEncoder(dev, k, 2, P_)
        cw = enc.encode_rows(hexes(case["row"]), hexes(case["mask"]) if P_ else None)
        assert list(cw[0]) == list(hexes(case["codeword"])), (k, P_)


@pytest.mark.parametrize("log_k", [0, 1, 2, 5, 8, 10, 11, 12, 13, 14, 16, 18, 21, 22])
@pytest.mark.parametrize("log_r", [1, 2])
def test_encoding_matches_oracle(dev, log_k, log_r):
    """Rows of 2^log_k against the oracle; 2^21 at rate 1/4 and 2^22 at rate 1/2 run ten stage
    bits in one strided pass each side, 2^22 at rate 1/4 eleven bits in two."""
    import tessera_amd as T
    k = 1 << log_k
    rows = 3 if log_k <= 14 else 2
    data = splitmix_uniform(rows * k, 1000 + log_k)
    for P_ in sorted({0, 1, k // 2, k} if log_k <= 12 else {0, 200 if k >= 200 else 1}):
        masks = splitmix_uniform(rows * P_, 77) if P_ else None
        want = O.encode_rows(data, log_k, log_r, masks=masks, mask_coeffs=P_, threads=8)
        got = T.RowEncoder(dev, k, log_r, P_).encode_rows(data, masks)
        assert np.array_equal(got, want), (log_k, log_r, P_)


def test_encoding_from_device_buffers(dev):
    import tessera_amd as T
    log_k, n_rows = 12, 5
    data = splitmix_uniform(n_rows << log_k, 5)
    src = dev.upload(data)
    out = dev.buffer(n_rows << (log_k + 2))
    T.RowEncoder(dev, 1 << log_k, 2).encode_rows(src, out=out)
    want = O.encode_rows(data, log_k, 2, threads=8).reshape(-1)
    assert np.array_equal(out.download(), want)


def test_shape_errors_before_work(dev):
    import tessera_amd as T
    with pytest.raises(T.BadShape, match="InverseRate"):
        T.RowEncoder(dev, 4, 0).encode_rows(np.zeros(4, dtype=np.uint64))
    with pytest.raises(T.BadShape, match="TooLong"):
        T.RowEncoder(dev, 1 << 31, 2).encode_rows(np.zeros(0, dtype=np.uint64))
    with pytest.raises(T.BadShape, match="MaskTooLarge"):
        T.RowEncoder(dev, 4, 1, 5).encode_rows(np.zeros(4, dtype=np.uint64), np.zeros(5, np.uint64))
    params = T.CommitmentParamsPy(4, 2, 1, 4)
    with pytest.raises(T.BadShape, match="UnexpectedMask"):
        T.ReferenceCommitment.commit_unmasked(dev, params, np.zeros(16, dtype=np.uint64))


def test_column_chains_golden(dev):
    import tessera_amd as T
    for case in load("hash.json")["columns"]:
        cols, block = case["columns"], case["block"]
        rows = np.array([[h(x) for x in r] for r in case["rows"]], dtype=np.uint64).reshape(-1)
        ch = T.ColumnChains(dev, bytes.fromhex(case["commitment_id"]), cols, block)
        m = len(case["rows"])
        for r in range(0, m, block):
            ch.absorb_block(rows[r * cols:min(m, r + block) * cols])
        assert [d.hex() for d in ch.finish()] == case["leaves"]


def test_column_chain_block_errors(dev):
    import tessera_amd as T
    ch = T.ColumnChains(dev, bytes(32), 3, 4)
    with pytest.raises(T.BadShape, match="BlockSize"):
        ch.absorb_block(np.zeros(15, dtype=np.uint64))
    ch.absorb_block(np.zeros(6, dtype=np.uint64))  # 2 rows: a short final block
    with pytest.raises(T.BadShape, match="AfterFinalBlock"):
        ch.absorb_block(np.zeros(3, dtype=np.uint64))


def test_merkle_golden(dev):
    import tessera_amd as T
    for case in load("hash.json")["merkle"]:
        leaves = [bytes.fromhex(x) for x in case["leaves"]]
        assert T.merkle_root(dev, leaves).hex() == case["root"]


@pytest.mark.parametrize("log_n", [11, 12, 13, 17, 20, 21])
def test_merkle_root_sizes(dev, log_n):
    """Trees around the build's launch boundaries (one launch per level above 2^20 nodes, nine
    levels per launch down to 2048 nodes, the rest in one workgroup) against the oracle."""
    import tessera_amd as T
    n = 1 << log_n
    raw = splitmix_uniform(4 * n, 1000 + log_n).view(np.uint8).tobytes()
    leaves = [raw[32 * i:32 * i + 32] for i in range(n)]
    assert T.merkle_root(dev, leaves) == O.merkle_tree(leaves, threads=8)[-32:]


def commit_and_check_small(dev, log_k, log_r, block, rows, seed):
    import tessera_amd as T
    k, n = 1 << log_k, 1 << (log_k + log_r)
    data = splitmix_uniform(rows * k, seed)
    cid = bytes((seed * 7 + i) % 256 for i in range(32))
    cw, nodes = O.commit_unmasked(data, log_k, log_r, block, cid, threads=8)
    c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(log_k, log_r, 0, block, cid), data)
    assert c.root() == nodes[-32:]
    assert b"".join(c.leaves()) == nodes[:32 * n]
    queries = [int(x) % n for x in splitmix_uniform(9, seed + 1)] + [0, n - 1, 0]
    ans = c.answer(queries)
    want_idx, want_cols, want_sib = O.answer(cw, nodes, n, queries)
    assert ans.indices == want_idx
    assert all(np.array_equal(a, b) for a, b in zip(ans.columns, want_cols))
    assert ans.siblings == want_sib
    c.free()


@pytest.mark.parametrize("log_k,log_r,block,rows", [(0, 1, 1, 1), (3, 2, 4, 10), (6, 1, 32, 33),
                                                    (10, 2, 32, 64), (12, 2, 124, 130)])
def test_commit_unmasked_small(dev, log_k, log_r, block, rows):
    commit_and_check_small(dev, log_k, log_r, block, rows, seed=log_k * 31 + rows)


def test_commit_no_rows_has_init_leaves(dev):
    import tessera_amd as T
    cid = bytes(range(32))
    c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(2, 1, 0, 4, cid),
                                              np.zeros(0, dtype=np.uint64))
    leaves = O.column_chains(cid, np.zeros(0, dtype=np.uint64), 8, 4)
    assert c.leaves() == leaves
    assert c.root() == O.merkle_tree(leaves)[-32:]


def test_commit_config1_shape(dev):
    """256 rows of k = 2^14 at rate 1/4, B = 32: the whole commitment against the oracle."""
    commit_and_check_small(dev, 14, 2, 32, 256, seed=2024)


def test_answer_errors(dev):
    import tessera_amd as T
    c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(3, 1, 0, 4), splitmix_uniform(16, 1))
    with pytest.raises(T.BadShape, match="NoQueries"):
        c.answer([])
    with pytest.raises(T.BadShape, match="QueryRange"):
        c.answer([1, 16])


def test_commit_large_properties(dev):
    """256 x 2^20 rate 1/4 (8 GiB of codewords, 2^22 leaves), device-resident input: sampled
    rows against the oracle encoder, sampled columns' chains recomputed from the device's
    codewords, the root over all leaves, and 64 verified openings."""
    import tessera_amd as T
    log_k, log_r, B, rows = 20, 2, 32, 256
    k, n = 1 << log_k, 1 << (log_k + log_r)
    src = dev.buffer(rows * k).fill_uniform(0x5EED)
    cid = bytes([7] * 32)
    c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(log_k, log_r, 0, B, cid), src)
    host_rows = src.download()
    # the device generator equals the host SplitMix64-with-rejection stream
    assert np.array_equal(host_rows[:4096], splitmix_uniform(4096, 0x5EED))
    ans = c.answer([0, 1, n - 1] + [int(x) % n for x in splitmix_uniform(61, 9)])
    # sampled rows: the oracle's encoding of the row equals the device codeword at every opened column
    for r in (0, 137, 255):
        want = O.encode_rows(host_rows[r * k:(r + 1) * k], log_k, log_r, threads=8)[0]
        for j, col in zip(ans.indices, ans.columns):
            assert int(col[r]) == int(want[j])
    # opened columns hash to the leaves the tree commits to, and the multiproof verifies
    leaves = c.leaves()
    for j, col in zip(ans.indices, ans.columns):
        assert O.column_chain_at(cid, j, col, B) == leaves[j]
    assert O.merkle_tree(leaves, threads=8)[-32:] == c.root()
    assert O.merkle_verify(c.root(), n, [(j, leaves[j]) for j in ans.indices], ans.siblings)
    c.free()
    src.free()
