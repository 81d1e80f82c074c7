"""Batched commitments: every root of a batch of matrices of mixed shapes equals the oracle's
commit_unmasked root and the plain per-matrix GPU call's, with and without autotuning and
hipGraph capture; graph replays interleaved with eager work all give the eager roots."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import batched_shapes, splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def make_batch(shapes, seed):
    mats = []
    for i, (lk, n) in enumerate(shapes):
        rows = splitmix_uniform(n << lk, seed + i)
        cid = bytes(((seed + i) * 13 + j) % 256 for j in range(32))
        mats.append((rows, lk, 2, 32, cid))
    return mats


@pytest.mark.parametrize("autotune,graph,serial", [(False, False, False), (True, False, False), (True, True, False),
                                                   (False, False, True), (True, True, True)])
def test_batch_roots_match_oracle(dev, autotune, graph, serial):
    from tessera_amd.batch import CommitBatch
    shapes = batched_shapes(12, seed=7, log_k=(1, 13), cols=(1, 70)) + [(0, 3), (12, 0)]
    mats = make_batch(shapes, 100)
    b = CommitBatch(dev, mats, autotune=autotune, graph=graph, serial=serial)
    b.run()
    got = b.roots()
    for (rows, lk, lr, blk, cid), root in zip(mats, got):
        _, nodes = O.commit_unmasked(rows, lk, lr, blk, cid, threads=8)
        assert root == nodes[-32:], (lk, rows.size >> lk)
    b.free()


@pytest.mark.parametrize("plan", [(13, 0), (13, 7), (12, 0), (12, 6), (13, 3)])
def test_every_plan_matches_oracle(dev, plan):
    """Each plan the tuner chooses among (the LDS tile and the pass split) gives the oracle's
    roots, on shapes whose passes it splits differently (log2 k up to 18)."""
    from tessera_amd.batch import CommitBatch
    shapes = [(18, 3), (16, 5), (14, 9), (11, 17), (6, 4)]
    mats = make_batch(shapes, 300)
    b = CommitBatch(dev, mats, autotune=False, graph=True, plan=plan)
    b.run()
    for (rows, lk, lr, blk, cid), root in zip(mats, b.roots()):
        _, nodes = O.commit_unmasked(rows, lk, lr, blk, cid, threads=8)
        assert root == nodes[-32:], (plan, lk)
    assert {p[2:4] for p in b.tuning()} == {plan}
    b.free()


def test_graph_replays_match_eager(dev):
    import tessera_amd as T
    from tessera_amd.batch import CommitBatch
    mats = make_batch(batched_shapes(6, seed=3, log_k=(10, 14), cols=(16, 40)), 7)
    eager = CommitBatch(dev, mats, autotune=False, graph=False)
    eager.run()
    want = eager.roots()
    g = CommitBatch(dev, mats, autotune=True, graph=True)
    for rep in range(4):
        g.run()
        assert g.roots() == want, rep
        eager.run()  # eager work between the replays
        c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(mats[0][1], 2, 0, 32, mats[0][4]), mats[0][0])
        assert c.root() == want[0]
        c.free()
    eager.free()
    g.free()
    # the plain per-matrix call agrees for every matrix
    for (rows, lk, lr, blk, cid), root in zip(mats, want):
        c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(lk, lr, 0, blk, cid), rows)
        assert c.root() == root
        c.free()


@pytest.fixture(scope="module")
def config4(dev):
    """The batched-commit workload at its real shapes: 64 matrices, log2 k uniform in [10, 20],
    N uniform in [16, 512] (about 2.0 G elements, 64 GiB of codewords), filled on the device,
    committed by one autotuned, graph-captured batch. Yields (shapes, row buffers, batch roots)."""
    from tessera_amd.batch import CommitBatch
    shapes = batched_shapes()
    bufs = [dev.buffer(n << lk).fill_uniform(0xC4 + i) for i, (lk, n) in enumerate(shapes)]
    cid = bytes([7] * 32)
    mats = [(buf, lk, 2, 32, cid) for buf, (lk, n) in zip(bufs, shapes)]
    print(f"config 4: {len(shapes)} matrices, {sum(n << lk for lk, n in shapes)} elements", flush=True)
    b = CommitBatch(dev, mats, autotune=True, graph=True)
    b.run()
    roots = b.roots()
    b.run()  # a replay gives the same roots
    assert b.roots() == roots
    b.free()
    yield shapes, bufs, cid, roots
    for buf in bufs:
        buf.free()


def test_config4_batch_equals_plain_commit(dev, config4):
    """Every batch root equals the plain per-matrix commit_unmasked root, including the
    matrices of log2 k 18..20 with N above 256 that run the 9- and 10-bit strided passes."""
    import tessera_amd as T
    shapes, bufs, cid, roots = config4
    assert sorted(set(lk for lk, _ in shapes)) == list(range(10, 21))
    assert any(lk >= 18 and n > 256 for lk, n in shapes)
    for (lk, n), buf, root in zip(shapes, bufs, roots):
        c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(lk, 2, 0, 32, cid), buf)
        assert c.root() == root, (lk, n)
        c.free()


@pytest.mark.parametrize("lo,hi", [(0, 22), (22, 24), (24, 26)])
def test_config4_batch_equals_oracle(dev, config4, lo, hi):
    """Batch roots against the oracle's commit_unmasked for every matrix of 2^lo < N k <= 2^hi
    elements (57 of the 64, log2 k 10..18 and 20)."""
    shapes, bufs, cid, roots = config4
    checked = 0
    for (lk, n), buf, root in zip(shapes, bufs, roots):
        if not (1 << lo) < (n << lk) <= (1 << hi):
            continue
        _, nodes = O.commit_unmasked(buf.download(), lk, 2, 32, cid, threads=16)
        assert root == nodes[-32:], (lk, n)
        checked += 1
    assert checked > 0
