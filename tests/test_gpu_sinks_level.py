"""The trait-shaped objects of the C ABI against the oracle: BlockSink (masked rows, blocks with a
short final block), WeightEncoding::matrix / MatrixSink (queried symbols and chains, declines), and
the level offload (DeviceLevel sums by brute force over small tables with Python integers, binds
to the oracle's multilinear evaluations)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("log_k,log_r,P_,block,rows", [(4, 1, 3, 4, 10), (8, 1, 200 % 257, 32, 70),
                                                      (6, 2, 0, 5, 5), (10, 1, 1, 124, 130)])
def test_block_sink_matches_oracle(dev, log_k, log_r, P_, block, rows):
    from tessera_amd.sinks import BlockSink
    k, n = 1 << log_k, 1 << (log_k + log_r)
    P_ = min(P_, k)
    data = splitmix_uniform(rows * k, 10 + log_k)
    masks = splitmix_uniform(rows * P_, 20 + log_k) if P_ else np.zeros(0, dtype=np.uint64)
    cid = bytes(range(32))
    sink = BlockSink(dev, log_k, log_r, P_, block, cid)
    for r0 in range(0, rows, block):
        r1 = min(rows, r0 + block)
        sink.absorb(data[r0 * k:r1 * k], masks[r0 * P_:r1 * P_] if P_ else None)
    got = sink.finish()
    cw = O.encode_rows(data, log_k, log_r, masks=masks if P_ else None, mask_coeffs=P_, threads=8)
    want = O.column_chains(cid, cw.reshape(-1), n, block)
    assert got == want
    sink.free()


def test_block_sink_errors(dev):
    import tessera_amd as T
    from tessera_amd.sinks import BlockSink
    sink = BlockSink(dev, 3, 1, 0, 4, bytes(32))
    sink.absorb(splitmix_uniform(2 * 8, 1))  # a short block: the last
    with pytest.raises(T.BadShape, match="AfterFinalBlock"):
        sink.absorb(splitmix_uniform(8, 2))
    with pytest.raises(T.BadShape):
        BlockSink(dev, 3, 1, 0, 125, bytes(32))
    with pytest.raises(T.BadShape, match="MaskTooLarge"):
        BlockSink(dev, 3, 1, 9, 4, bytes(32))


@pytest.mark.parametrize("log_k,log_r,block,rows,q", [(5, 2, 4, 9, 7), (12, 2, 32, 64, 40), (9, 1, 124, 124, 0)])
def test_matrix_sink_matches_oracle(dev, log_k, log_r, block, rows, q):
    from tessera_amd.sinks import WeightEncoding
    k, n = 1 << log_k, 1 << (log_k + log_r)
    data = splitmix_uniform(rows * k, 5 + log_k)
    queries = sorted({int(x) % n for x in splitmix_uniform(q, 3)}) if q else []
    cid = bytes([9] * 32)
    sink = WeightEncoding(dev).matrix(log_k, log_r, block, cid, n, queries)
    cw = O.encode_rows(data, log_k, log_r, threads=8).reshape(rows, n)
    for r0 in range(0, rows, block):
        r1 = min(rows, r0 + block)
        sym = sink.absorb(data[r0 * k:r1 * k])
        assert np.array_equal(sym, cw[r0:r1][:, queries].reshape(r1 - r0, len(queries)))
    assert sink.finish() == O.column_chains(cid, cw.reshape(-1), n, block)
    sink.free()


def test_matrix_sink_declines(dev):
    from tessera_amd.sinks import WeightEncoding
    w = WeightEncoding(dev)
    assert w.matrix(5, 2, 4, bytes(32), 64, [1]) is None      # columns != 4k
    assert w.matrix(5, 2, 125, bytes(32), 128, [1]) is None   # block above 124


# ---- level offload ----

def f2_mul(a, b):
    return ((a[0] * b[0] + 7 * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def brute_sums(shape, tabs, low, high, low_bits, lam):
    """The trait's definition, with Python integers."""
    pairs = len(tabs[0]) // 2
    out = []
    for x in (0, 2, 3):
        acc = (0, 0)
        for p in range(pairs):
            w = f2_mul(low[p & ((1 << low_bits) - 1)], high[p >> low_bits])
            v = [f2_add(t[2 * p], f2_mul((x, 0), f2_sub(t[2 * p + 1], t[2 * p]))) for t in tabs]
            if shape == 0:
                poly = f2_add(f2_add(f2_mul(v[0], v[3]), f2_mul(v[1], v[2])), f2_mul(lam, f2_mul(v[2], v[3])))
            else:
                poly = f2_mul(v[0], v[1])
            acc = f2_add(acc, f2_mul(w, poly))
        out.append(acc)
    return out


def rand_f2(rng, count):
    return [(int(a), int(b)) for a, b in rng.integers(0, P, (count, 2), dtype=np.uint64)]


@pytest.mark.parametrize("shape", [0, 1])
@pytest.mark.parametrize("m", [1, 2, 5, 7])
def test_level_sums_bind_firsts(dev, shape, m):
    from tessera_amd.level import LevelOffload
    rng = np.random.default_rng(100 * shape + m)
    nt = 4 if shape == 0 else 2
    tabs = [rand_f2(rng, 1 << m) for _ in range(nt)]
    lam = rand_f2(rng, 1)[0]
    level = LevelOffload(dev).load(shape, [np.array(t, dtype=np.uint64) for t in tabs], lam)
    point = []
    cur = [list(t) for t in tabs]
    for j in range(m):
        pairs = 1 << (m - j - 1)
        low_bits = min(j % 3, m - j - 1)
        low = rand_f2(rng, 1 << low_bits)
        high = rand_f2(rng, max(1, pairs >> low_bits))
        assert level.sums(low, high, low_bits) == brute_sums(shape, cur, low, high, low_bits, lam)
        r = rand_f2(rng, 1)[0]
        level.bind(r)
        point.append(r)
        cur = [[f2_add(t[2 * i], f2_mul(r, f2_sub(t[2 * i + 1], t[2 * i]))) for i in range(len(t) // 2)]
               for t in cur]
    # all variables bound: each table's first entry is its multilinear extension at the point
    want = [tuple(int(x) for x in O.mle_eval(np.array(t, dtype=np.uint64), True, np.array(point, dtype=np.uint64)))
            for t in tabs]
    assert level.firsts() == want
    level.free()


def test_level_shape_errors(dev):
    import tessera_amd as T
    from tessera_amd.level import LevelOffload
    lo = LevelOffload(dev)
    with pytest.raises(T.BadShape):
        lo.load(0, [np.zeros((4, 2), dtype=np.uint64)] * 2, (1, 0))
    level = lo.load(1, [np.ones((8, 2), dtype=np.uint64)] * 2)
    with pytest.raises(T.BadShape, match="cover the pairs"):
        level.sums([(1, 0)], [(1, 0)], 0)  # 4 pairs, weights cover 1
    level.free()
