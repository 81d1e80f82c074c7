"""tsr_merkle_root_noalloc, the entry point that allocates nothing (its tree lives in a fixed
device array of the library's own): through the C ABI its roots equal tsr_merkle_root's and the
oracle's, on calls in a row whose second is smaller (the array keeps the last call's contents),
with device and host leaves; it takes no memory from the context and declines trees above its
2^16-leaf maximum."""
import ctypes

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def roots(dev, leaves_bytes, n, on_device):
    from tessera_amd.lib import check, lib
    L = lib()
    host = np.frombuffer(leaves_bytes, dtype=np.uint8).copy()
    buf = dev.upload(host) if on_device else None
    src = buf.ptr if on_device else host.ctypes.data
    plain = ctypes.create_string_buffer(32)
    check(L.tsr_merkle_root(dev.handle, src, n, plain))
    before = dev.memory()
    got = ctypes.create_string_buffer(32)
    check(L.tsr_merkle_root_noalloc(dev.handle, src, n, got))
    assert dev.memory() == before  # nothing taken from the context's pool
    if buf is not None:
        buf.free()
    return plain.raw, got.raw


@pytest.mark.parametrize("on_device", [True, False])
def test_noalloc_root_equals_merkle_root(dev, on_device):
    rng = np.random.default_rng(7)
    # large then smaller: the second call reads none of the first call's tree
    for log_n in (16, 11, 3, 0, 13):
        n = 1 << log_n
        leaves = rng.integers(0, 256, 32 * n, dtype=np.uint8).tobytes()
        print(f"noalloc merkle root: n = 2^{log_n}, device leaves {on_device}", flush=True)
        plain, got = roots(dev, leaves, n, on_device)
        assert got == plain, log_n
        if log_n <= 13:
            nodes = O.merkle_tree([leaves[32 * i:32 * i + 32] for i in range(n)], threads=4)
            assert got == bytes(nodes[-32:]), log_n


def test_noalloc_limits(dev):
    from tessera_amd.lib import BAD_SHAPE, DECLINED, lib
    L = lib()
    out = ctypes.create_string_buffer(32)
    big = dev.buffer(8 * (1 << 17), dtype=np.uint32)  # 2^17 leaves: above the maximum
    assert L.tsr_merkle_root_noalloc(dev.handle, big.ptr, 1 << 17, out) == DECLINED
    assert b"2^16" in L.tsr_last_error()
    assert L.tsr_merkle_root_noalloc(dev.handle, big.ptr, 3, out) == BAD_SHAPE
    big.free()
