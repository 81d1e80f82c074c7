"""Failure handling forced from the host, so every kernel still exits normally: a memory budget
refuses allocations at every point of a call (TSR_OOM), and the call must leave no partial
result (no commitment, no moved transcript) and give back every byte it took. This is the
reference's decline contract (crates/tessera-gpu/src/level.rs:91-97,122-133: a device that
cannot hold the work declines before any side effect) and its error order (shape checks before
any launch or transcript move, crates/tessera-gpu/src/device.rs:30-53)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import splitmix_uniform

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def test_budget_refusal_leaves_no_partial_commitment(dev):
    """Sweep the budget from nothing to enough: each commit either fails with OOM, taking no
    memory with it, or succeeds with the oracle's root."""
    import tessera_amd as T
    from tessera_amd.lib import OOM, TsrError
    log_k, rows = 10, 40
    data = splitmix_uniform(rows << log_k, 5)
    cid = bytes([9] * 32)
    params = T.CommitmentParamsPy(log_k, 2, 0, 32, cid)
    _, nodes = O.commit_unmasked(data, log_k, 2, 32, cid, threads=4)
    d_rows = dev.upload(data)
    # the context caches the encoder's tables and its encode scratch on first use
    warm = T.ReferenceCommitment.commit_unmasked(dev, params, d_rows)
    assert warm.root() == nodes[-32:]
    warm.free()
    base = dev.memory()
    outcomes = []
    for budget in [1, 4096, 1 << 16, 1 << 18, 1 << 20, 3 << 20, 1 << 22, 6 << 20, 1 << 23, 1 << 24,
                   1 << 26, 1 << 30]:
        dev.set_budget(base + budget)
        try:
            c = T.ReferenceCommitment.commit_unmasked(dev, params, d_rows)
        except TsrError as e:
            assert e.status == OOM, e
            assert dev.memory() == base, budget  # every byte of the failed call came back
            outcomes.append("oom")
            continue
        assert c.root() == nodes[-32:], budget
        c.free()
        assert dev.memory() == base
        outcomes.append("ok")
    dev.set_budget(0)
    assert outcomes[0] == "oom" and outcomes[-1] == "ok", outcomes
    d_rows.free()


def test_budget_refusal_leaves_transcript_untouched(dev):
    """A sumcheck refused for memory leaves the transcript as it was: the same proof follows
    once the budget allows it."""
    from tessera_amd.lib import OOM, TsrError
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript
    n = 16
    a, b = splitmix_uniform(1 << n, 11), splitmix_uniform(1 << n, 12)
    want_t = Transcript(b"ctx")
    want = prove_batch(dev, want_t, "stage/0", 2, [[Table.base(a), Table.base(b)]])
    t = Transcript(b"ctx")
    state, counter = t.state, t.counter
    base = dev.memory()
    refused = 0
    for budget in [1, 1 << 12, 1 << 16, 1 << 19, 1 << 21]:
        dev.set_budget(base + budget)
        try:
            got = prove_batch(dev, t, "stage/0", 2, [[Table.base(a), Table.base(b)]])
        except TsrError as e:
            assert e.status == OOM, e
            assert (t.state, t.counter) == (state, counter), budget
            assert dev.memory() == base
            refused += 1
            continue
        break
    dev.set_budget(0)
    assert refused >= 1
    got = prove_batch(dev, t, "stage/0", 2, [[Table.base(a), Table.base(b)]])
    assert np.array_equal(got.rounds, want.rounds)
    assert (t.state, t.counter) == (want_t.state, want_t.counter)


def test_memory_accounting_and_foreign_free(dev):
    """tsr_free gives back exactly what tsr_alloc charged, and refuses a pointer it did not
    hand out without touching the accounting."""
    import ctypes

    from tessera_amd.lib import BAD_ARG, lib
    base = dev.memory()
    bufs = [dev.buffer(n) for n in (1, 1000, 4097, 1 << 20, 3 << 19)]
    assert dev.memory() > base
    for b in bufs:
        b.free()
    assert dev.memory() == base
    host = (ctypes.c_uint64 * 4)()
    assert lib().tsr_free(dev.handle, ctypes.addressof(host)) == BAD_ARG
    assert dev.memory() == base
