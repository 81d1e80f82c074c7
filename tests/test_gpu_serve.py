"""Served proving: a served request's tokens, witness root and proof transcript equal the oracle's
(the model, commitment and sumcheck restated on the CPU); every result is the same alone (the
plain per-request call) and in a concurrent mix; admission respects the device-memory budget."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

SEED = 0x7E55E7A


@pytest.fixture(scope="module")
def weights():
    return O.model_weights(SEED)


@pytest.mark.parametrize("rid,n_prompt,n_output", [(1, 1, 15), (2, 20, 80), (3, 100, 412)])
def test_served_request_matches_oracle(weights, rid, n_prompt, n_output):
    from tessera_amd.serve import Server
    rng = np.random.default_rng(rid)
    prompt = rng.integers(0, 4096, n_prompt).tolist()
    with Server(slots=1, model_seed=SEED) as srv:
        r = srv.prove(rid, prompt, n_output)
    toks, root, state, counter, _ = O.served_proof(weights, rid, prompt, n_output)
    assert list(r.tokens[:r.n_tokens]) == [int(x) for x in toks]
    assert bytes(r.root) == root
    assert bytes(r.transcript_state) == state and r.transcript_counter == counter


def test_alone_and_in_the_mix_are_equal():
    from tessera_amd.serve import Server, footprint, run_served, workload
    reqs = workload(16, seed=5, rate_hz=400.0, min_len=16, max_len=200)
    budget = 4 * footprint(200)
    with Server(slots=6, hbm_budget=budget, model_seed=SEED) as srv:
        alone = {rid: srv.prove(rid, p, n).key() for rid, p, n, _ in reqs}
        mixed, _, _ = run_served(srv, reqs)
        max_active, in_flight = srv.stats()
    assert in_flight == 0
    assert 1 < max_active <= 6
    for rid, r in mixed.items():
        assert r.status == 0
        assert r.key() == alone[rid], rid


def test_grouped_commits_equal_alone():
    """A burst to one prover: the decoded requests waiting for it are committed in shared
    launches (result.group > 1, group 16), and every result equals its plain call's; a server
    that commits one request at a time (group 1, the default) gives the same results."""
    from tessera_amd.serve import Server, workload
    reqs = workload(24, seed=11, rate_hz=1e9, min_len=16, max_len=300)
    with Server(slots=24, provers=1, chunk=64, group=16, model_seed=SEED) as srv:
        alone = {rid: srv.prove(rid, p, n).key() for rid, p, n, _ in reqs}
        for rid, p, n, _ in reqs:
            srv.submit(rid, p, n)
        grouped = [srv.wait() for _ in reqs]
    with Server(slots=24, provers=1, chunk=64, group=1, model_seed=SEED) as srv:
        for rid, p, n, _ in reqs:
            srv.submit(rid, p, n)
        single = [srv.wait() for _ in reqs]
    assert max(r.group for r in grouped) > 1
    assert all(r.group == 1 for r in single)
    for r in grouped + single:
        assert r.status == 0
        assert r.key() == alone[r.id], r.id


def _witness_commitment(weights, prompt, total):
    """The oracle's codewords and tree of a request's witness (rows of 2^13, rate 1/4, B = 32)."""
    _, wit, _ = O.model_decode(weights, prompt, total)
    k = 1 << 13
    rows = -(-wit.size // k)
    padded = np.zeros(rows * k, dtype=np.uint64)
    padded[:wit.size] = wit
    return O.commit_unmasked(padded, 13, 2, 32, bytes([0x5E] * 32), threads=8)


def test_resident_commitments_open_and_release(weights):
    """With resident commitments each proved request's witness codewords and Merkle tree stay on
    the device: openings equal the oracle's answer over the same witness (grouped and single
    commits alike), their bytes are held until released, a released request can no longer be
    opened, and a resident request's id is not reused until released."""
    from tessera_amd.lib import Declined, TsrError
    from tessera_amd.serve import Server
    rng = np.random.default_rng(21)
    reqs = [(200 + i, rng.integers(0, 4096, 3).tolist(), n) for i, n in enumerate([40, 130, 260, 20])]
    nsym = 1 << 15
    queries = [0, 7, nsym - 1, 12345, 7, 30000]
    for group in (16, 1):
        with Server(slots=4, provers=1, group=group, resident=True, model_seed=SEED) as srv:
            for rid, p, n in reqs:
                srv.submit(rid, p, n)
            res = {r.id: r for r in [srv.wait() for _ in reqs]}
            count, held = srv.resident()
            assert count == len(reqs) and held > 0
            for rid, p, n in reqs:
                r = res[rid]
                assert r.status == 0
                cw, nodes = _witness_commitment(weights, p, len(p) + n)
                assert bytes(r.root) == nodes[-32:]
                got = srv.open(r, queries)
                idx, cols, sib = O.answer(cw, nodes, nsym, queries)
                assert got.indices == idx and got.siblings == sib
                assert all(np.array_equal(a, b) for a, b in zip(got.columns, cols))
            with pytest.raises(TsrError):
                srv.submit(reqs[1][0], [1], 5)
            assert srv.release(reqs[0][0])
            assert not srv.release(reqs[0][0])
            with pytest.raises(Declined):
                srv.open(res[reqs[0][0]], [1])
            count2, held2 = srv.resident()
            assert count2 == len(reqs) - 1 and held2 < held
            for rid, _, _ in reqs[1:]:
                assert srv.release(rid)
            assert srv.resident() == (0, 0)


def test_admission_by_budget():
    from tessera_amd.serve import Server, footprint
    budget = footprint(300) + footprint(100) // 2  # room for one 300-token request at a time
    with Server(slots=4, hbm_budget=budget, model_seed=SEED) as srv:
        for rid in range(1, 5):
            srv.submit(rid, [rid], 299)
        res = [srv.wait() for _ in range(4)]
        max_active, _ = srv.stats()
    assert max_active == 1
    assert sorted(r.id for r in res) == [1, 2, 3, 4]
    # strict FIFO: with one request in flight at a time, each starts after the previous ended
    res.sort(key=lambda r: r.t_admit)
    for a, b in zip(res, res[1:]):
        assert b.t_admit >= a.t_done - 1e-3


def test_cancel_releases_slots_and_leaves_others_unchanged():
    """Cancelled requests (one still waiting for admission behind a long one, others decoding or
    waiting for the single prover) end with status CANCELLED and give back their slot, budget
    and memory; the remaining requests' results equal their plain calls, and the server goes on
    serving new requests."""
    from tessera_amd.lib import CANCELLED
    from tessera_amd.serve import Server, footprint
    rng = np.random.default_rng(9)
    reqs = [(100 + i, rng.integers(0, 4096, 4).tolist(), n) for i, n in enumerate([400, 60, 300, 30, 200, 50])]
    with Server(slots=2, provers=1, chunk=2, hbm_budget=4 * footprint(512), model_seed=SEED) as srv:
        alone = {rid: srv.prove(rid, p, n).key() for rid, p, n in reqs}
        for rid, p, n in reqs:
            srv.submit(rid, p, n)
        # the last request waits behind the first two (two slots): cancelled while queued
        assert srv.cancel(105)
        cancelled = {105}
        for rid in (100, 102):  # decoding, or already further along
            if srv.cancel(rid):
                cancelled.add(rid)
        assert not srv.cancel(999)  # never submitted
        res = {}
        for _ in reqs:
            r = srv.wait()
            res[r.id] = r
        assert set(res) == {rid for rid, _, _ in reqs}
        for rid, r in res.items():
            if rid in cancelled:
                assert r.status == CANCELLED, rid
            else:
                assert r.status == 0 and r.key() == alone[rid], rid
        max_active, in_flight = srv.stats()
        assert in_flight == 0 and max_active <= 2
        assert not srv.cancel(101)  # done
        # the freed slots serve a new request
        srv.submit(200, [5, 6], 20)
        r = srv.wait()
        assert r.status == 0 and r.id == 200 and r.n_tokens == 22
