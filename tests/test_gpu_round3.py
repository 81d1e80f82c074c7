"""The advice opening's answer on the device emitted as Round3Response bytes (crates/tessera-
protocol/src/round3.rs:100-175) and checked as the verifier checks it
(verifier/crates/tv-core/src/opening.rs:467-620): the bytes read back by an independent TCE
reader; each opened column re-hashed to its leaf (one chain per segment, then column-merge) and
the multiproof verified against the device root; then every combined row of the device's P4
(groups, then the proximity row), re-encoded with its combined mask, agrees at every opened
column with sum_r c(r) column[r]. A changed symbol fails the check, so the check itself is
pinned. Run over the reference's opening vectors and a larger random layout."""
import struct

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


def read_round3(b):
    """An independent reader of the advice part of TCE(Round3Response) with no opened matrix."""
    at = 0

    def u64():
        nonlocal at
        v = struct.unpack_from("<Q", b, at)[0]
        at += 8
        return v
    cols = [[u64() for _ in range(u64())] for _ in range(u64())]
    sib = []
    for _ in range(u64()):
        sib.append(bytes(b[at:at + 32]))
        at += 32
    assert u64() == 0 and at == len(b)  # no weight matrix, no byte left
    return cols, sib


def fp2_mul(a, b):
    return O.fp2_mul(a, b)


def verify(root, log_k, log_r, Pm, B, cid, segments, queries, cols, sib, combined, coef):
    """True iff the opened columns pass the verifier's Merkle and proximity checks."""
    n = 1 << (log_k + log_r)
    distinct = sorted(set(queries))
    if len(cols) != len(distinct):
        return False
    pairs = []
    for j, col in zip(distinct, cols):
        at, digests = 0, b""
        for s in segments:
            digests += O.column_chain_at(cid, j, col[at:at + s], B)
            at += s
        pairs.append((j, O.hash_domain(O.DOMAIN_INDEX["column-merge"], digests)))
    if not O.merkle_verify(root, n, pairs, sib):
        return False
    for g, (u, m) in enumerate(combined):
        parts = []
        for c in (0, 1):  # the encoding is Fp-linear: c0 and c1 components apart
            parts.append(O.encode_rows(np.array([x[c] for x in u], np.uint64), log_k, log_r,
                                       masks=np.array([x[c] for x in m], np.uint64), mask_coeffs=Pm)[0])
        for j, col in zip(distinct, cols):
            acc = (0, 0)
            for r, sym in enumerate(col):
                t = O.fp2_mul(tuple(int(x) for x in coef[g][r]), (int(sym), 0))
                acc = ((acc[0] + t[0]) % P, (acc[1] + t[1]) % P)
            if acc != (int(parts[0][j]), int(parts[1][j])):
                return False
    return True


def run_case(dev, log_k, log_r, Pm, B, cid, seg_rows, rows, masks, groups, xi, row_vars, queries):
    import tessera_amd as T
    from tessera_amd.codec import decode_round3, round3_from_answers
    flat_rows = [r for s in seg_rows for r in s]
    seg = [len(s) for s in seg_rows]
    params = T.CommitmentParamsPy(log_k, log_r, Pm, B, cid)
    c = T.ReferenceCommitment.commit(dev, params, seg, rows, masks)
    combined = c.open(groups, xi, flat_rows, row_vars)
    combined = [([tuple(int(x) for x in v) for v in u], [tuple(int(x) for x in v) for v in m]) for u, m in combined]
    ans = c.answer(queries)
    data = round3_from_answers(ans)
    assert decode_round3(data) == (len(ans.columns), 0)
    cols, sib = read_round3(data)
    root = c.root()
    c.free()
    coef = O.row_coefficients(row_vars, flat_rows, [[(np.asarray(rp, np.uint64), np.asarray(gm, np.uint64))
                                                      for rp, gm in cl] for _, cl in groups], xi)
    assert verify(root, log_k, log_r, Pm, B, cid, seg, queries, cols, sib, combined, coef)
    # a changed symbol of the answer fails
    bad = [list(col) for col in cols]
    bad[-1][-1] = (bad[-1][-1] + 1) % P
    assert not verify(root, log_k, log_r, Pm, B, cid, seg, queries, bad, sib, combined, coef)
    return data


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_opening_vectors_as_round3(dev, case):
    log_k, log_r, Pm, B = (int(case[x]) for x in ("log2_k", "log2_inv_rate", "mask_coeffs", "row_block"))
    rows = np.array([h(x) for r in case["witness"]["rows"] for x in r], dtype=np.uint64)
    masks = np.array([h(x) for m in case["witness"]["masks"] for x in m], dtype=np.uint64)
    groups = [([(h(a), h(b)) for a, b in g["column_point"]],
               [([(h(a), h(b)) for a, b in cl["row_point"]], (h(cl["gamma"][0]), h(cl["gamma"][1])))
                for cl in g["claims"]]) for g in case["groups"]]
    xi = [(h(a), h(b)) for a, b in case["xi_point"]]
    segments = [[int(r) for r in s] for s in case["segments"]]
    queries = [int(q) for q in case["queries"]]
    data = run_case(dev, log_k, log_r, Pm, B, bytes.fromhex(case["commitment_id"]), segments, rows, masks,
                    groups, xi, int(case["row_vars"]), queries)
    # the bytes hold the vectors' columns and siblings
    cols, sib = read_round3(data)
    assert cols == [[h(x) for x in col] for col in case["columns"]]
    assert [s.hex() for s in sib] == case["siblings"]


def test_random_layout_as_round3(dev):
    rng = np.random.default_rng(11)
    log_k, log_r, Pm, B, row_vars = 9, 1, 40, 8, 6
    segments = [list(range(0, 20)), [], list(range(24, 50)), [60, 61, 63]]
    N = sum(len(s) for s in segments)
    rows = splitmix_uniform(N << log_k, 21)
    masks = splitmix_uniform(N * Pm, 22)

    def f2():
        return tuple(int(x) for x in rng.integers(0, P, 2, dtype=np.uint64))
    groups = [([f2() for _ in range(log_k)], [([f2() for _ in range(row_vars)], f2()) for _ in range(c)])
              for c in (1, 3)]
    xi = [f2() for _ in range(row_vars)]
    queries = [int(q) for q in rng.integers(0, 1 << (log_k + log_r), 30)]
    run_case(dev, log_k, log_r, Pm, B, bytes([3] * 32), segments, rows, masks, groups, xi, row_vars, queries)
