"""The oracle's masked advice commitment (ReferenceCommitment::commit: masked encoding, one chain
set per segment, column-merge leaves, Merkle root) against the reference's opening vectors: the
root, the leaves at the queried columns, the answered columns and the canonical multiproof."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import h, hexes, load

CASES = load("pcs_opening.json")["openings"]


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_masked_commitment_matches_vectors(case):
    rows = [h(x) for r in case["witness"]["rows"] for x in r]
    masks = [h(x) for m in case["witness"]["masks"] for x in m]
    seg = [len(s) for s in case["segments"]]
    cid = bytes.fromhex(case["commitment_id"])
    cw, leaves, nodes = O.commit_masked(rows, masks, case["log2_k"], case["log2_inv_rate"],
                                        case["mask_coeffs"], case["row_block"], cid, seg)
    assert nodes[-32:].hex() == case["root"]
    distinct = sorted(set(case["queries"]))
    assert [leaves[j].hex() for j in distinct] == case["intermediate"]["leaves"]
    for j, col in zip(distinct, case["columns"]):
        assert [int(x) for x in cw[:, j]] == [h(x) for x in col]
    n = len(leaves)
    assert O.merkle_open(nodes, n, case["queries"]) == [bytes.fromhex(x) for x in case["siblings"]]


STREAMING = load("pcs_streaming.json")["cases"]


def streaming_inputs(case):
    k = 1 << case["layout"]["log2_k"]
    rows = [h(x) for r in case["rows"] for x in r]
    masks = [h(x) for m in case["masks"] for x in m]
    seg = [len(s) for s in case["segments"]]
    n = len(case["leaves"])
    return k, n.bit_length() - 1 - case["layout"]["log2_k"], rows, masks, seg


@pytest.mark.parametrize("case", STREAMING, ids=[c["name"] for c in STREAMING])
def test_streaming_commitments_match_vectors(case):
    """The reference's streaming advice commitments (canonical layouts of S = 1, 2, 8 segments,
    empty commitments, single rows): codewords, every segment's digests, leaves and root."""
    k, log_r, rows, masks, seg = streaming_inputs(case)
    log_k = case["layout"]["log2_k"]
    cid = bytes.fromhex(case["commitment_id"])
    cw, leaves, nodes = O.commit_masked(rows, masks, log_k, log_r, case["mask_coeffs"],
                                        case["row_block"], cid, seg)
    assert [[int(x) for x in r] for r in cw] == [[h(x) for x in r] for r in case["codewords"]]
    off = 0
    for s, count in enumerate(seg):
        digests = O.column_chains(cid, cw[off:off + count].reshape(-1), len(leaves), case["row_block"])
        assert [d.hex() for d in digests] == case["segment_digests"][s]
        off += count
    assert [x.hex() for x in leaves] == case["leaves"]
    assert nodes[-32:].hex() == case["root"]


def claim_groups(case):
    from tests.helpers import fp2
    return [[(np.array([fp2(x) for x in c["row_point"]], dtype=np.uint64).reshape(-1, 2),
              np.array(fp2(c["gamma"]), dtype=np.uint64)) for c in g["claims"]] for g in case["groups"]]


@pytest.mark.parametrize("case", CASES, ids=[c["name"][:40] for c in CASES])
def test_combined_rows_match_vectors(case):
    """RowLayout::row_coefficients and combine (RFC-0011 steps 2 to 5): the row coefficients
    and the combined rows u, M of every group and of the proximity row."""
    from tests.helpers import fp2
    rows_idx = [r for s in case["segments"] for r in s]
    xi = np.array([fp2(x) for x in case["xi_point"]], dtype=np.uint64).reshape(-1, 2)
    co = O.row_coefficients(case["row_vars"], rows_idx, claim_groups(case), xi)
    want = [[fp2(x) for x in col] for col in case["intermediate"]["row_coefficients"]]
    assert [[tuple(int(v) for v in x) for x in g] for g in co] == want
    k = 1 << case["log2_k"]
    rows = [h(x) for r in case["witness"]["rows"] for x in r]
    masks = [h(x) for m in case["witness"]["masks"] for x in m]
    u = O.combine(co, rows, k)
    mk = O.combine(co, masks, case["mask_coeffs"])
    for g, comb in enumerate(case["combined"]):
        assert [tuple(int(v) for v in x) for x in u[g]] == [fp2(x) for x in comb["u"]]
        assert [tuple(int(v) for v in x) for x in mk[g]] == [fp2(x) for x in comb["mask"]]
