"""The oracle's masked advice commitment (ReferenceCommitment::commit: masked encoding, one chain
set per segment, column-merge leaves, Merkle root) against the reference's opening vectors: the
root, the leaves at the queried columns, the answered columns and the canonical multiproof."""
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
