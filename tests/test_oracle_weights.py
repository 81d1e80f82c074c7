"""Pins the oracle's weight commitment (RFC-0040) against spec/vectors/weights/weights.json for
the four conformance models and the synthetic one: every matrix's message width, row bits,
committed rows, descriptor, commitment_id and root; the public artifact, R_W and the .twr
envelope byte for byte (and the two .twr files); and the envelope writer against
spec/vectors/codec/weight-commitment.json. The conformance models' weights are read at the
vectors' offsets in their IntegerModel artifacts (INT8 entries; u8 class indices or u16
multipliers)."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, load

WEIGHTS = load("weights.json")
MODELS = WEIGHTS["models"]


def model_inputs(model):
    """(tensors per weight type, channels, n_layers) of a model record."""
    arch = model["arch"]
    if "data" not in model["tensors"][0]:
        blob = open(os.path.join(GOLDEN, "models", os.path.basename(model["model"])), "rb").read()
    per_type = {}
    for t in model["tensors"]:
        if "data" in t:
            w = np.asarray(t["data"], dtype=np.int64)
        else:
            w = np.frombuffer(blob, dtype=np.int8, count=t["rows"] * t["cols"], offset=t["offset"]).astype(np.int64)
        per_type.setdefault(t["matrix"], []).append((t["layer"], w.reshape(t["rows"], t["cols"]), t["rows"], t["cols"]))
    tensors = []
    for name in O.WEIGHT_MATRICES[:9]:
        ws = sorted(per_type[name], key=lambda x: x[0])
        _, _, out, inn = ws[0]
        tensors.append((np.stack([w for _, w, _, _ in ws]), out, inn, len(ws)))
    channels = []
    for c in model["channels"]:
        t = O.WEIGHT_MATRICES.index(c["matrix"])
        dt = np.uint8 if c["kind"] == "class" else np.dtype("<u2")
        vals = np.frombuffer(blob, dtype=dt, count=c["count"], offset=c["offset"]).astype(np.uint64)
        channels.append((t, c["layer"], vals, O._pad2(tensors[t][1])))
    return tensors, channels, arch["n_layers"]


@pytest.mark.parametrize("model", MODELS, ids=[m["name"] for m in MODELS])
def test_weight_commitments(model):
    tensors, channels, n_layers = model_inputs(model)
    mats, public, r_w, twr = O.weight_commit(tensors, channels, n_layers, WEIGHTS["layout"],
                                             WEIGHTS["row_block"], WEIGHTS["epoch"])
    for want, (lay, desc, cid, root) in zip(model["matrices"], mats):
        name = want["matrix"]
        assert lay["log2_k"] == want["log2_k"] and lay["row_vars"] == want["row_vars"], name
        assert lay["rows"] == want["rows"] and lay["committed_rows"] == want["committed_rows"], name
        if name != "ChannelParams":
            assert (lay["a"], lay["h"]) == (want["a"], want["h"]), name
        else:
            assert lay["channel_width"] == want["channel_width"]
        assert desc.hex() == want["descriptor"], name
        assert cid.hex() == want["commitment_id"], name
        assert root.hex() == want["root"], name
    assert public.hex() == model["public"]
    assert r_w.hex() == model["weights_root"]
    assert len(twr) == model["bytes"] and twr[-32:].hex() == model["digest"]
    path = os.path.join(GOLDEN, "models", model["file"])
    if os.path.exists(path):
        assert open(path, "rb").read() == twr


def test_envelopes_of_public_artifacts():
    d = load("codec_weight_commitment.json")
    for case in d["cases"]:
        raw = bytes.fromhex(case["hex"])
        schema = int(case["schema"])
        payload = raw[20:-32]
        assert O.envelope(schema, int(case["version"]), payload) == raw, case["name"]
        if case["type"] == "WeightCommitmentPublic":
            assert O.hash_domain(O.DOMAIN_WEIGHTS_ROOT, payload).hex() == case["weights_root"]


def test_layout_closed_forms():
    """m_T = n_layers ceil(out / 2^a) and k_T = 2^a in_pad, for widths that are not powers of two."""
    for out, inn, layers, log2_kw in [(20, 12, 3, 8), (300, 12, 1, 6), (7, 1000, 5, 4), (1, 1, 1, 0)]:
        lay = O.weight_layout(out, inn, layers, log2_kw)
        a = lay["a"]
        assert lay["rows"] == layers * -(-out // (1 << a))
        assert 1 << lay["log2_k"] == (1 << a) * lay["in_pad"]
        assert all(np.diff(lay["committed_rows"]) > 0)
