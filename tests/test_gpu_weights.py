"""Parity of the gfx950 weight commitment (RFC-0040) with the reference's weight vectors for the
four conformance models and the synthetic one (every matrix's root, m_T, k_T and row bits; the
public artifact, R_W and the .twr bytes), with the oracle on a larger random model whose widths
are not powers of two (weights given in device memory), and the reference's error order."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import GOLDEN, load

pytestmark = pytest.mark.gpu

WEIGHTS = load("weights.json")
MODELS = WEIGHTS["models"]


@pytest.fixture(scope="module")
def dev():
    import tessera_amd as T
    d = T.Device(0)
    yield d
    d.close()


def raw_inputs(model):
    """(shapes, tensors [(matrix, layer, int8 array)], channels [(matrix, layer, values)])."""
    from tessera_amd.weights import WeightShapes
    blob = None
    if "data" not in model["tensors"][0]:
        blob = open(os.path.join(GOLDEN, "models", os.path.basename(model["model"])), "rb").read()
    tensors, shapes = [], {}
    for t in model["tensors"]:
        if blob is None:
            w = np.asarray(t["data"], dtype=np.int8)
        else:
            w = np.frombuffer(blob, dtype=np.int8, count=t["rows"] * t["cols"], offset=t["offset"])
        tensors.append((t["matrix"], t["layer"], w))
        shapes[t["matrix"]] = (t["rows"], t["cols"])
    channels = []
    for c in model["channels"]:
        dt = np.uint8 if c["kind"] == "class" else np.dtype("<u2")
        channels.append((c["matrix"], c["layer"],
                         np.frombuffer(blob, dtype=dt, count=c["count"], offset=c["offset"]).astype(
                             np.uint8 if c["kind"] == "class" else np.uint16)))
    ws = WeightShapes(model["arch"]["n_layers"], [shapes[m] for m in O.WEIGHT_MATRICES[:9]],
                      [(c[0], c[1]) for c in channels])
    return ws, tensors, channels


@pytest.mark.parametrize("model", MODELS, ids=[m["name"] for m in MODELS])
def test_weight_commitment_vectors(dev, model):
    from tessera_amd.weights import WeightCommitment
    shapes, tensors, channels = raw_inputs(model)
    wc = WeightCommitment(dev, WEIGHTS["layout"], shapes, WEIGHTS["row_block"], WEIGHTS["epoch"],
                          tensors, channels)
    for t, want in enumerate(model["matrices"]):
        root, log2_k, row_vars, rows = wc.matrix(t)
        assert (log2_k, row_vars, rows) == (want["log2_k"], want["row_vars"], want["rows"]), want["matrix"]
        assert root.hex() == want["root"], want["matrix"]
    assert wc.public().hex() == model["public"]
    assert wc.weights_root().hex() == model["weights_root"]
    twr = wc.artifact()
    assert len(twr) == model["bytes"] and twr[-32:].hex() == model["digest"]
    path = os.path.join(GOLDEN, "models", model["file"])
    if os.path.exists(path):
        assert open(path, "rb").read() == twr
    wc.free()


def test_random_model_against_oracle(dev):
    """d_model 96, d_ff 344, vocab 1000, 3 layers, k_W = 2^10, rate 1/4, B = 4; some weights
    with ChannelParams of both kinds; the tensors handed over in device memory."""
    import tessera_amd as T
    from tessera_amd.weights import WeightCommitment, WeightShapes
    rng = np.random.default_rng(40)
    d, ff, vocab, L, hd = 96, 344, 1000, 3, 24
    shape = {"Embed": (d, vocab), "Q": (4 * hd, d), "K": (2 * hd, d), "V": (2 * hd, d), "O": (d, 4 * hd),
             "Gate": (ff, d), "Up": (ff, d), "Down": (d, ff), "LmHead": (vocab, d)}
    layout = {"log2_k": 10, "log2_inv_rate": 2, "mask_coeffs": 64, "blinding_rows": 32}
    tensors, dev_tensors, per_type = [], [], {}
    for t, m in enumerate(O.WEIGHT_MATRICES[:9]):
        layers = L if 1 <= t <= 7 else 1
        ws = rng.integers(-128, 128, (layers,) + shape[m], dtype=np.int64).astype(np.int8)
        per_type[m] = ws
        for l in range(layers):
            buf = T.DeviceBuffer(dev, ws[l].size, np.int8).upload(ws[l].reshape(-1))
            dev_tensors.append(buf)
            tensors.append((m, l, buf))
    chans = [("Q", 0, rng.integers(0, 4, shape["Q"][0]).astype(np.uint8)),
             ("Up", 2, rng.integers(1 << 14, 1 << 16, shape["Up"][0]).astype(np.uint16)),
             ("Down", 1, rng.integers(0, 7, shape["Down"][0]).astype(np.uint8))]
    ws = WeightShapes(L, [shape[m] for m in O.WEIGHT_MATRICES[:9]], [(m, l) for m, l, _ in chans])
    wc = WeightCommitment(dev, layout, ws, 4, 3, tensors, chans)
    otens = [(per_type[m].astype(np.int64), shape[m][0], shape[m][1], per_type[m].shape[0])
             for m in O.WEIGHT_MATRICES[:9]]
    ochans = [(O.WEIGHT_MATRICES.index(m), l, v.astype(np.uint64), O._pad2(shape[m][0])) for m, l, v in chans]
    mats, public, r_w, twr = O.weight_commit(otens, ochans, L, layout, 4, 3)
    for t, (lay, _, _, root) in enumerate(mats):
        got = wc.matrix(t)
        assert got[0] == root and got[3] == lay["rows"] and got[1] == lay["log2_k"], O.WEIGHT_MATRICES[t]
    assert wc.public() == public and wc.weights_root() == r_w and wc.artifact() == twr
    wc.free()


def test_weight_errors_in_order(dev):
    import tessera_amd as T
    from tessera_amd.weights import WeightCommitment, WeightShapes
    model = MODELS[-1]  # synthetic: no ChannelParams
    shapes, tensors, channels = raw_inputs(model)
    lay, rb = WEIGHTS["layout"], WEIGHTS["row_block"]
    with pytest.raises(T.BadShape, match="WeightMissing: Q layer 2"):
        WeightCommitment(dev, lay, shapes, rb, 0, [t for t in tensors if (t[0], t[1]) != ("Q", 2)])
    with pytest.raises(T.BadShape, match="WeightTwice: K layer 0"):
        WeightCommitment(dev, lay, shapes, rb, 0, tensors + [t for t in tensors if (t[0], t[1]) == ("K", 0)])
    with pytest.raises(T.BadShape, match="WeightLength: V layer 1"):
        bad = [(m, l, w[:-1] if (m, l) == ("V", 1) else w) for m, l, w in tensors]
        WeightCommitment(dev, lay, shapes, rb, 0, bad)
    with pytest.raises(T.BadShape, match="WeightPlace: Embed layer 1"):
        WeightCommitment(dev, lay, shapes, rb, 0, tensors + [("Embed", 1, tensors[0][2])])
    with pytest.raises(T.BadShape, match="EmptyWeight: K"):
        empty = WeightShapes(shapes.layers, [s if i != 2 else (0, 12) for i, s in enumerate(shapes.shapes)])
        WeightCommitment(dev, lay, empty, rb, 0, tensors)
    with pytest.raises(T.BadShape, match="WeightMissing: ChannelParams of O layer 0"):
        withcp = WeightShapes(shapes.layers, shapes.shapes, [("O", 0)])
        WeightCommitment(dev, lay, withcp, rb, 0, tensors)


def test_envelope_seal_vectors(dev):
    from tessera_amd.weights import envelope_seal
    for case in load("codec_weight_commitment.json")["cases"]:
        raw = bytes.fromhex(case["hex"])
        assert envelope_seal(dev, int(case["schema"]), int(case["version"]), raw[20:-32]) == raw
