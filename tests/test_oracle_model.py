"""The served-proving model's oracle (oracle/model.c) against an independent vectorized numpy
restatement of the model's definition (a few tokens), the weight generator against SplitMix64
computed with Python integers, and the witness's own relations (rq links between stages, the
greedy choice, the prompt-then-generated token sequence) over a longer decode. The model has no
reference counterpart; these checks pin the oracle to its definition."""
import numpy as np
import pytest

from oracle import oracle as O

P = (1 << 64) - (1 << 32) + 1
L, D, H, DH, FF, V = 4, 256, 4, 64, 1024, 4096
LB = 4 * D * D + 2 * FF * D
OFF_LAYER, OFF_OUT = V * D, V * D + L * LB
WIT_LAYER = 7 * D + FF
SEED = 11


@pytest.fixture(scope="module")
def weights():
    return O.model_weights(SEED)


def test_weight_generator_is_splitmix(weights):
    M = (1 << 64) - 1
    for i in [0, 1, 2, 255, 4096 * 256, OFF_OUT - 1, weights.size - 1]:
        z = (SEED + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        z ^= z >> 31
        assert int(weights[i]) == z % 255 - 127
    assert weights.min() >= -127 and weights.max() <= 127


def restated_decode(W, prompt, total):
    """The model's definition, vectorized over the matrices (int64 arithmetic throughout)."""
    W = W.astype(np.int64)
    emb = W[:OFF_LAYER].reshape(V, D)
    wout = W[OFF_OUT:OFF_OUT + V * D].reshape(V, D)
    layers = []
    for l in range(L):
        b = OFF_LAYER + l * LB
        q, k, v, o = (W[b + i * D * D:b + (i + 1) * D * D].reshape(D, D) for i in range(4))
        w1 = W[b + 4 * D * D:b + 4 * D * D + FF * D].reshape(FF, D)
        w2 = W[b + 4 * D * D + FF * D:b + LB].reshape(D, FF)
        layers.append((q, k, v, o, w1, w2))
    rq = lambda x: np.clip(x >> 2, -127, 127)
    kc = [[] for _ in range(L)]
    vc = [[] for _ in range(L)]
    toks, wits = [], []
    cur = prompt[0]
    for t in range(total):
        toks.append(cur)
        x = emb[cur].copy()
        w = []
        for l, (q, k, v, o, w1, w2) in enumerate(layers):
            a = rq(x)
            qv, kv, vv = (np.clip((m @ a) >> 8, -127, 127) for m in (q, k, v))
            kc[l].append(kv)
            vc[l].append(vv)
            K, Vv = np.array(kc[l]), np.array(vc[l])
            att = np.zeros(D, dtype=np.int64)
            for h in range(H):
                sl = slice(h * DH, (h + 1) * DH)
                s = np.maximum((K[:, sl] @ qv[sl]) >> 6, 0)
                num = s @ Vv[:, sl]
                norm = 1 + s.sum()
                att[sl] = np.clip(np.sign(num) * (np.abs(num) // norm), -127, 127)  # C division
            x = x + o @ att
            m = rq(x)
            hh = np.clip((w1 @ m) >> 8, 0, 127)
            x = x + ((w2 @ hh) >> 8)
            w += [a, qv, kv, vv, att, m, hh, x.copy()]
        f = rq(x)
        nxt = int(np.argmax(wout @ f))
        w += [f, np.array([nxt])]
        wits.append([int(v) % P for v in np.concatenate(w)])
        cur = prompt[t + 1] if t + 1 < len(prompt) else nxt
    return toks, wits


def test_decode_matches_restatement(weights):
    prompt = [7, 4095]
    toks, wit, stride = O.model_decode(weights, prompt, 4)
    rt, rw = restated_decode(weights, prompt, 4)
    assert [int(t) for t in toks] == rt
    assert stride == L * WIT_LAYER + D + 1
    got = wit.reshape(4, stride)
    for t in range(4):
        assert [int(v) for v in got[t]] == [int(v) for v in rw[t]]


def signed(v):
    v = v.astype(object)
    return np.array([int(x) - P if int(x) > P // 2 else int(x) for x in v], dtype=np.int64)


def test_witness_relations(weights):
    prompt = [1, 2, 3]
    total = 24
    toks, wit, stride = O.model_decode(weights, prompt, total)
    w = wit.reshape(total, stride)
    rq = lambda x: np.clip(x >> 2, -127, 127)
    wout = weights[OFF_OUT:OFF_OUT + V * D].astype(np.int64).reshape(V, D)
    for t in range(total):
        row = signed(w[t])
        emb = weights[toks[t] * D:(toks[t] + 1) * D].astype(np.int64)
        assert np.array_equal(row[:D], rq(emb))  # layer 0's a = rq(embedding)
        for l in range(L):
            base = l * WIT_LAYER
            x = row[base + 6 * D + FF:base + 7 * D + FF]
            nxt = row[(l + 1) * WIT_LAYER:(l + 1) * WIT_LAYER + D] if l + 1 < L else row[L * WIT_LAYER:L * WIT_LAYER + D]
            assert np.array_equal(nxt, rq(x))  # the next stage reads rq(x)
            assert row[base + 6 * D:base + 6 * D + FF].min() >= 0  # h = clamp(relu(.), 0, 127)
        f = row[L * WIT_LAYER:L * WIT_LAYER + D]
        assert row[-1] == int(np.argmax(wout @ f))
        want_next = prompt[t + 1] if t + 1 < len(prompt) else row[-1]
        if t + 1 < total:
            assert toks[t + 1] == want_next
