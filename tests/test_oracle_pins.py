"""Pins the CPU oracle against the reference's golden vectors and against closed forms.

Every oracle function used as a checker elsewhere is tested here first: field arithmetic
(spec/vectors/field), BLAKE3 (the 35 official vectors) and the 19 domains, NTTs and rate-1/2 /
rate-1/4 encodings (spec/vectors/encoding, weights), column chains and Merkle multiproofs
(spec/vectors/hash) and the transcript (spec/vectors/transcript), plus invariants that a dropped
term, a wrong index or a transposed operand would break.
"""
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import P, context_from_json, fp2, h, hexes, load


# ---------------- field ----------------

def test_field_constants_and_roots():
    c = load("constants.json")["goldilocks"]
    assert h(c["modulus"]) == P
    roots = [h(x) for x in c["root_of_unity_2_k"]]
    for k, r in enumerate(roots):
        assert O.fp_root(k) == r
        # order exactly 2^k: r^(2^k) = 1 and r^(2^(k-1)) = -1 for k >= 1
        assert O.fp_pow(r, 1 << k) == 1
        if k:
            assert O.fp_pow(r, 1 << (k - 1)) == P - 1
    # 7 generates Fp^*: 7^((p-1)/q) != 1 for each prime q | p-1
    for q in c["p_minus_1_prime_factors"]:
        assert O.fp_pow(7, (P - 1) // q) != 1


def test_goldilocks_vectors():
    for a, b, s, d, m, inv in load("goldilocks.json")["cases"]:
        a, b = h(a), h(b)
        assert O.fp_add(a, b) == h(s)
        assert O.fp_sub(a, b) == h(d)
        assert O.fp_mul(a, b) == h(m)
        if inv is not None:
            assert O.fp_inv(a) == h(inv)
    # closed forms (goldilocks.rs:157-165)
    assert O.fp_add(P - 1, 1) == 0
    assert O.fp_mul(P - 1, P - 1) == 1
    assert O.fp_sub(1, P - 1) == 2


def test_goldilocks_ext2_vectors():
    for case in load("goldilocks_ext2.json")["cases"]:
        a0, a1, b0, b1, m0, m1, i0, i1, _norm = case
        a, b = (h(a0), h(a1)), (h(b0), h(b1))
        assert O.fp2_mul(a, b) == (h(m0), h(m1))
        if i0 is not None:
            assert O.fp2_inv(a) == (h(i0), h(i1))


def test_field_against_python_bigints():
    rng = random.Random(5)
    for _ in range(2000):
        a, b = rng.randrange(P), rng.randrange(P)
        assert O.fp_mul(a, b) == a * b % P
        assert O.fp_add(a, b) == (a + b) % P
        assert O.fp_sub(a, b) == (a - b) % P


# ---------------- BLAKE3 ----------------

def test_blake3_official_vectors():
    d = load("blake3_test_vectors.json")
    key = d["key"].encode()
    ctx = d["context_string"].encode()
    for case in d["cases"]:
        data = bytes(i % 251 for i in range(case["input_len"]))
        for mode, field, k in ((0, "hash", b""), (1, "keyed_hash", key), (2, "derive_key", ctx)):
            want = bytes.fromhex(case[field])
            assert O.blake3(data, mode, k, len(want)) == want, (case["input_len"], field)


def test_domain_digests():
    d = load("hash.json")
    assert [x["context"] for x in d["domains"]] == O.domain_contexts()
    for i, dom in enumerate(d["domains"]):
        for length, want in zip(d["input_lengths"], dom["digests"]):
            data = bytes(j % 251 for j in range(length))
            assert O.hash_domain(i, data).hex() == want, (dom["name"], length)


def test_prover_streams():
    """The per-proof streams of spec/vectors/hash prover_streams: the three prover domains'
    derive-key digests, the prover share (32 stream bytes), the NI session (16 bytes), 24 advice
    masks of the proof seed, and a seed whose seventh word is >= p (discarded whole)."""
    d = load("hash.json")
    v = d["prover_streams"]
    assert [x["name"] for x in v["domains"]] == list(O.PROVER_DOMAINS)
    for dom in v["domains"]:
        for length, want in zip(d["input_lengths"], dom["digests"]):
            data = bytes(j % 251 for j in range(length))
            assert O.blake3(data, mode=2, key=dom["context"].encode()).hex() == want, (dom["name"], length)
    seed = bytes.fromhex(v["proof_seed"])
    assert O.prover_stream("prover-share", seed, 32).hex() == v["prover_share"]
    assert O.prover_stream("ni-session", seed, 16).hex() == v["ni_session"]
    masks, end = O.advice_masks(seed, len(v["advice_masks"]))
    assert [int(m) for m in masks] == [h(x) for x in v["advice_masks"]]
    rej = v["rejection"]
    rseed = bytes.fromhex(rej["seed"])
    words = np.frombuffer(O.prover_stream("advice-masks", rseed, 8 * len(rej["words"])), dtype="<u8")
    assert [int(w) for w in words] == [h(x) for x in rej["words"]]
    assert int(words[rej["rejected_word_index"]]) >= P
    values, end = O.advice_masks(rseed, len(rej["values"]))
    assert [int(m) for m in values] == [h(x) for x in rej["values"]]
    assert end == len(rej["values"]) + 1  # one word discarded
    # draws in pieces continue the stream across the discarded word
    a, e1 = O.advice_masks(rseed, 5)
    b, e2 = O.advice_masks(rseed, 10, e1)
    assert [int(x) for x in np.concatenate([a, b])] == [int(m) for m in values] and e2 == end


# ---------------- NTT / encoding ----------------

def direct_eval(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % P
    return acc


def test_ntt_vectors():
    for case in load("encoding.json")["ntt"]:
        n = case["n"]
        log_n = n.bit_length() - 1
        coeffs = hexes(case["coefficients"])
        assert list(O.ntt(coeffs, log_n)) == list(hexes(case["values"]))
        assert list(O.ntt(coeffs, log_n, shift=7)) == list(hexes(case["coset_values"]))
        assert list(O.ntt(hexes(case["values"]), log_n, inverse=True)) == list(coeffs)
        assert list(O.ntt(hexes(case["coset_values"]), log_n, inverse=True, shift=7)) == list(coeffs)


@pytest.mark.parametrize("log_n", [0, 1, 5, 12, 13, 14])
def test_ntt_direct_and_roundtrip(log_n):
    rng = np.random.default_rng(log_n)
    n = 1 << log_n
    coeffs = rng.integers(0, P, n, dtype=np.uint64)
    vals = O.ntt(coeffs, log_n)
    w = O.fp_root(log_n)
    for j in rng.integers(0, n, 4):
        assert int(vals[j]) == direct_eval([int(c) for c in coeffs], O.fp_pow(w, int(j)))
    assert np.array_equal(O.ntt(vals, log_n, inverse=True), coeffs)
    cv = O.ntt(coeffs, log_n, shift=7)
    assert np.array_equal(O.ntt(cv, log_n, inverse=True, shift=7), coeffs)


def test_rate_half_encodings():
    for case in load("encoding.json")["encoding"]:
        k, P_ = case["k"], case["mask_coeffs"]
        log_k = k.bit_length() - 1
        cw = O.encode_rows(hexes(case["row"]), log_k, 1, masks=hexes(case["mask"]), mask_coeffs=P_)
        assert list(cw[0]) == list(hexes(case["codeword"])), (k, P_)


def test_rate_quarter_encodings():
    for case in load("weights_encoding.json")["encoding"]:
        k, P_ = case["k"], case["mask_coeffs"]
        log_k = k.bit_length() - 1
        cw = O.encode_rows(hexes(case["row"]), log_k, 2, masks=hexes(case["mask"]), mask_coeffs=P_)
        assert list(cw[0]) == list(hexes(case["codeword"])), (k, P_)


@pytest.mark.parametrize("log_k", [3, 10, 13])
def test_encoding_invariants(log_k):
    """coset_intt(codeword) = coefficients of f (deg < k, zero above): no mask, rate 1/4."""
    rng = np.random.default_rng(log_k)
    k = 1 << log_k
    row = rng.integers(0, P, k, dtype=np.uint64)
    cw = O.encode_rows(row, log_k, 2)[0]
    coeffs = O.ntt(cw, log_k + 2, inverse=True, shift=7)
    assert not coeffs[k:].any()
    assert np.array_equal(O.ntt(coeffs[:k], log_k), row)


# ---------------- column chains / Merkle ----------------

def test_column_chain_vectors():
    for case in load("hash.json")["columns"]:
        cols = case["columns"]
        rows = np.array([[h(x) for x in r] for r in case["rows"]], dtype=np.uint64).reshape(-1)
        got = O.column_chains(bytes.fromhex(case["commitment_id"]), rows, cols, case["block"])
        assert [g.hex() for g in got] == case["leaves"]


def test_merkle_vectors():
    for case in load("hash.json")["merkle"]:
        leaves = [bytes.fromhex(x) for x in case["leaves"]]
        n = len(leaves)
        nodes = O.merkle_tree(leaves)
        root = nodes[-32:]
        assert root.hex() == case["root"]
        for op in case["openings"]:
            sib = O.merkle_open(nodes, n, op["indices"])
            assert [s.hex() for s in sib] == op["siblings"]
            idx = sorted(set(op["indices"]))
            assert O.merkle_verify(root, n, [(i, leaves[i]) for i in idx], sib)
            if sib:
                bad = list(sib)
                bad[0] = bytes(32)
                assert not O.merkle_verify(root, n, [(i, leaves[i]) for i in idx], bad)


def test_merkle_brute_force_small():
    """Every subset of 8 leaves: the proof verifies and has the frontier's sibling count."""
    leaves = [O.hash_domain(0, bytes([i])) for i in range(8)]
    nodes = O.merkle_tree(leaves)
    root = nodes[-32:]
    for mask in range(1, 256):
        idx = [i for i in range(8) if mask >> i & 1]
        sib = O.merkle_open(nodes, 8, idx)
        assert O.merkle_verify(root, 8, [(i, leaves[i]) for i in idx], sib)
        known, count = set(idx), 0
        for _level in range(3):
            count += sum(1 for x in known if x ^ 1 not in known)
            known = {x // 2 for x in known}
        assert len(sib) == count


def test_commit_unmasked_composes():
    """commit = encode, chain each column over all rows in blocks of B, Merkle root."""
    rng = np.random.default_rng(3)
    log_k, log_r, B, rows = 4, 2, 4, 10
    data = rng.integers(0, P, rows << log_k, dtype=np.uint64)
    cid = bytes(range(32))
    cw, nodes = O.commit_unmasked(data, log_k, log_r, B, cid)
    n = 1 << (log_k + log_r)
    assert np.array_equal(cw, O.encode_rows(data, log_k, log_r))
    leaves = O.column_chains(cid, cw.reshape(-1), n, B)
    assert O.merkle_tree(leaves) == nodes


# ---------------- transcript ----------------

def test_transcript_vectors():
    d = load("transcript.json")
    for case in d["cases"]:
        t = O.OracleTranscript(bytes.fromhex(case["context_tce"]))
        from tests.helpers import context_from_json as cfj
        assert cfj(case["context"]) == bytes.fromhex(case["context_tce"])
        assert t.state.hex() == case["state"]
        for s in case["steps"]:
            op, label = s["op"], s["label"]
            if op == "absorb":
                t.absorb(label, bytes.fromhex(s["tce"]))
            elif op == "squeeze":
                assert t.squeeze(label).hex() == s["out"]
            elif op == "challenge_fp":
                v, rej = t.challenge_fp(label)
                assert (str(v), rej) == (s["value"], s["rejected"])
            elif op == "challenge_fp2":
                v, rej = t.challenge_fp2(label)
                assert ([str(v[0]), str(v[1])], rej) == (s["value"], s["rejected"])
            elif op == "challenge_indices":
                v = t.challenge_indices(label, s["log_n"], s["count"])
                assert [str(x) for x in v] == s["value"]
            elif op == "grind":
                assert str(t.grind(label, s["bits"])) == s["nonce"]
            elif op == "check_grind":
                assert t.check_grind(label, s["bits"], int(s["nonce"])) == s["valid"]
            else:
                raise AssertionError(op)
            assert t.state.hex() == s["state"], (case["name"], op, label)
            assert t.counter == s["counter"]
