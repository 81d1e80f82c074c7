"""Shared test helpers: golden fixture loading, hex parsing, seeded inputs, TCE encodings."""
import json
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = 0xFFFFFFFF00000001


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def h(x):
    """A field element written "0x" + 16 hex digits, or a decimal string."""
    if isinstance(x, int):
        return x
    return int(x, 16) if x.startswith("0x") else int(x)


def hexes(xs):
    return np.array([h(x) for x in xs], dtype=np.uint64)


def fp2(pair):
    return (h(pair[0]), h(pair[1]))


def splitmix_uniform(count, seed):
    """SplitMix64 stream (state += 0x9E3779B97F4A7C15, mix of crates/tessera-gpu/tests/commit.rs:19-30)
    with words >= p rejected, so every value is exactly uniform in [0, p)."""
    out = np.empty(count, dtype=np.uint64)
    state = np.uint64(seed)
    have = 0
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        while have < count:
            need = count - have + 16
            steps = np.arange(1, need + 1, dtype=np.uint64)
            z = state + steps * gamma
            state = z[-1]
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            z = z[z < np.uint64(P)]
            take = min(count - have, z.size)
            out[have:have + take] = z[:take]
            have += take
    return out


def ascii_tce(s: str) -> bytes:
    b = s.encode()
    return struct.pack("<Q", len(b)) + b


def context_tce(protocol="tessera-protocol/1", mode="NonInteractive", params="test-small/1",
                statement_digest=bytes(32), session_id=bytes(16)):
    """TCE(SessionContext) (crates/tessera-hash/src/transcript.rs:40-51): Ascii protocol, u16 mode tag,
    Ascii params, 32-byte statement digest, 16-byte session id."""
    tag = {"Live": 0, "NonInteractive": 1}[mode]
    return (ascii_tce(protocol) + struct.pack("<H", tag) + ascii_tce(params) + bytes(statement_digest)
            + bytes(session_id))


def context_from_json(c):
    return context_tce(c["protocol"], c["mode"], c["params"], bytes.fromhex(c["statement_digest"]),
                       bytes.fromhex(c["session_id"]))


def splitmix_raw(count, seed):
    """The raw SplitMix64 stream (no rejection)."""
    gamma = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.arange(1, count + 1, dtype=np.uint64) * gamma
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def zipf_rows(count, rows, s, seed):
    """Row indices drawn Zipf(s) over 0..rows-1 (rank k+1 has weight (k+1)^-s) by inverse CDF on
    u = (splitmix >> 11) * 2^-53."""
    w = np.arange(1, rows + 1, dtype=np.float64) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = (splitmix_raw(count, seed) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.minimum(np.searchsorted(cdf, u, side="right"), rows - 1).astype(np.uint64)


def logup_config3(lookups=1 << 22, rows=1 << 16, seed=0x1064):
    """The LogUp workload of the sumcheck/lookup config: table keys 0..rows-1 with SplitMix64
    values, lookups Zipf(1.1) over the rows, exact multiplicities."""
    keys = np.arange(rows, dtype=np.uint64)
    values = splitmix_uniform(rows, seed)
    idx = zipf_rows(lookups, rows, 1.1, seed + 1)
    return keys, values, idx


def batched_shapes(count=64, seed=0xBA7C, log_k=(10, 20), cols=(16, 512)):
    """The batched-commit workload: `count` matrices, log2 k uniform in [log_k], N rows uniform in
    [cols] (the configs' 'columns'), from a SplitMix64 stream."""
    r = splitmix_raw(2 * count, seed)
    out = []
    for i in range(count):
        lk = log_k[0] + int(r[2 * i] % np.uint64(log_k[1] - log_k[0] + 1))
        n = cols[0] + int(r[2 * i + 1] % np.uint64(cols[1] - cols[0] + 1))
        out.append((lk, n))
    return out
