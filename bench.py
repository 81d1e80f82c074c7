#!/usr/bin/env python3
"""Benchmark: Goldilocks elements committed per second on one MI355X.

Workload (one step): the large single commitment — 256 rows of k = 2^20 uniform Goldilocks
elements (2 GiB, resident in HBM before timing), Reed-Solomon encoded at rate 1/4 to n = 2^22
symbols per row (8 GiB of codewords), BLAKE3 column chains over all 256 rows in blocks of B = 32
(2^22 leaves), the Merkle root, then the root absorbed under `bench/root`, 64 column queries
drawn with challenge_indices("bench/queries", 22, 64) and answered (columns + multiproof).
value = 256 * 2^20 input elements per step / seconds per step.

Also reported: per-kernel times from HIP events recorded on the library's stream over the timed
steps (roofline of the dominant kernel), and the CPU baseline (the oracle's C restatement of the
reference's CPU algorithms, all host threads) on a bounded sample of the same workload.

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LOG_K, LOG_R, ROWS, BLOCK, QUERIES = 20, 2, 256, 32, 64
SEED = 0x7E55E7A
HBM_PEAK_GBS = 8000.0


def kernel_bytes(name, rows, log_k, log_r):
    """Algorithmic HBM bytes of one launch of each kernel of the workload."""
    k, n = 1 << log_k, 1 << (log_k + log_r)
    return {
        "k_tile_pass<gs>": 2 * 8 * rows * n,  # read + write every codeword symbol
        "k_tile_pass<ct>": 2 * 8 * rows * k,  # read + write every row value
        "k_encode_mid": 8 * rows * k + 8 * rows * n,  # read a row tile, write its codeword tile
        "k_chain_leaves": 8 * rows * n + 32 * n,  # read every symbol, write n leaves
        "k_merkle_level": 32 * (2 * n - 2) + 32 * (n - 1),
    }.get(name)


def run_gpu(steps, warmup):
    import tessera_amd as T
    from tessera_amd.lib import lib
    from tessera_amd.transcript import Transcript, session_context
    import ctypes

    dev = T.Device(0)
    k, n = 1 << LOG_K, 1 << (LOG_K + LOG_R)
    src = dev.buffer(ROWS * k).fill_uniform(SEED)
    params = T.CommitmentParamsPy(LOG_K, LOG_R, 0, BLOCK, bytes([7] * 32))
    ctx_tce = session_context()

    def step():
        c = T.ReferenceCommitment.commit_unmasked(dev, params, src)
        root = c.root()
        t = Transcript(ctx_tce)
        t.absorb("bench/root", root)
        q = t.challenge_indices("bench/queries", LOG_K + LOG_R, QUERIES)
        ans = c.answer(q)
        c.free()
        return root, ans

    for _ in range(warmup):
        step()
    dev.synchronize()
    lib().tsr_ctx_profile(dev.handle, 2)
    lib().tsr_ctx_profile(dev.handle, 1)
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        root, ans = step()
    dev.synchronize()
    t1 = time.perf_counter()
    lib().tsr_ctx_profile(dev.handle, 0)
    kernels = {}
    for name in ("k_tile_pass<ct>", "k_encode_mid", "k_tile_pass<gs>", "k_chain_leaves",
                 "k_merkle_level"):
        ms = ctypes.c_double(0)
        cnt = ctypes.c_uint64(0)
        lib().tsr_ctx_kernel_time(dev.handle, name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        if cnt.value:
            kernels[name] = {"total_ms": ms.value, "launches": cnt.value,
                             "avg_ms": ms.value / cnt.value}
    src.free()
    dev.close()
    return (t1 - t0) / steps, kernels, root.hex(), len(ans.siblings)


def traffic_from_profiles(kernel):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel's reading), when a committed summary holds it; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel)
    except (OSError, ValueError):
        return None


def cpu_baseline():
    """The oracle (C, all host threads) on a bounded sample of the same workload, scaled per unit:
    encode 8 of the 256 rows (per-row cost is uniform), chain 2^16 of the 2^22 columns over all 256
    rows (per-column cost is uniform), and a Merkle tree over 2^16 leaves (per-node cost uniform)."""
    import numpy as np
    from oracle import oracle as O
    from tests.helpers import splitmix_uniform

    threads = os.cpu_count() or 1
    k, n = 1 << LOG_K, 1 << (LOG_K + LOG_R)
    enc_rows, cols = 8, 1 << 16
    rows = splitmix_uniform(enc_rows * k, SEED)
    t0 = time.perf_counter()
    cw = O.encode_rows(rows, LOG_K, LOG_R, threads=threads)
    t_enc = (time.perf_counter() - t0) * (ROWS / enc_rows)
    # chains over all 256 rows for `cols` columns: a (256 x cols) block of codeword symbols
    block = np.ascontiguousarray(np.tile(cw[:, :cols], (ROWS // enc_rows, 1)))
    t0 = time.perf_counter()
    leaves = O.column_chains(bytes([7] * 32), block, cols, BLOCK, threads=threads)
    t_chain = (time.perf_counter() - t0) * (n / cols)
    t0 = time.perf_counter()
    O.merkle_tree(leaves, threads=threads)
    t_merkle = (time.perf_counter() - t0) * (n / cols)
    total = t_enc + t_chain + t_merkle
    return {
        "value": ROWS * k / total,
        "unit": "elem/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C restatement, {threads} threads: encode {enc_rows}/256 rows "
                   f"({t_enc:.2f} s scaled), chains over 256 rows for 2^16/2^22 columns "
                   f"({t_chain:.2f} s scaled), Merkle over 2^16 leaves ({t_merkle:.2f} s scaled); "
                   f"est. {total:.1f} s per commit"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X; "
              "multi-GPU sharding is described in DESIGN.md, not run", file=sys.stderr)
        sys.exit(2)

    sec, kernels, root, n_sib = run_gpu(a.steps, a.warmup)
    elems = ROWS * (1 << LOG_K)
    value = elems / sec
    dom = max(kernels, key=lambda x: kernels[x]["total_ms"]) if kernels else None
    roofline = None
    if dom:
        b = kernel_bytes(dom, ROWS, LOG_K, LOG_R)
        achieved = b / (kernels[dom]["avg_ms"] * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": traffic_from_profiles(dom), "algorithmic_bytes": b,
                    "avg_ms": round(kernels[dom]["avg_ms"], 4)}
    cpu = None if a.no_cpu_baseline else cpu_baseline()
    line = {
        "metric": "Goldilocks elems committed/s",
        "value": round(value, 1),
        "unit": "elem/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(sec * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "large single commit: 256 rows x k=2^20 Goldilocks, rate 1/4 -> "
                               "n=2^22, BLAKE3 column chains B=32, Merkle root, 64 openings",
                   "rows": ROWS, "log2_k": LOG_K, "log2_inv_rate": LOG_R, "row_block": BLOCK,
                   "queries": QUERIES, "parallelism": "single GPU"},
        "root": root,
        "kernels_ms": {k: round(v["avg_ms"], 4) for k, v in kernels.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
