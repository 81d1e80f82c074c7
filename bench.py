#!/usr/bin/env python3
"""Benchmark: Goldilocks elements committed per second on one MI355X, and served proofs/s with
p50/p99 latency against plain per-request GPU calls.

Headline workload (one step): the large single commitment — 256 rows of k = 2^20 uniform
Goldilocks elements (2 GiB, resident in HBM before timing), Reed-Solomon encoded at rate 1/4 to
n = 2^22 symbols per row (8 GiB of codewords), BLAKE3 column chains over all 256 rows in blocks of
B = 32 (2^22 leaves), the Merkle root, then the root absorbed under `bench/root`, 64 column queries
drawn with challenge_indices("bench/queries", 22, 64) and answered (columns + multiproof).
value = 256 * 2^20 input elements per step / seconds per step.

Also on the same JSON line:
  roofline       the dominant kernel of the headline step (HIP events on the library's stream)
  cpu_baseline   the oracle's C restatement of the reference's CPU algorithms, all host threads,
                 on a bounded sample of the headline workload
  plain          256 x 2^14 commit + 14-variable Sum a*b sumcheck (the small plain drop-in)
  sumcheck       24-variable degree-3 Sum eq(r,x) a b; LogUp of 2^22 Zipf(1.1) lookups into 2^16 rows
  batched        64 matrices (log2 k ~ U[10,20], N ~ U[16,512]) committed by one planned,
                 autotuned, graph-captured batch vs 64 sequential plain calls; roots compared
  served         Poisson arrivals, prompt+output lengths U[16,512], continuous batching vs the
                 plain per-request call at the same arrival times; results compared

python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--only headline]
python bench.py --only refshapes   (the reference's three H100 unit-cost workloads, ~1 minute)
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LOG_K, LOG_R, ROWS, BLOCK, QUERIES = 20, 2, 256, 32, 64
SEED = 0x7E55E7A
HBM_PEAK_GBS = 8000.0
SERVE_RATE_HZ, SERVE_REQUESTS = 400.0, 256
SC_BYTES_PER_INDEX = 80  # the n = 24 eq*a*b sumcheck's algorithmic HBM bytes per index
LOGUP_ROUNDS = 22 * 21 // 2  # level l of the 2^22-leaf lookup tree has l variables
PROFILE_STEPS = 3  # headline steps of the separate, instrumented pass that times each kernel
# BLAKE3 compressions of one column of the headline: the column-init input (40 B, one block) and
# ceil(256 / 32) = 8 column-step inputs of 32 + 32 * 8 = 288 B (5 blocks each)
CHAIN_COMPRESSIONS_PER_COLUMN = 1 + (ROWS // BLOCK) * 5


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def kernel_butterflies(name, rows, log_k, log_r, tile_log=13):
    """Radix-2 butterflies of one launch of each NTT kernel of the encoder's plan: the fused
    middle runs tk CT stages over k and tk + r GS stages over n (tk = tile_log - r), the strided
    passes the remaining log_k - tk CT stages and log_k - tk GS stages."""
    k, n = 1 << log_k, 1 << (log_k + log_r)
    tk = min(log_k, tile_log - log_r)
    return {
        "k_encode_mid": rows * (k // 2 * tk + n // 2 * (tk + log_r)),
        "k_tile_pass<ct>": rows * k // 2 * (log_k - tk),
        "k_tile_pass<gs>": rows * n // 2 * (log_k - tk),
    }.get(name)


def butterfly_peak():
    """The device's measured register-resident butterfly rate (tsr_butterfly_peak)."""
    import tessera_amd as T
    from tessera_amd.lib import check, lib
    dev = T.Device(0)
    rate = ctypes.c_double(0)
    check(lib().tsr_butterfly_peak(dev.handle, ctypes.byref(rate)))
    dev.close()
    return rate.value


def timed(fn, reps):
    """Median seconds of `reps` calls of fn (each call ends synchronized)."""
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def run_headline(steps, warmup):
    import tessera_amd as T
    from tessera_amd.lib import lib
    from tessera_amd.transcript import Transcript, session_context

    dev = T.Device(0)
    k = 1 << LOG_K
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
    # the timed steps run with no instrumentation
    t0 = time.perf_counter()
    for _ in range(steps):
        root, ans = step()
    dev.synchronize()
    t1 = time.perf_counter()
    # per-kernel times in a separate pass: HIP events around every launch on the library's stream
    lib().tsr_ctx_profile(dev.handle, 2)
    lib().tsr_ctx_profile(dev.handle, 1)
    for _ in range(PROFILE_STEPS):
        step()
    dev.synchronize()
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


def run_hostfed(steps=3):
    """The headline commit handed host rows (2 GiB of pageable host memory per step, copied by
    the library): the PCIe-inclusive rate, reported beside the headline, never as its value."""
    import tessera_amd as T
    dev = T.Device(0)
    k = 1 << LOG_K
    src = dev.buffer(ROWS * k).fill_uniform(SEED)
    host = src.download()
    src.free()
    params = T.CommitmentParamsPy(LOG_K, LOG_R, 0, BLOCK, bytes([7] * 32))
    T.ReferenceCommitment.commit_unmasked(dev, params, host).free()  # warm-up
    dev.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        c = T.ReferenceCommitment.commit_unmasked(dev, params, host)
        root = c.root()
        c.free()
    dev.synchronize()
    sec = (time.perf_counter() - t0) / steps
    dev.close()
    print(json.dumps({"workload": "headline commit from host rows (pageable)", "steps": steps,
                      "ms_per_step": round(1e3 * sec, 3), "elems_per_s": round(ROWS * k / sec, 1),
                      "host_bytes_per_step": host.nbytes, "root": root.hex()}), flush=True)


def run_plain(reps=5):
    """256 rows x k = 2^14, rate 1/4, B = 32, the root absorbed under bench/root, then stage/0
    Sum a*b over two uniform tables of 2^14 (degree 2), from a fresh transcript each proof."""
    import tessera_amd as T
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript, session_context
    dev = T.Device(0)
    rows = dev.buffer(256 << 14).fill_uniform(SEED + 1)
    a = dev.buffer(1 << 14).fill_uniform(SEED + 2)
    b = dev.buffer(1 << 14).fill_uniform(SEED + 3)
    params = T.CommitmentParamsPy(14, 2, 0, 32, bytes([7] * 32))
    ctx_tce = session_context()
    out = {}

    def proof():
        c = T.ReferenceCommitment.commit_unmasked(dev, params, rows)
        t = Transcript(ctx_tce)
        t.absorb("bench/root", c.root())
        c.free()
        prove_batch(dev, t, "stage/0", 2, [[Table(a), Table(b)]])
        out["state"] = t.state.hex()

    proof()
    sec = timed(proof, reps)
    dev.close()
    return {"workload": "256 x 2^14 commit (rate 1/4, B 32) + 14-variable Sum a*b sumcheck",
            "ms_per_proof": round(1e3 * sec, 3), "elems_per_s": round((256 << 14) / sec, 1),
            "transcript_state": out["state"]}


def run_sumcheck(reps=3):
    """(i) stage/0 Sum eq(r,x) a(x) b(x), n = 24, degree 3, eq public (given by its point
    r = 24 challenge_fp2("bench/r/{i}")), a and b uniform and resident; (ii) prove_logup of 2^22
    lookups Zipf(1.1) into a 2^16-row table with values, multiplicities counted on the device."""
    import tessera_amd as T
    from tessera_amd.logup import LogUpInstance, multiplicities, prove_logup
    from tessera_amd.sumcheck import Table, prove_batch
    from tessera_amd.transcript import Transcript, session_context
    from tests.helpers import logup_config3
    dev = T.Device(0)
    n = 24
    a = dev.buffer(1 << n).fill_uniform(SEED + 4)
    b = dev.buffer(1 << n).fill_uniform(SEED + 5)
    ctx_tce = session_context()

    def prod3():
        t = Transcript(ctx_tce)
        r = [t.challenge_fp2(f"bench/r/{i}") for i in range(n)]
        # eq(r, .) public, given by its point (the rounds never build its 2^24 entries)
        prove_batch(dev, t, "stage/0", 3, [[Table.eq(r).public(), Table(a), Table(b)]])

    prod3()
    s_prod = timed(prod3, reps)
    a.free(); b.free()

    keys, vals, idx = logup_config3()
    d_keys, d_vals = dev.upload(keys), dev.upload(vals)
    d_lk, d_lv = dev.upload(keys[idx]), dev.upload(vals[idx])
    d_idx = dev.upload(idx.astype(np.uint64))

    d_m = dev.buffer(keys.size)  # 2^16 rows: the counts stay on the device

    def lookups():
        m = multiplicities(dev, keys.size, d_idx, out=d_m)
        t = Transcript(ctx_tce)
        prove_logup(dev, t, [LogUpInstance(d_keys, d_vals, d_lk, d_lv, m)])

    lookups()
    s_logup = timed(lookups, reps)
    dev.close()
    return {"eq_a_b_n24_ms": round(1e3 * s_prod, 3),
            "eq_a_b_ns_per_index": round(1e9 * s_prod / (1 << n), 4),
            "logup_2p22_into_2p16_ms": round(1e3 * s_logup, 3),
            "logup_ns_per_lookup": round(1e9 * s_logup / (1 << 22), 4)}


def run_batched(reps=3):
    """64 matrices (the batched-commit workload, seed 0xBA7C), uniform entries generated on the
    device, rate 1/4, B = 32: one CommitBatch (plan, autotune, graph capture at creation, shape
    groups as parallel graph branches; each run replays the graph) vs 64 sequential
    commit_unmasked calls; every root compared. Also: a second batch of the same shapes created
    from the context's cached tuning, the serial variant (one stream), and each matrix's own
    sequential time by log2 k, which bounds what any batching can gain."""
    import tessera_amd as T
    from tessera_amd.batch import CommitBatch
    from tests.helpers import batched_shapes
    dev = T.Device(0)
    shapes = batched_shapes()
    bufs = [dev.buffer(n << lk).fill_uniform(SEED + 100 + i) for i, (lk, n) in enumerate(shapes)]
    total = sum(n << lk for lk, n in shapes)
    cid = bytes([7] * 32)
    mats = [(buf, lk, 2, 32, cid) for buf, (lk, n) in zip(bufs, shapes)]
    t0 = time.perf_counter()
    first = CommitBatch(dev, mats, autotune=True, graph=True)
    t_create_first = time.perf_counter() - t0
    tuning = first.tuning()
    first.free()
    t0 = time.perf_counter()
    batch = CommitBatch(dev, mats, autotune=True, graph=True)  # the tuning comes from the cache
    t_create = time.perf_counter() - t0
    batch.run()
    dev.synchronize()
    s_batch = timed(lambda: (batch.run(), dev.synchronize()), reps)
    roots_b = batch.roots()
    batch.free()
    serial = CommitBatch(dev, mats, autotune=True, graph=True, serial=True)
    serial.run()
    s_serial = timed(lambda: (serial.run(), dev.synchronize()), reps)
    roots_serial = serial.roots()
    serial.free()

    def one(buf, lk):
        c = T.ReferenceCommitment.commit_unmasked(dev, T.CommitmentParamsPy(lk, 2, 0, 32, cid), buf)
        r = c.root()
        c.free()
        return r

    def sequential():
        return [one(buf, lk) for buf, lk, _, _, _ in mats]

    roots_s = sequential()
    s_seq = timed(sequential, reps)
    # each matrix alone, by log2 k: where the sequential time goes
    by_k = {}
    for (buf, lk, _, _, _), (_, n) in zip(mats, shapes):
        sec = timed(lambda: one(buf, lk), 3)
        e = by_k.setdefault(lk, {"matrices": 0, "elements": 0, "ms": 0.0})
        e["matrices"] += 1
        e["elements"] += n << lk
        e["ms"] += 1e3 * sec
    per_matrix_ms = sum(e["ms"] for e in by_k.values())
    small_ms = sum(e["ms"] for lk, e in by_k.items() if lk <= 15)
    large = {lk: e for lk, e in by_k.items() if lk > 15}
    large_rate = sum(e["elements"] for e in large.values()) / (1e-3 * sum(e["ms"] for e in large.values()))
    dev.close()
    return {"matrices": len(shapes), "elements": total,
            "batched_ms": round(1e3 * s_batch, 3), "sequential_ms": round(1e3 * s_seq, 3),
            "batched_serial_ms": round(1e3 * s_serial, 3),
            "batched_elems_per_s": round(total / s_batch, 1),
            "sequential_elems_per_s": round(total / s_seq, 1),
            "speedup": round(s_seq / s_batch, 3),
            "roots_equal": roots_b == roots_s and roots_serial == roots_s,
            "create_ms_first": round(1e3 * t_create_first, 1), "create_ms": round(1e3 * t_create, 1),
            "plans": [list(t) for t in tuning],
            "by_log2_k": {str(lk): {k: round(v, 3) if isinstance(v, float) else v for k, v in e.items()}
                          for lk, e in sorted(by_k.items())},
            "bound": {"log2_k_le_15_ms": round(small_ms, 3), "per_matrix_total_ms": round(per_matrix_ms, 3),
                      "large_k_elems_per_s": round(large_rate, 1),
                      "max_speedup_if_small_k_were_free": round(per_matrix_ms / (per_matrix_ms - small_ms), 3),
                      "max_speedup_at_large_k_rate": round(per_matrix_ms / (1e3 * total / large_rate), 3)}}


def run_served(cpu_sample):
    """SERVE_REQUESTS requests, Poisson arrivals at SERVE_RATE_HZ, prompt+output lengths
    U[16,512] (prompt a quarter): the server (continuous batching, 128 slots, 4 prover streams,
    chunks of 32 positions, each proved request's witness codewords and Merkle tree kept resident
    on the device until the client releases it) vs plain per-request calls on the same arrivals
    (each starts at the later of its arrival and the previous call's end). Latency = completion -
    arrival. Capacity: the same requests offered all at once. Both with each request's witness
    committed on its own (group 1, the default) and with the decoded requests waiting for a
    prover committed together (group 16)."""
    from tessera_amd.serve import Server, report, run_plain as plain_calls, run_served as served_calls, workload
    reqs = workload(SERVE_REQUESTS, seed=0x5E7E, rate_hz=SERVE_RATE_HZ)
    burst = [(rid + 10_000, p, n, 0.0) for rid, p, n, _ in reqs]
    warm = workload(8, seed=1, rate_hz=1e6)
    runs = {}
    for group in (1, 16):
        with Server(slots=128, provers=4, chunk=32, group=group, resident=True, model_seed=SEED) as srv:
            for phase, wl in (("warm", warm), ("served", reqs), ("capacity", burst)):
                res, wall, t0 = served_calls(srv, wl)
                rep = report(res, wl, wall, t0)
                rep["mean_group"] = float(np.mean([r.group for r in res.values()]))
                rep["max_in_flight"] = srv.stats()[0]
                rep["resident_gib"] = round(srv.resident()[1] / 2**30, 2)
                for rid in res:
                    srv.release(rid)
                runs[(group, phase)] = (rep, res)
            if group == 1:
                res_p, lat_p, wall_p = plain_calls(srv, reqs)
    p50, p99 = np.percentile(lat_p, 50), np.percentile(lat_p, 99)
    service = [r.t_done - r.t_admit for r in res_p.values()]
    plain = {"proofs_per_s": len(res_p) / wall_p, "p50_ms": 1e3 * p50, "p99_ms": 1e3 * p99,
             "service_p50_ms": 1e3 * float(np.median(service)), "wall_s": wall_p}
    equal = True
    for group in (1, 16):
        res_s, res_b = runs[(group, "served")][1], runs[(group, "capacity")][1]
        equal = equal and all(res_s[i].status == 0 and res_s[i].key() == res_p[i].key() for i in res_p)
        equal = equal and all(res_b[i].status == 0 and res_b[i].key() == runs[(1, "capacity")][1][i].key()
                              for i in res_b)
    rnd = lambda d: {k: round(v, 3) if isinstance(v, float) else v for k, v in d.items()}
    served, capacity = runs[(1, "served")][0], runs[(1, "capacity")][0]
    out = {"rate_hz": SERVE_RATE_HZ, "requests": len(reqs),
           "served": rnd(served),
           "plain": rnd(plain),
           "capacity": rnd(capacity),
           "served_commits_grouped": rnd(runs[(16, "served")][0]),
           "capacity_commits_grouped": rnd(runs[(16, "capacity")][0]),
           "speedup_proofs_per_s": round(served["proofs_per_s"] / plain["proofs_per_s"], 3),
           "capacity_vs_plain": round(capacity["proofs_per_s"] / plain["proofs_per_s"], 3),
           "results_equal": bool(equal)}
    if cpu_sample:
        out["cpu"] = cpu_served_sample(reqs[:3])
    return out


def cpu_served_sample(reqs):
    """The oracle (C: decode, commit, sumcheck; all host threads) on the first requests of the
    same workload: proofs/s for one request at a time."""
    from oracle import oracle as O
    threads = host_threads()
    w = O.model_weights(SEED)
    t0 = time.perf_counter()
    tokens = 0
    for rid, prompt, n_out, _ in reqs:
        O.served_proof(w, rid, prompt, n_out, threads=threads)
        tokens += len(prompt) + n_out
    sec = time.perf_counter() - t0
    return {"proofs_per_s": round(len(reqs) / sec, 4), "cores": threads, "kind": "port",
            "sample": f"{len(reqs)} requests, {tokens} tokens, {sec:.1f} s"}


def copy_probe(nbytes=2 << 30, reps=4):
    """Device-to-device copies of a known size: the reference reading of FETCH_SIZE and WRITE_SIZE
    (each copy fetches and writes `nbytes`)."""
    import tessera_amd as T
    from tessera_amd.lib import check, lib
    dev = T.Device(0)
    a = dev.buffer(nbytes // 8).fill_uniform(SEED)
    b = dev.buffer(nbytes // 8)
    t0 = time.perf_counter()
    for _ in range(reps):
        check(lib().tsr_copy(dev.handle, b.ptr, a.ptr, nbytes))
    dev.synchronize()
    sec = (time.perf_counter() - t0) / reps
    print(json.dumps({"copy_bytes": nbytes, "copies": reps, "ms_per_copy": round(1e3 * sec, 3),
                      "GBps_read_plus_write": round(2 * nbytes / sec / 1e9, 1)}), flush=True)
    dev.close()


def traffic_tables(fetch_head, fetch_copy, write_head, write_copy, out_json, out_md, command):
    """HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE CSVs (KiB per dispatch):
    each counter is scaled by what it reads for the calibration copy (bench.py --only copy: 2 GiB
    fetched and written per copy); writes the per-kernel table to out_json (read by the roofline)
    and a markdown summary to out_md."""
    import csv
    import re

    def per_kernel(path):
        agg = {}
        for r in csv.DictReader(open(path)):
            name = re.sub(r"\(.*", "", r["Kernel_Name"]).replace("void ", "").replace("tsr::", "")
            # the bench's launch names: k_tile_pass<true|false>, tile sizes dropped
            name = re.sub(r"<(true|false)[^>]*>", r"<\1>", re.sub(r"<\d[^>]*>", "", name))
            # the fused middle over two rows per workgroup is timed under the fused middle's name
            name = name.replace("k_encode_mid_rows", "k_encode_mid")
            agg.setdefault(name, []).append(float(r["Counter_Value"]) * 1024.0)
        return agg

    def copy_factor(path):
        vals = per_kernel(path).get("__amd_rocclr_copyBuffer", [])
        big = max(vals)
        return (2 << 30) / big, big

    f_fac, f_raw = copy_factor(fetch_copy)
    w_fac, w_raw = copy_factor(write_copy)
    fetch, write = per_kernel(fetch_head), per_kernel(write_head)
    table = {}
    lines = [f"# HBM traffic per launch, headline commit (256 x 2^20, rate 1/4)\n",
             "Command (each pass its own run, from the repository root on the GPU box):\n",
             "```\n" + command + "\n```\n",
             f"Calibration: one 2 GiB device copy reads FETCH_SIZE = {f_raw / 2**30:.4f} GiB and "
             f"WRITE_SIZE = {w_raw / 2**30:.4f} GiB, so FETCH_SIZE is scaled by {f_fac:.4f} and "
             f"WRITE_SIZE by {w_fac:.4f}.\n",
             "| kernel | launches | fetch (GiB/launch) | write (GiB/launch) | traffic (GiB/launch) |",
             "|---|---|---|---|---|"]
    for name in fetch:
        fv = sum(fetch[name]) / len(fetch[name]) * f_fac
        wv = sum(write.get(name, [0])) / max(1, len(write.get(name, [0]))) * w_fac
        if name.startswith("k_"):
            # every launch of the headline commit covers all its rows; the run makes one launch
            # per timed step and per instrumented step, averaged above
            table[name] = {"rows": ROWS, "bytes": round(fv + wv)}
        lines.append(f"| `{name}` | {len(fetch[name])} | {fv / 2**30:.4f} | {wv / 2**30:.4f} | "
                     f"{(fv + wv) / 2**30:.4f} |")
    with open(out_json, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    with open(out_md, "w") as f:
        f.write("\n".join(lines) + "\n")


def traffic_from_profiles(kernel, rows):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE + WRITE_SIZE, scaled by a
    copy kernel's reading) for launches of `rows` rows, when a committed summary holds it; else
    None. The summary maps kernel -> {"rows": rows per launch, "bytes": bytes per launch}."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    alias = {"k_tile_pass<gs>": "k_tile_pass<true>", "k_tile_pass<ct>": "k_tile_pass<false>"}
    try:
        with open(path) as f:
            entry = json.load(f).get(alias.get(kernel, kernel))
    except (OSError, ValueError):
        return None
    if not isinstance(entry, dict) or entry.get("rows") != rows:
        return None
    return entry["bytes"]


def trace_summary(db, title, command):
    """Markdown of a rocprofv3 --kernel-trace --stats database: per kernel calls, total, average,
    min and max duration, share, registers, LDS and launch shape."""
    import re
    import sqlite3
    c = sqlite3.connect(db)
    rows = list(c.execute("select * from top_kernels"))
    extra = {r[0]: r[1:] for r in c.execute(
        "select name, min(duration), max(duration), max(vgpr_count), max(sgpr_count), max(lds_size), "
        "max(grid_x*grid_y*grid_z), max(workgroup_x) from kernels group by name")}
    durs = {}
    for name, d in c.execute("select name, duration from kernels"):
        durs.setdefault(name, []).append(d)
    out = [f"# {title}\n", "Command (from the repository root on the GPU box):\n",
           "```\n" + command + "\n```\n",
           "The median leaves out the slower first launches of a run (cold caches and clocks).\n",
           "| kernel | calls | total (us) | avg (us) | median (us) | min (us) | max (us) | % | VGPR | SGPR | LDS (B) | grid | WG |",
           "|---|---|---|---|---|---|---|---|---|---|---|---|---|"]
    for name, calls, tot, avg, pct in rows:
        short = re.sub(r"\(.*", "", name)
        mn, mx, vg, sg, lds, grid, wg = extra.get(name, (0,) * 7)
        med = float(np.median(durs.get(name, [0]))) / 1000
        out.append(f"| `{short}` | {calls} | {tot:.1f} | {avg:.1f} | {med:.1f} | {mn / 1000:.1f} | {mx / 1000:.1f} | "
                   f"{pct:.1f} | {vg} | {sg} | {lds} | {grid} | {wg} |")
    return "\n".join(out) + "\n"


def rfc0020(fn, units, warmups=3, runs=10):
    """RFC-0020's measurement protocol (crates/tessera-bench/src/measure.rs:16-33): warm-up runs,
    measured runs, ns per unit as median, min, max and the coefficient of variation (sample
    standard deviation over the mean); a run whose CV is above 3% is repeated once."""
    def once():
        for _ in range(warmups):
            fn()
        ts = []
        for _ in range(runs):
            t0 = time.perf_counter()
            fn()
            ts.append(1e9 * (time.perf_counter() - t0) / units)
        ts = np.array(ts)
        cv = float(ts.std(ddof=1) / ts.mean()) if runs > 1 and ts.mean() > 0 else 0.0
        return {"median": round(float(np.median(ts)), 4), "min": round(float(ts.min()), 4),
                "max": round(float(ts.max()), 4), "cv": round(cv, 4), "runs": runs, "warmups": warmups}
    st = once()
    if st["cv"] > 0.03:
        st = once()
        st["repeated"] = True
    st["noisy"] = st["cv"] > 0.03
    return st


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# The reference's single-thread CPU microbenchmarks (docs/perf/baselines/github-ubuntu-24.04.json,
# GitHub ubuntu-24.04 x86_64 runner; sizes from crates/tessera-bench/benches/primitives.rs:106-113).
REF_CPU_NS = {"field/fp-mul": 1.5947, "ntt/fp": 29.4802, "hash/column-chains": 27.9673,
              "sumcheck/product-3": 66.5614, "logup/prove": 207.9136, "trees/forest": 53.1683}


def cpu_calibration():
    """The oracle single-threaded at the reference's microbenchmark sizes, beside the reference's
    published numbers, so that the CPU baseline can be judged against the reference's own CPU
    path: 2^20 Fp products, a forward NTT of 2^20 points, one block of 32 rows into 2^14 column
    chains, a degree-3 sumcheck of three Fp tables of 2^20, the LogUp stage of 2^18 lookups into
    2^16 rows with values, a product forest of 2^10 trees of 2^12 leaves. Runs above 0.2 s use
    one warm-up and three runs."""
    from oracle import oracle as O
    from tests.helpers import context_tce, splitmix_uniform
    out = {}
    ctx = context_tce(statement_digest=bytes([0x65] * 32), session_id=bytes([0x65] * 16))

    def add(name, fn, units):
        t0 = time.perf_counter()
        fn()
        long = time.perf_counter() - t0 > 0.2
        st = rfc0020(fn, units, *((1, 3) if long else (3, 10)))
        st["reference_ns"] = REF_CPU_NS[name]
        st["ratio_to_reference"] = round(st["median"] / REF_CPU_NS[name], 3)
        out[name] = st

    n = 1 << 20
    a, b, c = splitmix_uniform(n, 1), splitmix_uniform(n, 2), splitmix_uniform(n, 3)
    add("field/fp-mul", lambda: O.fp_mul_vec(a, b), n)
    add("ntt/fp", lambda: O.ntt(a.copy(), 20), n)
    rows = splitmix_uniform(32 << 14, 4)
    add("hash/column-chains", lambda: O.column_chains(bytes([0x65] * 32), rows, 1 << 14, 32), 32 << 14)
    add("sumcheck/product-3", lambda: O.prove_batch(O.OracleTranscript(ctx), "stage/0", 3,
                                                    [[(a, 0, 0), (b, 0, 0), (c, 0, 0)]], threads=1), n)
    R, N = 1 << 16, 1 << 18
    keys = np.arange(R, dtype=np.uint64)
    vals = splitmix_uniform(R, 5)
    picks = (splitmix_uniform(N, 6) % np.uint64(R)).astype(np.uint64)
    inst = {"keys": keys, "values": vals, "lookup_keys": keys[picks], "lookup_values": vals[picks],
            "multiplicities": O.multiplicities(R, picks)}
    add("logup/prove", lambda: O.prove_logup(O.OracleTranscript(ctx), [inst], [0, 1]), N)
    leaves = splitmix_uniform(1 << 22, 7)
    add("trees/forest", lambda: O.prove_trees(O.OracleTranscript(ctx), [(leaves, 12)]), 1 << 22)
    return out


def cpu_baseline():
    """The oracle (C, all host threads) on a bounded sample of the same workload, scaled per unit,
    measured with RFC-0020's protocol: encode 32 of the 256 rows (per-row cost is uniform), chain
    2^17 of the 2^22 columns over all 256 rows (per-column cost is uniform), and a Merkle tree over
    2^17 leaves (per-node cost uniform). value = input elements per second of the whole commit at
    the sample's median rate. Also the single-thread calibration against the reference's own
    CPU microbenchmarks."""
    from oracle import oracle as O
    from tests.helpers import splitmix_uniform

    threads = host_threads()
    k, n = 1 << LOG_K, 1 << (LOG_K + LOG_R)
    enc_rows, cols = 32, 1 << 17
    rows = splitmix_uniform(enc_rows * k, SEED)
    cw = O.encode_rows(rows, LOG_K, LOG_R, threads=threads)
    # chains over all 256 rows for `cols` columns: a (256 x cols) block of codeword symbols
    block = np.ascontiguousarray(np.tile(cw[:, :cols], (ROWS // enc_rows, 1)))
    leaves = O.column_chains(bytes([7] * 32), block, cols, BLOCK, threads=threads)
    parts = {
        "encode": (lambda: O.encode_rows(rows, LOG_K, LOG_R, threads=threads), ROWS / enc_rows),
        "chains": (lambda: O.column_chains(bytes([7] * 32), block, cols, BLOCK, threads=threads), n / cols),
        "merkle": (lambda: O.merkle_tree(leaves, threads=threads), n / cols),
    }
    stats, total = {}, 0.0
    for name, (fn, scale) in parts.items():
        st = rfc0020(fn, 1)  # ns per run of the sample
        stats[name] = {"sample_ms": round(st["median"] / 1e6, 3), "cv": st["cv"], "noisy": st["noisy"],
                       "scaled_s": round(st["median"] * 1e-9 * scale, 3)}
        total += st["median"] * 1e-9 * scale
    return {
        "value": round(ROWS * k / total, 1),
        "unit": "elem/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": cpu_model(),
        "protocol": "RFC-0020: 3 warm-ups + 10 measured runs per part, median (CV > 3% repeated once)",
        "sample": (f"oracle C restatement, {threads} threads: encode {enc_rows}/256 rows, chains over "
                   f"256 rows for 2^17/2^22 columns, Merkle over 2^17 leaves, each scaled per unit; "
                   f"est. {total:.2f} s per commit"),
        "parts": stats,
        "single_thread_calibration": cpu_calibration(),
    }


# The reference's published H100 unit costs for the same three workloads (1x H100 SXM 80 GB,
# docs/perf/v0.1/cc-h100.md:9,13, pi-h100.md:9-11,19, l-h100.md:18-21).
H100_PS = {"encode_and_chains": 134.8, "forest": 84.6, "logup": 297.0}


def run_refshapes(reps=3):
    """The reference's three unit-cost workloads on this back end, per unit as the reference
    quotes them: (i) encode at rate 1/2 with 200 mask coefficients + BLAKE3 column chains of
    B = 32 rows, rows of k = 2^22 (8 row blocks, 2^30 committed elements; the rows are drawn on
    the device each block and that time is subtracted); (ii) a product forest of 2^16 trees of
    2^14 leaves (2^30 leaves): build + every level sumcheck; (iii) LogUp of 2^30 lookups into a
    2^24-row table with values: both fraction trees + every level sumcheck (multiplicities
    counted beforehand). Best of `reps` after a warm-up."""
    import tessera_amd as T
    from tessera_amd.lib import lib
    from tessera_amd.logup import Forest, LogUpInstance, multiplicities, prove_logup, prove_trees
    from tessera_amd.pcs import ColumnChains, RowEncoder
    from tessera_amd.transcript import Transcript
    dev = T.Device(0)

    def best(fn):
        fn()
        dev.synchronize()
        out = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            dev.synchronize()
            out.append(time.perf_counter() - t0)
        return min(out)

    res = {}
    log_k, masks_n, B, blocks = 22, 200, 32, 8
    k, n = 1 << log_k, 1 << (log_k + 1)
    rows, masks, cw = dev.buffer(B * k), dev.buffer(B * masks_n), dev.buffer(B * n)
    enc = RowEncoder(dev, k, 1, masks_n)

    def cc(work=True):
        ch = ColumnChains(dev, bytes([7] * 32), n, B) if work else None
        for i in range(blocks):
            rows.fill_uniform(SEED + 1000 + i)
            masks.fill_uniform(SEED + 2000 + i)
            if work:
                enc.encode_rows(rows, masks, out=cw)
                ch.absorb_block(cw)
        if ch:  # the chain digests stay on the device
            dev.synchronize()
            lib().tsr_chains_free(ch.handle)
            ch.handle = None
    t_all, t_gen = best(cc), best(lambda: cc(False))
    elems = blocks * B * k
    res["encode_and_chains"] = {"ps_per_elem": round(1e12 * (t_all - t_gen) / elems, 2),
                                "elements": elems}
    rows.free(); masks.free(); cw.free()

    leaves = dev.buffer(1 << 30).fill_uniform(SEED + 77)
    t = best(lambda: prove_trees(dev, Transcript(b"forest"), [Forest(leaves, 14)]))
    res["forest"] = {"ps_per_leaf": round(1e12 * t / (1 << 30), 2), "trees": 1 << 16, "leaves_per_tree": 1 << 14}
    leaves.free()

    L, R = 30, 24
    rng = np.random.default_rng(SEED)
    vals = rng.integers(0, 1 << 62, 1 << R, dtype=np.uint64)
    d_keys, d_vals = dev.upload(np.arange(1 << R, dtype=np.uint64)), dev.upload(vals)
    idx = rng.integers(0, 1 << R, 1 << L, dtype=np.uint64)
    d_idx = dev.upload(idx)  # the keys are the row indices
    d_lv = dev.upload(vals[idx])
    del idx
    m = dev.upload(multiplicities(dev, 1 << R, d_idx))
    inst = [LogUpInstance(d_keys, d_vals, d_idx, d_lv, m)]
    t = best(lambda: prove_logup(dev, Transcript(b"logup"), inst))
    res["logup"] = {"ps_per_lookup": round(1e12 * t / (1 << L), 2), "lookups": 1 << L, "rows": 1 << R}
    dev.close()
    for key, unit in (("encode_and_chains", "ps_per_elem"), ("forest", "ps_per_leaf"), ("logup", "ps_per_lookup")):
        res[key]["h100_ps"] = H100_PS[key]
        res[key]["speedup_vs_h100"] = round(H100_PS[key] / res[key][unit], 3)
    return res


def headline_roofline(dom, kernels, rows_profiled):
    """The dominant kernel's roofline. The encoder's kernels are integer-VALU bound (64-bit
    modular multiplies from 32-bit multiply-adds), so the binding roof is the butterfly rate
    against the rate the same butterfly reaches register-resident on this device
    (tsr_butterfly_peak); the HBM fraction (algorithmic bytes and measured traffic) is beside it."""
    launches = kernels[dom]["launches"]
    rows_per_launch = rows_profiled // launches  # the encoder may split rows over launches
    avg_s = kernels[dom]["avg_ms"] * 1e-3
    b = kernel_bytes(dom, rows_per_launch, LOG_K, LOG_R)
    gbs = b / avg_s / 1e9
    hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": b,
           "traffic": traffic_from_profiles(dom, rows_per_launch)}
    bf = kernel_butterflies(dom, rows_per_launch, LOG_K, LOG_R)
    out = {"kernel": dom, "rows_per_launch": rows_per_launch, "avg_ms": round(kernels[dom]["avg_ms"], 4)}
    if bf:
        peak = butterfly_peak()
        rate = bf / avg_s
        out.update({"bound": "valu", "achieved": round(rate / 1e9, 1), "peak": round(peak / 1e9, 1),
                    "unit": "G butterflies/s", "frac": round(rate / peak, 4),
                    "peak_kind": "measured: the same Goldilocks butterfly register-resident "
                                 "(tsr_butterfly_peak), ~42 VALU instructions per butterfly",
                    "butterflies": bf, "traffic": hbm["traffic"], "hbm": hbm})
    else:
        out.update({"bound": "hbm", **{k: hbm[k] for k in ("achieved", "peak", "unit", "frac", "traffic")},
                    "algorithmic_bytes": b})
    return out


def chain_roofline(kernels):
    """k_chain_leaves of the headline: BLAKE3 compressions per second and its HBM fraction (it
    reads every codeword symbol once and writes one leaf per column)."""
    c = kernels.get("k_chain_leaves")
    if not c:
        return None
    n = 1 << (LOG_K + LOG_R)
    launches_per_step = max(1, c["launches"] // PROFILE_STEPS)
    per_step_s = c["avg_ms"] * 1e-3 * launches_per_step
    comp = CHAIN_COMPRESSIONS_PER_COLUMN * n
    b = kernel_bytes("k_chain_leaves", ROWS, LOG_K, LOG_R)
    gbs = b / per_step_s / 1e9
    return {"kernel": "k_chain_leaves", "bound": "valu", "ms_per_step": round(per_step_s * 1e3, 4),
            "compressions": comp, "achieved": round(comp / per_step_s / 1e9, 2),
            "unit": "G BLAKE3 compressions/s",
            "hbm": {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": b,
                    "traffic": traffic_from_profiles("k_chain_leaves", ROWS)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", nargs=4, metavar=("FETCH_HEAD", "FETCH_COPY", "WRITE_HEAD", "WRITE_COPY"),
                    help="rocprofv3 --pmc CSVs -> profiles/traffic.json and its summary")
    ap.add_argument("--traffic-command", default="", help="the profiling command, for the summary")
    ap.add_argument("--trace-summary", nargs=3, metavar=("DB", "TITLE", "COMMAND"),
                    help="print a markdown summary of a rocprofv3 kernel-trace database")
    ap.add_argument("--only", choices=["headline", "copy", "hostfed", "refshapes", "plain", "sumcheck", "batched",
                                       "served"],
                    default=None,
                    help="time the headline commit only; run the device copy that calibrates "
                         "the HBM byte counters (4 copies of 2 GiB); time the headline commit "
                         "from host rows (the PCIe-inclusive rate); or time the reference's three "
                         "H100 unit-cost workloads; or one of the other sections alone")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU (a single MI355X); running the path sharded over "
              "several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)

    if a.trace_summary:
        print(trace_summary(*a.trace_summary), end="")
        return
    if a.traffic:
        traffic_tables(*a.traffic, os.path.join(ROOT, "profiles", "traffic.json"),
                       os.path.join(ROOT, "profiles", "traffic_commit_large_256x2p20.md"),
                       a.traffic_command)
        return
    if a.only == "copy":
        copy_probe()
        return
    if a.only == "hostfed":
        run_hostfed()
        return
    if a.only == "batched":
        print(json.dumps({"workload": "batched commit", **run_batched()}), flush=True)
        return
    if a.only == "plain":
        print(json.dumps(run_plain()), flush=True)
        return
    if a.only == "sumcheck":
        print(json.dumps({"workload": "sumcheck/lookup configuration", **run_sumcheck()}), flush=True)
        return
    if a.only == "served":
        print(json.dumps({"workload": "served proving", **run_served(cpu_sample=False)}), flush=True)
        return
    if a.only == "refshapes":
        print(json.dumps({"workload": "the reference's H100 unit-cost workloads", **run_refshapes()}), flush=True)
        return
    log("headline: 256 x 2^20 commit + 64 openings")
    sec, kernels, root, n_sib = run_headline(a.steps, a.warmup)
    elems = ROWS * (1 << LOG_K)
    value = elems / sec
    dom = max(kernels, key=lambda x: kernels[x]["total_ms"]) if kernels else None
    roofline = None
    if dom:
        roofline = headline_roofline(dom, kernels, ROWS * PROFILE_STEPS)
    extra = {}
    if a.only is None:
        log("plain: 256 x 2^14 commit + Sum a*b")
        extra["plain"] = run_plain()
        log("sumcheck: n=24 eq*a*b, LogUp 2^22 lookups")
        extra["sumcheck"] = run_sumcheck()
        log("batched: 64 matrices")
        extra["batched"] = run_batched()
        log("served: Poisson arrivals")
        extra["served"] = run_served(cpu_sample=not a.no_cpu_baseline)
    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline()
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
    line.update(extra)
    line["rooflines"] = {"k_chain_leaves": chain_roofline(kernels)}
    if "sumcheck" in extra:
        sc = extra["sumcheck"]
        gbs = SC_BYTES_PER_INDEX * (1 << 24) / (sc["eq_a_b_n24_ms"] * 1e-3) / 1e9
        line["rooflines"]["sumcheck_eq_a_b_n24"] = {
            "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": SC_BYTES_PER_INDEX * (1 << 24),
            "note": f"{SC_BYTES_PER_INDEX} B per hypercube index: two Fp tables read in round 0, "
                    "then fused bind + sums over the folded Fp2 halves, eq(r, x) as a point"}
        line["rooflines"]["logup_2p22_into_2p16"] = {
            "bound": "latency", "rounds": LOGUP_ROUNDS,
            "us_per_round": round(1e3 * sc["logup_2p22_into_2p16_ms"] / LOGUP_ROUNDS, 2),
            "note": "22 levels of the lookup tree, a level of l variables proven in l rounds"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
