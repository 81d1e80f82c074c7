"""Served proving with continuous batching (the native server of serve.hip) and a Poisson load
generator that reports proofs/s and p50/p99 latency."""
import ctypes
import time

import numpy as np

from .lib import check, lib, u32, u64, vp

MAX_TOKENS = 512


class Request(ctypes.Structure):
    _fields_ = [("id", u64), ("n_prompt", u32), ("prompt", ctypes.POINTER(u32)), ("n_output", u32)]


class Result(ctypes.Structure):
    _fields_ = [("id", u64), ("status", ctypes.c_int), ("slot", u32), ("group", u32),
                ("root", ctypes.c_uint8 * 32),
                ("transcript_state", ctypes.c_uint8 * 32), ("transcript_counter", u64),
                ("final_a", u64 * 2), ("witness_rows", u64), ("sumcheck_vars", u64),
                ("n_tokens", u32), ("tokens", u32 * MAX_TOKENS), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double)]

    def key(self):
        """What must be equal alone and in the mix: tokens, root, transcript state and counter."""
        return (bytes(self.root), bytes(self.transcript_state), int(self.transcript_counter),
                tuple(self.tokens[:self.n_tokens]))


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", u32), ("provers", u32), ("chunk", u32), ("group", u32), ("hbm_budget", u64),
                ("model_seed", u64), ("resident", u32)]


def _declare():
    L = lib()
    L.tsr_server_create.restype = ctypes.c_int
    L.tsr_server_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
    L.tsr_server_submit.restype = ctypes.c_int
    L.tsr_server_submit.argtypes = [vp, ctypes.POINTER(Request)]
    L.tsr_server_wait.restype = ctypes.c_int
    L.tsr_server_wait.argtypes = [vp, ctypes.POINTER(Result), u32]
    L.tsr_server_prove.restype = ctypes.c_int
    L.tsr_server_prove.argtypes = [vp, ctypes.POINTER(Request), ctypes.POINTER(Result)]
    L.tsr_server_cancel.restype = ctypes.c_int
    L.tsr_server_cancel.argtypes = [vp, u64]
    L.tsr_server_stats.restype = ctypes.c_int
    L.tsr_server_stats.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u64)]
    L.tsr_server_open.restype = ctypes.c_int
    L.tsr_server_open.argtypes = [vp, u64, vp, u64, vp, u64, vp, vp, u64, ctypes.POINTER(u64),
                                  ctypes.POINTER(u64)]
    L.tsr_server_release.restype = ctypes.c_int
    L.tsr_server_release.argtypes = [vp, u64]
    L.tsr_server_resident.restype = ctypes.c_int
    L.tsr_server_resident.argtypes = [vp, ctypes.POINTER(u32), ctypes.POINTER(u64)]
    L.tsr_server_footprint.restype = u64
    L.tsr_server_footprint.argtypes = [u32]
    L.tsr_server_clock.restype = ctypes.c_double
    L.tsr_server_clock.argtypes = [vp]
    L.tsr_server_destroy.restype = None
    L.tsr_server_destroy.argtypes = [vp]
    return L


def footprint(tokens):
    return int(_declare().tsr_server_footprint(tokens))


class Server:
    def __init__(self, device=0, slots=64, hbm_budget=0, model_seed=0x7E55E7A, provers=4, chunk=32,
                 group=1, resident=False):
        """group: the most decoded requests a prover commits in one launch set (1: one by one).
        resident: proved requests keep their witness codewords and Merkle tree on the device
        (within hbm_budget) for open() until release()."""
        L = _declare()
        cfg = ServerConfig(slots, provers, chunk, group, hbm_budget, model_seed, 1 if resident else 0)
        h = vp()
        check(L.tsr_server_create(device, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h.value

    @staticmethod
    def _request(rid, prompt, n_output):
        p = (u32 * len(prompt))(*[int(x) for x in prompt])
        return Request(rid, len(prompt), ctypes.cast(p, ctypes.POINTER(u32)), n_output), p

    def submit(self, rid, prompt, n_output):
        req, p = self._request(rid, prompt, n_output)
        check(lib().tsr_server_submit(self.handle, ctypes.byref(req)))

    def cancel(self, rid):
        """True when the request was cancelled (its result arrives with status CANCELLED), False
        when it is being proved or is no longer in flight."""
        from .lib import DECLINED
        rc = lib().tsr_server_cancel(self.handle, rid)
        if rc == DECLINED:
            return False
        check(rc)
        return True

    def wait(self, timeout_ms=600000):
        r = Result()
        check(lib().tsr_server_wait(self.handle, ctypes.byref(r), timeout_ms))
        return r

    def prove(self, rid, prompt, n_output):
        """The plain per-request call, synchronous."""
        req, p = self._request(rid, prompt, n_output)
        r = Result()
        check(lib().tsr_server_prove(self.handle, ctypes.byref(req), ctypes.byref(r)))
        return r

    def open(self, result, queries):
        """Columns of a resident request's witness commitment (tsr_server_open), as
        ReferenceCommitment.answer gives them, for the finished Result `result`."""
        from .pcs import ColumnAnswer
        rows = int(result.witness_rows)
        q = np.ascontiguousarray(np.array(queries, dtype=np.uint64))
        cap_d = max(1, len(set(int(x) for x in q)))
        cap_s = 64 * cap_d + 64
        distinct = np.zeros(cap_d, dtype=np.uint64)
        cols = np.zeros(cap_d * max(rows, 1), dtype=np.uint64)
        sib = ctypes.create_string_buffer(32 * cap_s)
        nd, ns = u64(), u64()
        check(lib().tsr_server_open(self.handle, int(result.id), q.ctypes.data if q.size else None, q.size,
                                    distinct.ctypes.data, cap_d, cols.ctypes.data, sib, cap_s,
                                    ctypes.byref(nd), ctypes.byref(ns)))
        d = nd.value
        raw = sib.raw
        return ColumnAnswer([int(x) for x in distinct[:d]], [cols[i * rows:(i + 1) * rows].copy() for i in range(d)],
                            [raw[32 * i:32 * i + 32] for i in range(ns.value)])

    def release(self, rid):
        """Frees request rid's resident commitment: True, or False when it has none."""
        from .lib import DECLINED
        rc = lib().tsr_server_release(self.handle, rid)
        if rc == DECLINED:
            return False
        check(rc)
        return True

    def resident(self):
        """(resident commitments, device bytes they hold)."""
        c, b = u32(), u64()
        check(lib().tsr_server_resident(self.handle, ctypes.byref(c), ctypes.byref(b)))
        return int(c.value), int(b.value)

    def clock(self):
        return float(lib().tsr_server_clock(self.handle))

    def stats(self):
        m, b = u32(), u64()
        check(lib().tsr_server_stats(self.handle, ctypes.byref(m), ctypes.byref(b)))
        return int(m.value), int(b.value)

    def close(self):
        if self.handle:
            lib().tsr_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def workload(count, seed, rate_hz, min_len=16, max_len=512, vocab=4096):
    """Requests with prompt+output lengths uniform in [min_len, max_len] (the prompt a quarter of
    it, at least one token) and Poisson arrival times at `rate_hz`, from a seeded generator."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(count):
        total = int(rng.integers(min_len, max_len + 1))
        n_prompt = max(1, total // 4)
        prompt = rng.integers(0, vocab, n_prompt).tolist()
        reqs.append((i + 1, prompt, total - n_prompt, float(arrivals[i])))
    return reqs


def _percentiles(lat):
    lat = np.sort(np.asarray(lat))
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


def run_served(server, reqs):
    """Submits each request at its arrival time; returns (results by id, wall seconds)."""
    t0 = server.clock()
    pending = len(reqs)
    results = {}
    wall0 = time.perf_counter()
    for rid, prompt, n_out, at in reqs:
        delay = at - (server.clock() - t0)
        if delay > 0:
            time.sleep(delay)
        server.submit(rid, prompt, n_out)
    while pending:
        r = server.wait()
        results[r.id] = r
        pending -= 1
    wall = time.perf_counter() - wall0
    return results, wall, t0


def run_plain(server, reqs):
    """The plain per-request GPU calls: one at a time, each request starting at the later of its
    arrival and the previous request's end. Returns (results, latencies, wall seconds)."""
    t0 = server.clock()
    results, lat = {}, []
    wall0 = time.perf_counter()
    for rid, prompt, n_out, at in reqs:
        delay = at - (server.clock() - t0)
        if delay > 0:
            time.sleep(delay)
        r = server.prove(rid, prompt, n_out)
        results[rid] = r
        lat.append((r.t_done - t0) - at)
    return results, lat, time.perf_counter() - wall0


def report(results, reqs, wall, t0):
    arrive = {rid: at for rid, _, _, at in reqs}
    lat = [(r.t_done - t0) - arrive[rid] for rid, r in results.items()]
    p50, p99 = _percentiles(lat)
    return {"proofs_per_s": len(results) / wall, "p50_ms": 1e3 * p50, "p99_ms": 1e3 * p99,
            "requests": len(results), "wall_s": wall}
