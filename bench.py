#!/usr/bin/env python3
"""Benchmark: batched ragged semi-local LCS (GCUPS) on one MI355X.

A "step" is one pass of the hot path over one batch: all 4,096 ragged pairs
(log-normal lengths, median 2,000, clipped to [100, 50,000], alphabet 20, half
of the pairs mutated copies) combed into their bit-exact kernel permutations
in ONE persistent launch.  Inputs are resident in HBM before the timed region.
The same pairs issued one launch per pair are timed beside it, and the
reference's OpenMP CPU path is timed on a bounded sample on the host cores.

Prints one JSON line (rank 0).  --gpus N with N > 1 exits: this engine runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "GCUPS batched vs one-pair-per-launch, 1 MI355X, bit-exact; serving p50/p99"
VALU_OPS_PER_CELL = 5          # 2 compares + 1 or + 2 selects: the algorithmic work of one cell
BITPAR_OPS_PER_BIT_CELL = 0.93 # bit-parallel combing: 127 steps x 15 64-bit ops per 64 x 64 word-cell, as int32
HYYRO_OPS_PER_WORD_ROW = 8     # Hyyro's row: 64-bit AND, ADD, XOR, OR = 8 int32 lane-ops per 64 cells
LANES_PER_CU_CLK = 64          # int32 VALU lane-ops per CU per clock on CDNA4
PEAK_CLOCK_GHZ = 2.4           # MI355X peak engine clock
CUS = 256


def peak_tops(cus: int = CUS) -> float:
    return cus * LANES_PER_CU_CLK * PEAK_CLOCK_GHZ * 1e9 / 1e12


def valu_roofline(ops: float, ms: float, what: str, cus: int = CUS) -> dict:
    """Roofline object of a VALU-bound kernel: algorithmic int32 lane-ops
    per second against the int32 VALU peak."""
    achieved = ops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": round(peak_tops(cus), 2),
            "unit": f"T lane-ops/s (algorithmic: {what})", "frac": round(achieved / peak_tops(cus), 4),
            "traffic": None}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# The reference's OpenMP runtime reads its placement once, when libgomp
# loads: threads bound to cores, packed (the CPU baseline's settings).
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")


def host_info() -> dict:
    """The host the CPU baselines run on: CPU model, logical CPUs, the CPUs
    this process may use, and the thread count used for "all cores" (the
    OMP_NUM_THREADS share where set, else the usable CPUs)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS") or usable)
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable, "threads_all": max(1, threads),
            "omp_proc_bind": os.environ.get("OMP_PROC_BIND"), "omp_places": os.environ.get("OMP_PLACES")}


def _median_time(fn, runs: int, warmup: int = 1) -> float:
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def _reference():
    from oracle.binding import Oracle, Reference

    if Reference.available():
        r = Reference()
        return r, "reference", r.march
    return Oracle(), "port", "x86-64-v3"


def cpu_baseline(pairs, target_s: float = 12.0, target_1t_s: float = 4.0):
    """Reference OpenMP path (oracle/_ref/libref*.so) on a bounded sample of the
    same pairs, on all cores and on one thread; falls back to the oracle's
    restatement when the reference build is absent.  Returns a dict for the
    JSON line."""
    host = host_info()
    threads = host["threads_all"]
    eng, kind, march = _reference()
    order = np.random.default_rng(99).permutation(len(pairs))

    def sample(cells_target):
        sel, tot = [], 0
        for i in order:
            sel.append(int(i))
            tot += len(pairs[i][0]) * len(pairs[i][1])
            if tot >= cells_target:
                break
        return sel, tot

    def timed(th, tgt):
        sel, cells = sample(5e8)
        sub = [(pairs[i][0].astype(np.int32), pairs[i][1].astype(np.int32)) for i in sel]
        t = time.perf_counter()
        eng.comb_batch(sub, threads=th)
        dt = time.perf_counter() - t
        want = min(cells / dt * tgt, sum(len(a) * len(b) for a, b in pairs))
        if want > 2 * cells:
            sel, cells = sample(want)
            sub = [(pairs[i][0].astype(np.int32), pairs[i][1].astype(np.int32)) for i in sel]
            t = time.perf_counter()
            eng.comb_batch(sub, threads=th)
            dt = time.perf_counter() - t
        return sel, cells, dt

    sel, cells, dt = timed(threads, target_s)
    sel1, cells1, dt1 = timed(1, target_1t_s)
    return {"value": cells / dt / 1e9, "unit": "GCUPS", "cores": threads, "kind": kind,
            "sample": f"{len(sel)} of the {len(pairs)} pairs (seeded random order), {cells:.3e} cells, "
                      f"{dt:.1f} s; sticky_braid_mpi per pair in an OpenMP parallel-for over pairs",
            "seconds": dt,
            "one_thread": {"value": round(cells1 / dt1 / 1e9, 3), "unit": "GCUPS", "cores": 1,
                           "sample": f"{len(sel1)} pairs, {cells1:.3e} cells, {dt1:.1f} s"},
            "build": f"g++ -O3 -march={march} -fopenmp", "host": host}


def cpu_baselines_configs(golden):
    """The reference's OpenMP paths for configs 1, 2, 3 and 5 on bounded
    samples of the same seeded inputs, on all cores and on one thread (one
    warm-up, then the median of several runs; one run for the long
    one-thread samples).  Bit-exactness of these functions against the
    engine is covered by the fixtures; here they are timed only."""
    from oracle.binding import RefAMatch
    from semilocal_amd import workloads as W

    host = host_info()
    T = host["threads_all"]
    eng, kind, march = _reference()
    meta = {"kind": kind, "build": f"g++ -O3 -march={march} -fopenmp", "threads_all": T}
    out = {}

    a, b = W.plain_pair()
    a, b = a.astype(np.int32), b.astype(np.int32)
    cells = len(a) * len(b)
    tA = _median_time(lambda: eng.comb(a, b, threads=T), 5)
    t1 = _median_time(lambda: eng.comb(a, b, threads=1), 5)
    out["plain_pair_10k"] = {"unit": "GCUPS", "all_cores": round(cells / tA / 1e9, 3), "one_thread":
                             round(cells / t1 / 1e9, 3), "function": "sticky_braid_mpi<int,false,true>",
                             "sample": "the whole 10,000 x 10,000 pair", **meta}

    a, b = W.binary_pair()
    if kind == "reference":
        sa = a[:1 << 16]
        bc = len(sa) * len(b)
        tA = _median_time(lambda: eng.llcs_binary(sa, b, threads=T), 3)
        t1 = _median_time(lambda: eng.llcs_binary(sa, b, threads=1), 1, warmup=0)
        out["binary_pair_2p20_score"] = {
            "unit": "G bit-cells/s", "all_cores": round(bc / tA / 1e9, 1), "one_thread": round(bc / t1 / 1e9, 1),
            "function": "binary::llcs_2symbol_smart_combing<uint64_t>",
            "sample": "a[:65536] x b (all 2^20 bits): 2^36 bit-cells", **meta}
    pa, pb = a[:1 << 13].astype(np.int32), b.astype(np.int32)
    pc = len(pa) * len(pb)
    tA = _median_time(lambda: eng.comb(pa, pb, threads=T), 3, warmup=0)
    t1 = _median_time(lambda: eng.comb(pa, pb, threads=1), 1, warmup=0)
    out["binary_pair_2p20_permutation"] = {
        "unit": "GCUPS", "all_cores": round(pc / tA / 1e9, 3), "one_thread": round(pc / t1 / 1e9, 3),
        "function": "sticky_braid_mpi<int,false,true>", "sample": "a[:8192] x b (2^20): 8.6e9 cells", **meta}

    a, b = W.similar_pair()
    ha, hb = a[:100_000].astype(np.int32), b[:100_000].astype(np.int32)
    hc = len(ha) * len(hb)
    if kind == "reference":
        t0 = time.perf_counter()
        eng.L.ref_precalc_warmup()
        pre = time.perf_counter() - t0
    else:
        pre = 0.0
    tA = _median_time(lambda: eng.hybrid_iterative(ha, hb, 8, 8, threads=T), 3, warmup=0)
    t1 = _median_time(lambda: eng.hybrid_iterative(ha, hb, 8, 8, threads=1), 1, warmup=0)
    out["similar_pair_1e6_hybrid"] = {
        "unit": "GCUPS", "all_cores": round(hc / tA / 1e9, 3), "one_thread": round(hc / t1 / 1e9, 3),
        "function": "hybrid_iterative<int,false>, 8x8 blocks", "precalc_map5_s": round(pre, 3),
        "sample": "a[:100000] x b[:100000] of the similar pair: 1e10 cells (precalc timed separately)", **meta}

    reqs = W.serving_stream(count=2000, rate_hz=1000.0)
    sel, tot = [], 0
    for r in reqs:
        if r.kind == "compare":
            sel.append(r)
            tot += len(r.a) * len(r.b)
            if tot >= 4e9:
                break
    sub = [(r.a.astype(np.int32), r.b.astype(np.int32)) for r in sel]
    t0 = time.perf_counter()
    perms = eng.comb_batch(sub, threads=T)
    tA = time.perf_counter() - t0
    t0 = time.perf_counter()
    eng.comb_batch(sub[:max(1, len(sub) // 4)], threads=1)
    t1 = time.perf_counter() - t0
    c1 = sum(len(x) * len(y) for x, y in sub[:max(1, len(sub) // 4)])
    rec = {"unit": "GCUPS", "all_cores": round(tot / tA / 1e9, 3), "one_thread": round(c1 / t1 / 1e9, 3),
           "function": "sticky_braid_mpi<int,false,false> per request, OpenMP parallel-for over requests",
           "sample": f"the first {len(sel)} comparisons of the serving stream ({tot:.2e} cells; one thread: "
                     f"the first {max(1, len(sub) // 4)})", **meta}
    if RefAMatch.available():
        ra = RefAMatch()
        pos = {r.rid: k for k, r in enumerate(sel)}
        qs = [r for r in reqs if r.kind == "query" and r.target in pos]
        nq, dq = 0, 0.0
        for r in qs:
            k = pos[r.target]
            m, n = len(sel[k].a), len(sel[k].b)
            lo, hi = r.window
            i, j = (lo, hi) if r.qkind == "string_substring" else (-lo, n + m - hi)
            dq += ra.semi_local_H(perms[k], m, np.array([i]), np.array([j]))[1]
            nq += 1
        rec["queries_per_s_one_thread"] = round(nq / dq, 1) if nq else None
        rec["query_function"] = "SemiLocalStringSubstringWrapper::dominanceSumRandom (O(m+n) scan)"
        rec["queries_timed"] = nq
    out["serving"] = rec
    out["host"] = host
    return out


def percentile(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def bench_serving(S, W, count: int, rate_hz: float, pool_gb: float):
    """Config 5: replay a Poisson request stream in real time against the
    request loop; latency = submission (at the scheduled arrival) to
    completion.  30% of requests are follow-up queries on earlier braids."""
    reqs = W.serving_stream(count=count, rate_hz=rate_hz)
    srv = S.Server(pool_bytes=int(pool_gb * (1 << 30)))
    ids = {}
    t0 = time.perf_counter()
    lag = 0.0
    for r in reqs:
        delay = r.arrival_s - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        else:
            lag = max(lag, -delay)
        if r.kind == "compare":
            ids[r.rid] = srv.submit_compare(r.a, r.b)
        else:
            kind = srv.SUBSTRING_STRING if r.qkind == "substring_string" else srv.STRING_SUBSTRING
            ids[r.rid] = srv.submit_query(ids[r.target], kind, *r.window)
    lat = {"compare": [], "query": []}
    vals = {}
    for r in reqs:
        v, ms = srv.result(ids[r.rid])
        lat[r.kind].append(ms)
        vals[r.rid] = v
    wall = time.perf_counter() - t0
    st = srv.stats()
    # spot parity: a sample of comparisons and queries against the one-shot path
    g = np.random.default_rng(0)
    cmp_ids = [r for r in reqs if r.kind == "compare"]
    checked = 0
    for r in g.choice(cmp_ids, size=min(8, len(cmp_ids)), replace=False):
        perm = S.semi_local_comb(r.a, r.b)
        assert vals[r.rid] == S.llcs_from_kernel(perm, len(r.a), len(r.b)), "serving parity failure"
        checked += 1
    qs = [r for r in reqs if r.kind == "query"][:8]
    for r in qs:
        t = reqs[r.target]
        perm = S.semi_local_comb(t.a, t.b)
        m, n = len(t.a), len(t.b)
        lo, hi = r.window
        if r.qkind == "string_substring":    # H(lo, hi)
            h = hi - lo - int(np.count_nonzero(perm[lo + m:] < hi))
        else:                                 # H(-lo, n+m-hi) - lo - (m-hi)
            jj = n + m - hi
            h = jj + lo - int(np.count_nonzero(perm[m - lo:] < jj)) - lo - (m - hi)
        assert vals[r.rid] == h, "serving query parity failure"
        checked += 1
    srv.close()
    allv = lat["compare"] + lat["query"]
    return {
        "workload": f"{count} requests, Poisson {rate_hz:g}/s, lengths log-normal (median 5000, sigma_ln 1.0) "
                    f"clipped to [200, 200000], alphabet 4, 30% follow-up queries (half substring-vs-string, half "
                    f"string-vs-substring), seed 5",
        "rate_hz": rate_hz, "requests": count,
        "p50_ms": round(percentile(allv, 50), 3), "p99_ms": round(percentile(allv, 99), 3),
        "compare_p50_ms": round(percentile(lat["compare"], 50), 3),
        "compare_p99_ms": round(percentile(lat["compare"], 99), 3),
        "query_p50_ms": round(percentile(lat["query"], 50), 3) if lat["query"] else None,
        "query_p99_ms": round(percentile(lat["query"], 99), 3) if lat["query"] else None,
        "gcups": round(st["cells_done"] / st["busy_seconds"] / 1e9, 2) if st["busy_seconds"] else None,
        "offered_gcups": round(st["cells_done"] / reqs[-1].arrival_s / 1e9, 2),
        "wall_s": round(wall, 3), "max_submit_lag_ms": round(lag * 1e3, 3),
        "launches": st["launches"], "deferrals": st["deferrals"], "evictions": st["evictions"],
        "peak_inflight": st["peak_inflight"], "peak_inflight_compares": st["peak_inflight_compares"],
        "pool_gb": round(st["pool_bytes"] / 2**30, 2), "pool_peak_mb": round(st["pool_peak"] / 2**20, 1),
        "parity_spot_checks": checked,
    }


def bench_serving_sweep(S, W, rates=(1000, 4000, 16000, 32000, 64000), slo_ms: float = 20.0,
                        pool_gb: float = 16.0):
    """Config 5 at several offered rates: each rate replays its own Poisson
    stream (about 1.5 s of arrivals, 2,000 to 16,000 requests) in real time;
    p50/p99 latency and served GCUPS per rate, and the highest rate whose p99
    stays within `slo_ms`."""
    rows = []
    for rate in rates:
        count = int(min(16000, max(2000, rate * 1.5)))
        r = bench_serving(S, W, count, float(rate), pool_gb)
        rows.append({k: r[k] for k in ("rate_hz", "requests", "p50_ms", "p99_ms", "compare_p99_ms", "query_p99_ms",
                                       "gcups", "offered_gcups", "max_submit_lag_ms", "launches", "deferrals",
                                       "peak_inflight_compares")})
        log(f"serving at {rate}/s: p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms, {r['gcups']} GCUPS served")
    ok = [row["rate_hz"] for row in rows if row["p99_ms"] is not None and row["p99_ms"] <= slo_ms]
    return {"workload": "the serving stream (seed 5) at each offered rate, replayed in real time",
            "slo_p99_ms": slo_ms, "max_rate_within_slo_hz": max(ok) if ok else None, "rates": rows}


def bench_index(S, W):
    """Resident-braid queries on one large braid (200k x 200k, alphabet 4):
    20,000 string-substring queries through the request loop with the braid
    scanned per query vs answered from its wavelet-matrix index, and the
    index's own build and query rates."""
    g = np.random.default_rng(9)
    m = n = 200_000
    a = g.integers(0, 4, m).astype(np.uint8)
    b = a.copy()
    sub = g.random(n) < 0.05                      # a similar string: 5% substitutions
    b[sub] = g.integers(0, 4, int(np.count_nonzero(sub))).astype(np.uint8)
    nq = 20_000
    lo = g.integers(0, n, nq)
    hi = np.minimum(n, lo + g.integers(1, 50_000, nq))
    out = {}
    answers = {}
    for mode, index_min in (("scan", 0), ("index", 1)):
        srv = S.Server(pool_bytes=8 << 30)
        srv.set_index(index_min)
        rid = srv.submit_compare(a, b)
        srv.result(rid)
        t0 = time.perf_counter()
        while index_min and srv.stats()["indexed_braids"] < 1 and time.perf_counter() - t0 < 30:
            time.sleep(0.0005)
        build_ms = (time.perf_counter() - t0) * 1e3
        t0 = time.perf_counter()
        qids = [srv.submit_query(rid, srv.STRING_SUBSTRING, int(x), int(y)) for x, y in zip(lo, hi)]
        vals = [srv.result(q)[0] for q in qids]
        dt = time.perf_counter() - t0
        st = srv.stats()
        answers[mode] = vals
        out[mode] = {"queries": nq, "seconds": round(dt, 3), "queries_per_s": round(nq / dt, 1),
                     "index_queries": st["index_queries"]}
        if index_min:
            out[mode]["index_build_ms"] = round(build_ms, 2)
        srv.close()
    perm = S.semi_local_comb(a, b)
    ix = S.BraidIndex(perm, m, n)
    kinds = np.full(100_000, 1, np.int32)
    qi = g.integers(0, n, 100_000).astype(np.int32)
    qj = np.minimum(n, qi + g.integers(0, 50_000, 100_000)).astype(np.int32)
    ix.query(kinds[:10], qi[:10], qj[:10])
    t0 = time.perf_counter()
    ix.query(kinds, qi, qj)
    qdt = time.perf_counter() - t0
    st = ix.stats()
    ix.close()
    same = answers["scan"] == answers["index"]
    return {"workload": "one resident braid of a 200,000 x 200,000 pair (alphabet 4); 20,000 string-substring "
                        "queries through the request loop",
            "scan": out["scan"], "index": out["index"],
            "standalone_index": {"levels": st["levels"], "bytes": st["bytes"], "queries": 100_000,
                                 "ms_incl_transfers": round(qdt * 1e3, 2)},
            "parity": "index answers equal scan answers" if same else "MISMATCH"}


def amatch_workload(n: int = 1_000_000, m: int = 1_000):
    """Pattern of 1,000 symbols (alphabet 4) and a text of n symbols holding
    n / 3,000 mutated copies of it (10% substitutions), seed 21."""
    g = np.random.default_rng(21)
    p = g.integers(0, 4, m).astype(np.uint8)
    t = g.integers(0, 4, n).astype(np.uint8)
    for pos in g.integers(0, n - m, n // 3000):
        c = p.copy()
        sub = g.random(m) < 0.10
        c[sub] = g.integers(0, 4, int(np.count_nonzero(sub))).astype(np.uint8)
        t[pos:pos + m] = c
    return p, t


def bench_amatch(S, k: float = 0.9):
    """Approximate matching: InteractiveDuplicateSearchSemiDense's windows on a
    1,000 x 1,000,000 pattern/text, k = 0.9 (w = 1,111, substrings of length
    >= 900: 212 columns per window triangle)."""
    p, t = amatch_workload()
    m, n = len(p), len(t)
    w, L = int(m / k), int(m * k)
    t0 = time.perf_counter()
    am = S.AMatch(p, t, w)
    create_s = time.perf_counter() - t0
    st = am.stats()
    comb_ms, win_ms = am.time(w, L, warmup=1, iters=5)

    def timed(variant, reps=5):
        am.duplicate_search(k, variant)                 # warm: pinned staging, host sets
        walls, splits = [], []
        for _ in range(reps):
            t0 = time.perf_counter()
            found = am.duplicate_search(k, variant)
            walls.append((time.perf_counter() - t0) * 1e3)
            splits.append(am.last_times())
        i = int(np.argsort(walls)[reps // 2])          # the median call
        return found, walls[i], splits[i]

    res, dense_ms, dense_split = timed(0)
    res_sparse, sparse_ms, sparse_split = timed(1)
    am.close()
    nw = n - w + 1
    S_cols = w - L + 1
    return {"workload": f"pattern {m} x text {n:,} (alphabet 4, {n // 3000} planted copies with 10% "
                        f"substitutions), k = {k}: w = {w}, min length {L}",
            "chunks": st["chunks"], "chunk_windows": st["chunk_windows"], "combed_cells": st["cells"],
            "create_s_incl_upload": round(create_s, 3),
            "comb_ms": round(comb_ms, 3), "comb_gcups": round(st["cells"] / (comb_ms * 1e-3) / 1e9, 1),
            "windows_ms": round(win_ms, 3), "windows": nw,
            "window_column_updates_per_s": round(nw * S_cols / (win_ms * 1e-3), 1),
            "search_timing": "median of 5 warm calls, wall time of the C entry point incl. host phases",
            "semi_dense_search_ms_incl_host": round(dense_ms, 2), "semi_dense_found": len(res),
            "semi_dense_device_and_copy_ms": round(dense_split[0], 2), "semi_dense_host_phases_ms":
                round(dense_split[1], 2),
            "semi_sparse_search_ms_incl_host": round(sparse_ms, 2), "semi_sparse_found": len(res_sparse),
            "semi_sparse_device_and_copy_ms": round(sparse_split[0], 2), "semi_sparse_host_phases_ms":
                round(sparse_split[1], 2)}


def cpu_baseline_amatch(k: float = 0.9, n_sample: int = 300_000):
    """The reference's InteractiveDuplicateSearchSemiDense::find (compiled in
    place, one thread) on the first n_sample symbols of the same text."""
    from oracle.binding import AMatchOracle, RefAMatch

    p, t = amatch_workload()
    t = t[:n_sample]
    eng, kind = (RefAMatch(), "reference") if RefAMatch.available() else (AMatchOracle(), "port")
    t0 = time.perf_counter()
    eng.dup_search(p, t, k, 0)
    dt = time.perf_counter() - t0
    nw = n_sample - int(len(p) / k) + 1
    return {"value": round(nw / dt, 1), "unit": "windows/s", "cores": 1, "kind": kind,
            "sample": f"text prefix of {n_sample:,} symbols ({nw:,} windows), {dt:.2f} s incl. its "
                      f"sequential combing"}


def bench_hyyro(S, W, pairs=None):
    """Score-only LLCS by Hyyro's bit-vector recurrence: the ragged batch's
    4096 pairs (LLCS of every pair), and one 20,000 x 1,000,000 pair."""
    pairs = pairs if pairs is not None else W.ragged_batch()
    t0 = time.perf_counter()
    hb = S.HyyroBatch(pairs)
    create_s = time.perf_counter() - t0
    scores = hb.run()
    default_ms = hb.time(warmup=1, iters=5)
    # lane width (32-bit words per lane) per length bucket: the autotuner times every offered
    # form on each bucket's own pairs and keeps the fastest
    t0 = time.perf_counter()
    hb.autotune(rounds=3)
    tune_s = time.perf_counter() - t0
    buckets = hb.buckets()
    if not np.array_equal(hb.run(), scores):
        raise RuntimeError("hyyro: scores changed with the autotuned words per lane")
    ms = hb.time(warmup=1, iters=5)
    st = hb.stats()
    hb.close()
    cells = int(sum(len(a) * len(b) for a, b in pairs))
    g = np.random.default_rng(17)
    a = g.integers(0, 4, 20_000).astype(np.uint8)
    b = g.integers(0, 4, 1_000_000).astype(np.uint8)
    one = S.HyyroBatch([(a, b)])
    one_default_ms = one.time(warmup=1, iters=3)
    one.autotune(rounds=2)
    one_wpl = [bk["words_per_lane"] for bk in one.buckets() if bk["pairs"]]
    one_ms = one.time(warmup=1, iters=3)
    one_st = one.stats()
    one.close()
    # short pairs: 4096 x (2000 x 2000), alphabet 20 -- Hyyro's LLCS against
    # the semi-local batch's full kernels (default tile shape) on the same pairs
    short = [(g.integers(0, 20, 2000).astype(np.uint8), g.integers(0, 20, 2000).astype(np.uint8))
             for _ in range(4096)]
    hs = S.HyyroBatch(short)
    short_default_ms = hs.time(warmup=1, iters=5)
    hs.autotune(rounds=3)
    short_wpl = [bk["words_per_lane"] for bk in hs.buckets() if bk["pairs"]]
    short_ms = hs.time(warmup=1, iters=5)
    short_llcs = hs.run()
    hs.close()
    sb = S.Batch(short)
    sb_ms, _ = sb.time(per_pair=False, warmup=1, iters=5)
    sb_llcs = [S.llcs_from_kernel(r, 2000, 2000) for r in sb.run().results()]
    sb.close()
    short_cells = 4096 * 2000 * 2000
    return {"workload": "LLCS of the ragged batch's 4096 pairs (lengths log-normal, alphabet 20); one 20,000 x "
                        "1,000,000 pair (alphabet 4); 4096 pairs of 2000 x 2000 (alphabet 20)",
            "batch_create_s_incl_recode_upload_tables": round(create_s, 3),
            "batch_words_per_lane_by_bucket": [{k: bk[k] for k in ("max_words", "pairs", "words_per_lane", "waves")}
                                               for bk in buckets],
            "batch_autotune_s": round(tune_s, 2), "batch_default_wpl_ms": round(default_ms, 3),
            "batch_ms": round(ms, 3), "batch_cells": cells, "batch_gcups": round(cells / (ms * 1e-3) / 1e9, 1),
            "batch_word_rows": st["word_rows"],
            "batch_word_rows_per_s": round(st["word_rows"] / (ms * 1e-3), 1),
            "batch_roofline": valu_roofline(st["word_rows"] * HYYRO_OPS_PER_WORD_ROW, ms, "8 per 64-bit word-row"),
            "batch_llcs_sum": int(scores.sum()),
            "long_pair_words_per_lane": one_wpl, "long_pair_default_wpl_ms": round(one_default_ms, 2),
            "long_pair_ms": round(one_ms, 2), "long_pair_gcups": round(2e10 / (one_ms * 1e-3) / 1e9, 1),
            "long_pair_word_rows": one_st["word_rows"],
            "short_pairs_words_per_lane": short_wpl, "short_pairs_default_wpl_ms": round(short_default_ms, 3),
            "short_pairs_4096x2000x2000_ms": round(short_ms, 3),
            "short_pairs_gcups": round(short_cells / (short_ms * 1e-3) / 1e9, 1),
            "short_pairs_semi_local_batch_ms": round(sb_ms, 3),
            "short_pairs_llcs_equal_semi_local": bool(np.array_equal(short_llcs, np.array(sb_llcs)))}


def cpu_baseline_hyyro(pairs, target_s: float = 5.0):
    """The reference's lllcs_hyyro (preprocess + hyyro_magic<uint64_t>,
    compiled in place, one thread) on a seeded sample of the same pairs;
    the oracle's DP when the reference build is absent."""
    from oracle.binding import Oracle, Reference

    eng, kind = (Reference(), "reference") if Reference.available() else (Oracle(), "port")
    fn = eng.hyyro if kind == "reference" else eng.llcs_dp
    order = np.random.default_rng(98).permutation(len(pairs))
    cells, done, t0 = 0, 0, time.perf_counter()
    for i in order:
        a, b = pairs[i]
        fn(a.astype(np.int32), b.astype(np.int32))
        cells += len(a) * len(b)
        done += 1
        if time.perf_counter() - t0 > target_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(cells / dt / 1e9, 2), "unit": "GCUPS", "cores": 1, "kind": kind,
            "sample": f"{done} of the {len(pairs)} pairs (seeded random order), {cells:.3e} cells, {dt:.1f} s"}


def bench_hybrid(S, W, golden, grid_autotune: bool = True):
    """Config 3: 8x8 blocks combed in one launch, merged by 6 levels of batched
    steady-ant multiplications; then (grid_autotune) the block split per
    length bucket."""
    a, b = W.similar_pair()
    h = S.Hybrid(a, b, 8, 8)
    t0 = time.perf_counter()
    tiles = h.autotune(iters=1)            # the combing launch's tile shape and queue order
    tune_s = time.perf_counter() - t0
    perm = h.run()
    ok = W.sha256(perm) == golden["similar_pair_1e6"]["perm_sha256"]
    tot, comb, merge = h.time(warmup=1, iters=3)
    cells, levels = h.stats()
    h.set_graph(True)                      # the same run captured once into a HIP graph, then replayed
    ok = ok and W.sha256(h.run()) == golden["similar_pair_1e6"]["perm_sha256"]
    gtot, _, _ = h.time(warmup=1, iters=3)
    h.close()
    out = {"workload": "similar pair: a 1e6 symbols (alphabet 4), b = a with 5% subs and 1% indels (seed 3); "
                       "8x8 blocks", "ms": round(tot, 3), "comb_ms": round(comb, 3), "merge_ms": round(merge, 3),
           "graph_ms": round(gtot, 3), "merge_levels": levels, "gcups": round(cells / (tot * 1e-3) / 1e9, 2),
           "comb_roofline": valu_roofline(cells * VALU_OPS_PER_CELL, comb, "5 per cell"),
           "graph_gcups": round(cells / (gtot * 1e-3) / 1e9, 2),
           "comb_tiles": {"rows_per_lane": tiles[0], "waves_per_cu": tiles[1], "autotune_s": round(tune_s, 1)}}
    if not grid_autotune:
        out["parity"] = "bit-exact vs the reference hybrid's fixture (eager and graph)" if ok else "MISMATCH"
        return out
    # block split per length bucket: the grid autotuner times every candidate
    # grid on a pair of the bucket's length and keeps the fastest for it
    grids = [(1, 1), (1, 2), (2, 1), (2, 2), (4, 4), (8, 8), (16, 16), (8, 16), (16, 8), (32, 32)]
    by_bucket = []
    t0 = time.perf_counter()
    for length in (100_000, 300_000, 1_000_000):
        x, y = (a, b) if length == 1_000_000 else W.similar_pair(seed=5, length=length)
        best, ms = S.hybrid_autotune_blocks(x, y, grids, iters=2)
        by_bucket.append({"length": length, "bucket_log2": int(np.log2(max(len(x), len(y)))),
                          "grid": list(best), "ms_by_grid": {f"{gm}x{gn}": round(t, 2) for (gm, gn), t in zip(grids, ms)}})
    split_s = time.perf_counter() - t0
    best = S.hybrid_bucket_blocks(len(a), len(b))
    hb = S.Hybrid(a, b, *best)
    ok = ok and W.sha256(hb.run()) == golden["similar_pair_1e6"]["perm_sha256"]
    hb.close()
    out["block_split_by_bucket"] = by_bucket
    out["block_split_autotune_s"] = round(split_s, 1)
    out["parity"] = ("bit-exact vs the reference hybrid's fixture (eager, graph and the autotuned grid)" if ok
                     else "MISMATCH")
    return out


def bench_multiply(S):
    """Kernel (c) alone: one product of two random permutations of 2^20."""
    n = 1 << 20
    g = np.random.default_rng(0)
    p, q = g.permutation(n).astype(np.int32), g.permutation(n).astype(np.int32)
    ms, el = S.braid_multiply_time([(p, q)], warmup=1, iters=5)
    # algorithmic bytes: read two int32 maps and write one per element per level
    gbs = 12.0 * el / (ms * 1e-3) / 1e9
    return {"workload": "R = P (.) Q, n = 2^20 random permutations", "ms": round(ms, 3),
            "levels_x_elements": el, "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": 8000.0,
                                                  "unit": "GB/s", "frac": round(gbs / 8000.0, 5),
                                                  "bytes_per_element_level": 12}}


def bench_binary(S, W, golden):
    """Config 2: bit-parallel LLCS of two 2^20-bit strings, and their full
    permutation by kernel (a): one launch, and the hybrid at its measured
    block grid (the faster of the two is `permutation_ms`)."""
    a, b = W.binary_pair()
    bp = S.BitParallel(a, b)
    form = bp.form                   # the plan's default: the carry form
    llcs = bp.llcs()
    ms = bp.time(warmup=1, iters=3)
    bp.form = S.BITPAR_COMBING       # the same score by sliding-diagonal combing
    comb_llcs = bp.llcs()
    comb_ms = bp.time(warmup=1, iters=3)
    bp.close()
    rec = golden["binary_pair_2p20"]
    t = time.perf_counter()
    perm = S.semi_local_comb(a, b)
    pms = (time.perf_counter() - t) * 1e3
    cells = len(a) * len(b)
    bt = S.Batch([(a, b)])           # the same permutation with inputs resident
    kms, _ = bt.time(warmup=1, iters=2)
    bt.close()
    # the same permutation by the hybrid (blocks combed in one launch, merged
    # by kernel (c)): its block grid measured, one block included.  Halving a
    # block's rows halves its chain of passes; finer grids put two waves on
    # a SIMD.
    grids = [(1, 1), (1, 2), (2, 1), (2, 2)]
    best, grid_ms = S.hybrid_autotune_blocks(a, b, grids, iters=2)
    hy = S.Hybrid(a, b, *best)
    hy_perm = hy.run()
    hms, hcomb, hmerge = hy.time(warmup=1, iters=2)
    hy.close()
    hy_ok = W.sha256(hy_perm) == rec["perm_sha256"]
    best_ms = min(kms, hms)
    return {"workload": "binary pair, |a| = |b| = 2^20 (seed 2)",
            "bitparallel_ms": round(ms, 3), "bitparallel_gcups": round(cells / (ms * 1e-3) / 1e9, 1),
            "bitparallel_form": "carry" if form == S.BITPAR_CARRY else "combing",
            "bitparallel_combing_ms": round(comb_ms, 3),
            "permutation_ms": round(best_ms, 2), "permutation_gcups": round(cells / (best_ms * 1e-3) / 1e9, 1),
            "permutation_one_launch_ms": round(kms, 2),
            "permutation_hybrid": {"grid": list(best), "ms": round(hms, 2), "comb_ms": round(hcomb, 2),
                                   "merge_ms": round(hmerge, 2),
                                   "grid_ms": {f"{g[0]}x{g[1]}": round(t, 2) for g, t in zip(grids, grid_ms)}},
            "permutation_ms_incl_transfers": round(pms, 1),
            "bitparallel_roofline": valu_roofline(cells * BITPAR_OPS_PER_BIT_CELL, ms, "0.93 per bit-cell"),
            "permutation_roofline": valu_roofline(cells * VALU_OPS_PER_CELL, best_ms, "5 per cell"),
            "parity": "bit-exact" if (llcs == comb_llcs == rec["llcs_bitparallel"]
                                      and W.sha256(perm) == rec["perm_sha256"] and hy_ok)
            else "MISMATCH"}


def bench_plain(S, W, golden):
    """Config 1: one 10k x 10k pair, one launch."""
    a, b = W.plain_pair()
    bt = S.Batch([(a, b)])
    res = bt.run().results()[0]
    ms, kms = bt.time(warmup=2, iters=10)
    bt.close()
    return {"workload": "plain pair 10000 x 10000, alphabet 4 (seed 1)", "ms": round(ms, 4),
            "gcups": round(len(a) * len(b) / (ms * 1e-3) / 1e9, 1), "kernel_ms": round(kms, 4),
            "roofline": valu_roofline(len(a) * len(b) * VALU_OPS_PER_CELL, kms, "5 per cell"),
            "parity": "bit-exact" if W.sha256(res) == golden["plain_pair_10k"]["perm_sha256"] else "MISMATCH"}



def tile_buckets(batch, pairs):
    """The batch's tile choice per length bucket (rows = the shorter string):
    pairs at each tile height and the critical chains (raised priority)."""
    rows, prio = batch.pair_tiles()
    m = np.array([min(len(x), len(y)) for x, y in pairs], np.int64)
    n = np.array([max(len(x), len(y)) for x, y in pairs], np.int64)
    edges = [0, 256, 512, 1024, 2048, 4096, 8192, 16384, 1 << 31]
    out = []
    for lo, hi in zip(edges[:-1], edges[1:]):
        k = (m >= lo) & (m < hi) & (rows > 0)
        if not k.any():
            continue
        out.append({"rows": f"[{lo}, {hi})" if hi < (1 << 31) else f">= {lo}", "pairs": int(k.sum()),
                    "cells": int((m[k] * n[k]).sum()),
                    "rows_per_lane": {str(R): int((rows[k] == R).sum()) for R in (4, 8, 16) if (rows[k] == R).any()},
                    "critical": int(prio[k].sum())})
    return out


def copy_probe(mb):
    """One plain device copy of `mb` MiB each way: the scale for the
    rocprofv3 FETCH_SIZE / WRITE_SIZE byte counters, profiled in the same
    process as the kernel they calibrate."""
    from semilocal_amd._lib import probe_copy

    nbytes = mb << 20
    pms_copy = probe_copy(nbytes)
    log(f"copy probe: {nbytes} bytes read, {nbytes} bytes written in {pms_copy:.3f} ms "
        f"({2 * nbytes / pms_copy / 1e6:.0f} GB/s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--per-pair-steps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="fixed tile shape (--rows-per-lane, --waves-per-cu, --schedule, --critical-fraction) "
                         "instead of the autotuner")
    ap.add_argument("--rows-per-lane", type=int, default=0, help="with --no-autotune: 4, 8, 16, or 0 = per pair")
    ap.add_argument("--waves-per-cu", type=int, default=8)
    ap.add_argument("--schedule", type=int, default=3, help="queue order 0-8 (with --no-autotune)")
    ap.add_argument("--critical-fraction", type=float, default=0.6, help="with --no-autotune and per-pair tiles")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="also run one plain device copy of this many MiB (PMC byte-counter scale)")
    ap.add_argument("--no-extras", action="store_true", help="only the ragged batch (configs 1, 2, 3, 5 skipped)")
    ap.add_argument("--only", choices=["plain", "binary", "hybrid", "multiply", "index", "serving", "amatch",
                                       "hyyro", "serving_sweep"],
                    help="run one extra section alone (for profiling) and print its JSON")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rate", type=float, default=1000.0)
    ap.add_argument("--serve-pool-gb", type=float, default=16.0)
    ap.add_argument("--serve-slo-ms", type=float, default=20.0, help="p99 latency SLO of the serving rate sweep")
    ap.add_argument("--no-grid-autotune", action="store_true",
                    help="--only hybrid: skip the block-split autotuner (the profiled runs stay config 3's own)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the batch shards by pair across GPUs by design, but no multi-GPU path is built).",
              file=sys.stderr)
        sys.exit(2)

    import semilocal_amd
    from semilocal_amd import workloads as W

    dev = semilocal_amd.device_info()
    log("device:", dev)
    if args.only:
        golden = json.load(open(os.path.join(ROOT, "tests", "golden", "configs.json")))
        fn = {"plain": lambda: bench_plain(semilocal_amd, W, golden),
              "binary": lambda: bench_binary(semilocal_amd, W, golden),
              "hybrid": lambda: bench_hybrid(semilocal_amd, W, golden, not args.no_grid_autotune),
              "multiply": lambda: bench_multiply(semilocal_amd),
              "index": lambda: bench_index(semilocal_amd, W),
              "amatch": lambda: bench_amatch(semilocal_amd),
              "hyyro": lambda: bench_hyyro(semilocal_amd, W),
              "serving": lambda: bench_serving(semilocal_amd, W, args.serve_requests, args.serve_rate,
                                               args.serve_pool_gb),
              "serving_sweep": lambda: bench_serving_sweep(semilocal_amd, W, slo_ms=args.serve_slo_ms,
                                                           pool_gb=args.serve_pool_gb)}[args.only]
        print(json.dumps({"section": args.only, **fn()}), flush=True)
        if args.copy_probe_mb:
            copy_probe(args.copy_probe_mb)
        return
    t0 = time.perf_counter()
    pairs = W.ragged_batch()
    log(f"workload generated in {time.perf_counter() - t0:.1f}s")
    batch = semilocal_amd.Batch(pairs, rows_per_lane=args.rows_per_lane)
    if args.no_autotune:
        batch.set_schedule(args.schedule)
        batch.set_critical_fraction(args.critical_fraction)
        batch.set_tiles(args.rows_per_lane, args.waves_per_cu)
    else:
        t0 = time.perf_counter()
        tiles = batch.autotune(iters=2)
        log(f"autotuner picked rows_per_lane={tiles[0]} waves_per_cu={tiles[1]} schedule={batch.schedule()} "
            f"critical_fraction={batch.critical_fraction()} in {time.perf_counter() - t0:.1f}s")
    tiles = batch.tiles()
    st = batch.stats()
    cells = st["cells"]
    buckets = tile_buckets(batch, pairs)
    log("batch:", st, "tile buckets:", buckets)

    parity = None
    if not args.no_parity:
        golden = os.path.join(ROOT, "tests", "golden", "configs.json")
        ref = json.load(open(golden))["ragged_batch_4096"]
        res = batch.run().results()
        ok = sum(W.sha256(p) == h for p, h in zip(res, ref["perm_sha256_each"]))
        parity = f"{ok}/{len(res)} pairs bit-exact vs the reference's fixture"
        log("parity:", parity)
        if ok != len(res):
            print(json.dumps({"metric": METRIC, "error": "parity failure", "parity": parity}))
            sys.exit(1)

    ms, kms = batch.time(per_pair=False, warmup=args.warmup, iters=args.steps)
    log(f"batched: {ms:.3f} ms/step (kernel {kms:.3f} ms)")
    pp_rows = None
    if args.per_pair_steps > 0:
        # the plain form at its own best tile height (one pair's passes run in
        # parallel, so it favours short passes), then back to the tuned shape
        pms = float("inf")
        for r in (4, 8, 16):
            batch.set_tiles(r, tiles[1])
            t_r, k_r = batch.time(per_pair=True, warmup=1, iters=args.per_pair_steps)
            log(f"one launch per pair, rows_per_lane={r}: {t_r:.3f} ms/step (kernels {k_r:.3f} ms)")
            if t_r < pms:
                pms, pp_rows = t_r, r
        batch.set_tiles(tiles[0], tiles[1])
    else:
        pms = float("nan")
    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)

    gcups = cells / (ms * 1e-3) / 1e9
    gcups_pp = cells / (pms * 1e-3) / 1e9
    peak_tops = dev["compute_units"] * LANES_PER_CU_CLK * PEAK_CLOCK_GHZ * 1e9 / 1e12
    # HBM bytes per launch of this kernel, measured by two rocprofv3 --pmc
    # passes and scaled by a copy kernel of known size (profiles/)
    # and the executed VALU instructions and VALU busy of the same kernel at
    # this tile shape from a third pass (SQ_INSTS_VALU, GRBM_GUI_ACTIVE)
    traffic = executed = None
    tf = os.path.join(ROOT, "profiles", "comb_batch_kernel_ragged4096.traffic.json")
    if os.path.exists(tf):
        rec = json.load(open(tf)).get("kernels", {}).get(f"comb_batch_kernel<uint8_t,{tiles[0]}>")
        if rec:
            traffic = rec["traffic_bytes_per_launch"]
            if rec.get("sq_insts_valu_per_dispatch"):
                ex_tops = rec["sq_insts_valu_per_dispatch"] * 64 / (kms * 1e-3) / 1e12
                executed = {"lane_ops_per_cell": round(rec["sq_insts_valu_per_dispatch"] * 64 / cells, 3),
                            "achieved": round(ex_tops, 3), "frac": round(ex_tops / peak_tops, 4),
                            "valu_busy": rec.get("valu_busy"),
                            "source": "rocprofv3 --pmc SQ_INSTS_VALU (profiles/comb_batch_kernel_ragged4096.md) "
                                      "over this run's kernel time"}
    achieved_tops = cells * VALU_OPS_PER_CELL / (kms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": round(gcups, 2),
        "unit": "GCUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "ragged batch: 4096 pairs, lengths log-normal (median 2000, sigma_ln 1.0) "
                               "clipped to [100, 50000], alphabet 20, half mutated (10% subs, 2% indels), seed 4",
                   "pairs": st["pairs"], "cells": cells, "tiles": st["tiles"],
                   "rows_per_lane": tiles[0] if tiles[0] else "per pair", "waves_per_cu": tiles[1],
                   "queue_schedule": batch.schedule(),
                   "critical_fraction": batch.critical_fraction() if tiles[0] == 0 else None,
                   "tile_buckets": buckets},
        "one_pair_per_launch": ({"value": round(gcups_pp, 2), "unit": "GCUPS", "ms_per_step": round(pms, 3),
                                 "rows_per_lane": pp_rows} if args.per_pair_steps > 0 else None),
        "batched_vs_per_pair": round(gcups / gcups_pp, 2) if args.per_pair_steps > 0 else None,
        "parity": parity,
        "roofline": {"bound": "valu", "achieved": round(achieved_tops, 3), "peak": round(peak_tops, 2),
                     "unit": "T lane-ops/s (algorithmic: 5 int32 lane-ops per cell, the generic branchless "
                             "cell; the gfx950 cell executes 3)",
                     "frac": round(achieved_tops / peak_tops, 4),
                     "traffic": traffic,
                     "kernel": f"comb_batch_kernel<uint8_t,{tiles[0]}>", "kernel_ms": round(kms, 4),
                     "ops_per_cell": VALU_OPS_PER_CELL, "executed": executed},
    }
    batch.close()
    if not args.no_extras:
        golden = json.load(open(os.path.join(ROOT, "tests", "golden", "configs.json")))
        for key, fn in (("plain_pair_10k", lambda: bench_plain(semilocal_amd, W, golden)),
                        ("binary_pair_2p20", lambda: bench_binary(semilocal_amd, W, golden)),
                        ("similar_pair_1e6_hybrid", lambda: bench_hybrid(semilocal_amd, W, golden)),
                        ("braid_multiply_2p20", lambda: bench_multiply(semilocal_amd)),
                        ("braid_index_queries", lambda: bench_index(semilocal_amd, W)),
                        ("approximate_matching_1e6", lambda: bench_amatch(semilocal_amd)),
                        ("hyyro_llcs", lambda: bench_hyyro(semilocal_amd, W, pairs)),
                        ("serving", lambda: bench_serving(semilocal_amd, W, args.serve_requests, args.serve_rate,
                                                          args.serve_pool_gb)),
                        ("serving_sweep", lambda: bench_serving_sweep(semilocal_amd, W, slo_ms=args.serve_slo_ms,
                                                                      pool_gb=args.serve_pool_gb))):
            t0 = time.perf_counter()
            out[key] = fn()
            log(f"{key}: {out[key]}  ({time.perf_counter() - t0:.1f}s)")
    if not args.no_cpu_baseline:
        cb = cpu_baseline(pairs)
        log("cpu baseline:", cb)
        out["cpu_baseline"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in cb.items() if k != "seconds"}
        if "hyyro_llcs" in out:
            out["hyyro_llcs"]["cpu_reference"] = cpu_baseline_hyyro(pairs)
            log("hyyro cpu reference:", out["hyyro_llcs"]["cpu_reference"])
        if "approximate_matching_1e6" in out:
            out["approximate_matching_1e6"]["cpu_reference"] = cpu_baseline_amatch()
            log("approximate matching cpu reference:", out["approximate_matching_1e6"]["cpu_reference"])
        if not args.no_extras:
            t0 = time.perf_counter()
            cfg = cpu_baselines_configs(golden)
            for key in ("plain_pair_10k", "similar_pair_1e6_hybrid", "serving"):
                if key in out and key in cfg:
                    out[key]["cpu_baseline"] = cfg[key]
            if "binary_pair_2p20" in out:
                out["binary_pair_2p20"]["cpu_baseline"] = {k: cfg[k] for k in
                                                           ("binary_pair_2p20_score", "binary_pair_2p20_permutation")
                                                           if k in cfg}
            log(f"config cpu baselines: {cfg}  ({time.perf_counter() - t0:.1f}s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
