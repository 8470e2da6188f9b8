"""Writes the golden vectors under tests/golden/ from the reference's OWN
implementation (oracle/_ref/libref.so: the reference headers compiled in
place by oracle/Makefile).  Run in the container that holds the reference:

    make -C oracle && python tests/golden/make_golden.py [--full]

Outputs (data only: inputs or their generator seed + SHA-256, and the
reference's outputs):
  comb_sweep.npz     small pairs (m, n in 1..300, sigma 2..26, m > n too):
                     full row->col permutations from sticky_braid_mpi
  steady_ant.npz     random permutation pairs and their products
                     (steady_ant_parallel_wrapper with precalc(map, 5))
  hybrid_small.npz   hybrid_iterative on small pairs with several block grids
  bitpar.json        LLCS of binary pairs from llcs_2symbol_smart_combing
  configs.json       full-size workloads: seed, input SHA-256, SHA-256 of the
                     row->col permutation, LLCS, sampled entries
  serving_stream.json  the serving parity stream (workloads.serving_parity_stream:
                     config 5's law, 320 requests, both query kinds): per
                     comparison the input SHA-256, the permutation's SHA-256
                     and LLCS from sticky_braid_mpi; per follow-up query the
                     LCS read from the reference's permutation by the semi-local
                     H formulas (approximate_matching/utils.h:281-293, 331-342)
--full also runs the long ones (the 2^20 binary pair and the 10^6 hybrid).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.binding import Oracle, Reference  # noqa: E402
from semilocal_amd import workloads as W  # noqa: E402


def perm_record(perm, m, n, seed_sample=0):
    g = np.random.default_rng(seed_sample)
    idx = np.sort(g.choice(len(perm), size=min(64, len(perm)), replace=False))
    return {
        "perm_sha256": W.sha256(np.asarray(perm, np.int32)),
        "llcs": int(m - np.count_nonzero(np.asarray(perm)[:m] >= n)),
        "sample_idx": idx.tolist(),
        "sample_val": np.asarray(perm)[idx].astype(int).tolist(),
    }


def comb_sweep(ref: Reference):
    g = np.random.default_rng(12345)
    shapes = [(1, 1), (1, 7), (7, 1), (2, 300), (300, 2), (64, 64), (65, 63), (63, 65), (127, 129),
              (128, 128), (129, 1000), (1000, 129), (1024, 1024), (1025, 2049), (3000, 70)]
    for _ in range(150):
        shapes.append((int(g.integers(1, 301)), int(g.integers(1, 301))))
    A, B, P, meta = [], [], [], []
    for m, n in shapes:
        sigma = int(g.choice([2, 4, 5, 11, 20, 26]))
        a = g.integers(0, sigma, m).astype(np.int32)
        b = g.integers(0, sigma, n).astype(np.int32)
        p = ref.comb(a, b, threads=2)
        A.append(a)
        B.append(b)
        P.append(p)
        meta.append((m, n, sigma))
    np.savez_compressed(os.path.join(HERE, "comb_sweep.npz"), a=np.concatenate(A), b=np.concatenate(B),
                        perm=np.concatenate(P), meta=np.array(meta, np.int32))
    print("comb_sweep:", len(meta), "pairs")


def steady_ant(ref: Reference):
    g = np.random.default_rng(777)
    sizes = list(range(1, 41)) + [63, 64, 65, 100, 127, 128, 129, 255, 256, 300, 1000, 4096, 5000, 20000]
    Ps, Qs, Rs, ns = [], [], [], []
    for n in sizes:
        p = g.permutation(n).astype(np.int32)
        q = g.permutation(n).astype(np.int32)
        r = ref.steady_ant(p, q, nested=0)
        Ps.append(p)
        Qs.append(q)
        Rs.append(r)
        ns.append(n)
    np.savez_compressed(os.path.join(HERE, "steady_ant.npz"), p=np.concatenate(Ps), q=np.concatenate(Qs),
                        r=np.concatenate(Rs), n=np.array(ns, np.int32))
    print("steady_ant:", len(ns), "products")


def hybrid_small(ref: Reference):
    g = np.random.default_rng(4242)
    A, B, P, meta = [], [], [], []
    for m, n, sm, sn in [(100, 100, 2, 2), (300, 200, 3, 2), (200, 300, 1, 4), (513, 257, 4, 1), (1000, 1000, 8, 8),
                         (777, 1234, 3, 5), (64, 2000, 1, 16), (2000, 64, 16, 1), (5, 9, 1, 1), (31, 17, 2, 3)]:
        sigma = int(g.choice([2, 4, 20]))
        a = g.integers(0, sigma, m).astype(np.int32)
        b = g.integers(0, sigma, n).astype(np.int32)
        p = ref.hybrid_iterative(a, b, sm, sn, threads=2)
        A.append(a)
        B.append(b)
        P.append(p)
        meta.append((m, n, sigma, sm, sn))
    np.savez_compressed(os.path.join(HERE, "hybrid_small.npz"), a=np.concatenate(A), b=np.concatenate(B),
                        perm=np.concatenate(P), meta=np.array(meta, np.int32))
    print("hybrid_small:", len(meta), "cases")


def bitpar(ref: Reference):
    g = np.random.default_rng(99)
    cases = []
    for la, lb in [(64, 64), (64, 128), (128, 64 * 7), (64 * 5, 64 * 5), (1024, 2048), (4096, 4096), (64 * 33, 64 * 70)]:
        a = g.integers(0, 2, la).astype(np.uint8)
        b = g.integers(0, 2, lb).astype(np.uint8)
        llcs = ref.llcs_binary(a, b, threads=2)
        cases.append({"a": a.tolist(), "b": b.tolist(), "llcs": int(llcs)})
    with open(os.path.join(HERE, "bitpar.json"), "w") as f:
        json.dump(cases, f)
    print("bitpar:", len(cases), "cases")


def configs(ref: Reference, full: bool):
    path = os.path.join(HERE, "configs.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    threads = os.cpu_count() or 1

    a, b = W.plain_pair()
    t = time.time()
    p = ref.comb(a.astype(np.int32), b.astype(np.int32), threads=1)
    out["plain_pair_10k"] = {"generator": "workloads.plain_pair(seed=1, length=10000, sigma=4)",
                             "input_sha256": W.sha256(a, b), "m": len(a), "n": len(b),
                             "reference": "semi_local::sticky_braid_mpi<int,false,true>", **perm_record(p, len(a), len(b))}
    print("plain_pair_10k", time.time() - t)

    pairs = W.ragged_batch()
    t = time.time()
    perms = ref.comb_batch([(x.astype(np.int32), y.astype(np.int32)) for x, y in pairs], threads=threads)
    out["ragged_batch_4096"] = {
        "generator": "workloads.ragged_batch(seed=4, count=4096, median=2000, sigma_ln=1.0, lo=100, hi=50000, sigma=20)",
        "input_sha256": W.sha256(*[np.concatenate([x, y]) for x, y in pairs]),
        "pairs": len(pairs),
        "cells": int(sum(len(x) * len(y) for x, y in pairs)),
        "perm_sha256_each": [W.sha256(q) for q in perms],
        "llcs_each": [int(len(x) - np.count_nonzero(q[:len(x)] >= len(y))) for (x, y), q in zip(pairs, perms)],
        "reference": "semi_local::sticky_braid_mpi<int,false,false> per pair, OpenMP parallel-for over pairs",
    }
    print("ragged_batch_4096", time.time() - t)

    if full:
        a, b = W.binary_pair()
        t = time.time()
        p = ref.comb(a.astype(np.int32), b.astype(np.int32), threads=threads)
        o = Oracle()
        llcs_bits = ref.llcs_binary(a, b, threads=threads)
        out["binary_pair_2p20"] = {"generator": "workloads.binary_pair(seed=2, length=1<<20)",
                                   "input_sha256": W.sha256(a, b), "m": len(a), "n": len(b),
                                   "reference": "semi_local::sticky_braid_mpi<int,false,true> (permutation); "
                                                "binary::llcs_2symbol_smart_combing<uint64_t> (score)",
                                   "llcs_bitparallel": int(llcs_bits), **perm_record(p, len(a), len(b))}
        del o
        print("binary_pair_2p20", time.time() - t)

        a, b = W.similar_pair()
        t = time.time()
        p = ref.hybrid_iterative(a.astype(np.int32), b.astype(np.int32), 8, 8, threads=threads)
        out["similar_pair_1e6"] = {"generator": "workloads.similar_pair(seed=3, length=1000000, sigma=4, sub=0.05, indel=0.01)",
                                   "input_sha256": W.sha256(a, b), "m": len(a), "n": len(b),
                                   "reference": "semi_local::hybrid_iterative<int,false> with 8x8 blocks and precalc(map,5)",
                                   **perm_record(p, len(a), len(b))}
        print("similar_pair_1e6", time.time() - t)

    with open(path, "w") as f:
        json.dump(out, f, indent=1)


def H(perm, m, n, i, j):
    """H(i, j) = j - i - #{r >= i + m : P[r] < j} (approximate_matching/utils.h:281-293)."""
    return j - i - int(np.count_nonzero(perm[i + m:] < j))


def serving(ref: Reference):
    reqs = W.serving_parity_stream()
    cmp = [r for r in reqs if r.kind == "compare"]
    t = time.time()
    perms = ref.comb_batch([(r.a.astype(np.int32), r.b.astype(np.int32)) for r in cmp],
                           threads=os.cpu_count() or 1)
    by_rid = {}
    out_cmp = []
    for r, p in zip(cmp, perms):
        m, n = len(r.a), len(r.b)
        by_rid[r.rid] = (p, m, n)
        out_cmp.append({"rid": r.rid, "m": m, "n": n, "input_sha256": W.sha256(r.a, r.b),
                        "perm_sha256": W.sha256(p), "llcs": int(m - np.count_nonzero(p[:m] >= n))})
    out_q = []
    for r in reqs:
        if r.kind != "query":
            continue
        p, m, n = by_rid[r.target]
        s_, e_ = r.window
        if r.qkind == "string_substring":          # LCS(a, b[s:e)) = H(s, e)
            v = H(p, m, n, s_, e_)
        else:                                      # LCS(a[s:e), b) = H(-s, n+m-e) - s - (m-e)
            v = H(p, m, n, -s_, n + m - e_) - s_ - (m - e_)
        out_q.append({"rid": r.rid, "target": r.target, "qkind": r.qkind, "start": s_, "stop": e_, "lcs": int(v)})
    rec = {"generator": "workloads.serving_parity_stream(count=320, rate_hz=2000.0)",
           "reference": "semi_local::sticky_braid_mpi<int,false,false> per comparison; queries by the H formulas "
                        "of approximate_matching/utils.h:281-342 on the reference's permutation",
           "requests": len(reqs), "compares": out_cmp, "queries": out_q}
    with open(os.path.join(HERE, "serving_stream.json"), "w") as f:
        json.dump(rec, f, indent=0)
    print("serving:", len(out_cmp), "comparisons,", len(out_q), "queries", time.time() - t)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--full", action="store_true")
    ap.add_argument("--only", default="")
    args = ap.parse_args()
    ref = Reference()
    steps = {"comb": comb_sweep, "steady": steady_ant, "hybrid": hybrid_small, "bitpar": bitpar, "serving": serving}
    for k, fn in steps.items():
        if not args.only or k in args.only.split(","):
            fn(ref)
    if not args.only or "configs" in args.only.split(","):
        configs(ref, args.full)


if __name__ == "__main__":
    main()
