"""Readers for the golden vectors in tests/golden (written by make_golden.py)."""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def comb_sweep():
    z = np.load(os.path.join(GOLDEN, "comb_sweep.npz"))
    a, b, p, meta = z["a"], z["b"], z["perm"], z["meta"]
    out, ia, ib, ip = [], 0, 0, 0
    for m, n, sigma in meta.tolist():
        out.append((a[ia:ia + m], b[ib:ib + n], p[ip:ip + m + n], sigma))
        ia, ib, ip = ia + m, ib + n, ip + m + n
    return out


def steady_ant():
    z = np.load(os.path.join(GOLDEN, "steady_ant.npz"))
    out, s = [], 0
    for n in z["n"].tolist():
        out.append((z["p"][s:s + n], z["q"][s:s + n], z["r"][s:s + n]))
        s += n
    return out


def hybrid_small():
    z = np.load(os.path.join(GOLDEN, "hybrid_small.npz"))
    out, ia, ib, ip = [], 0, 0, 0
    for m, n, sigma, sm, sn in z["meta"].tolist():
        out.append((z["a"][ia:ia + m], z["b"][ib:ib + n], z["perm"][ip:ip + m + n], sm, sn))
        ia, ib, ip = ia + m, ib + n, ip + m + n
    return out


def bitpar():
    with open(os.path.join(GOLDEN, "bitpar.json")) as f:
        return [(np.array(c["a"], np.uint8), np.array(c["b"], np.uint8), c["llcs"]) for c in json.load(f)]


def configs():
    with open(os.path.join(GOLDEN, "configs.json")) as f:
        return json.load(f)


def serving_stream():
    with open(os.path.join(GOLDEN, "serving_stream.json")) as f:
        return json.load(f)
