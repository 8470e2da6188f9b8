"""ctypes bindings for the CPU checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
(as the checker) and by the ``cpu_baseline`` leg of ``bench.py``.  The product
package ``semilocal_amd`` never imports this module.

``Oracle`` wraps liboracle.so (the C restatement in oracle.c);
``Reference`` wraps _ref/libref.so (the reference's own headers compiled by
oracle/Makefile), which exists only where it was built.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")
REF_V4_SO = os.path.join(HERE, "_ref", "libref_v4.so")
_V4_FLAGS = {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"}


def host_has_avx512() -> bool:
    """The x86-64-v4 feature set (AVX-512 F/BW/CD/DQ/VL) on this host."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return _V4_FLAGS <= set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return False


def reference_so() -> str:
    """The reference build for this host: x86-64-v4 where AVX-512 is there."""
    return REF_V4_SO if os.path.exists(REF_V4_SO) and host_has_avx512() else REF_SO

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (and _ref/libref.so where the reference exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _i32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        L.orc_comb_rowmajor.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p]
        L.orc_comb_antidiag.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p, C.c_int]
        L.orc_comb_batch.argtypes = [_i32p, _i64p, _i32p, _i32p, _i64p, _i32p, _i64p, _i32p, C.c_int, C.c_int]
        L.orc_llcs_from_perm.argtypes = [_i32p, C.c_int, C.c_int]
        L.orc_llcs_dp.argtypes = [_i32p, C.c_int, _i32p, C.c_int]
        L.orc_kernel_bruteforce.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p]
        L.orc_semi_local_H.argtypes = [_i32p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_mult_naive.argtypes = [_i32p, _i32p, C.c_int, _i32p]
        L.orc_steady_ant.argtypes = [_i32p, _i32p, C.c_int, _i32p]
        L.orc_staggered.argtypes = [_i32p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int, _i32p]
        L.orc_hybrid_iterative.argtypes = [_i32p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int, _i32p, C.c_int]
        L.orc_bitpar_llcs.argtypes = [_u64p, C.c_int, _u64p, C.c_int, C.c_int, C.c_int]
        L.orc_pack_bits_reverse.argtypes = [_u8p, C.c_int, _u64p]
        L.orc_pack_bits.argtypes = [_u8p, C.c_int, _u64p]
        L.orc_openmp_max_threads.restype = C.c_int

    def comb_rowmajor(self, a, b):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.orc_comb_rowmajor(a, len(a), b, len(b), out)
        return out

    def comb(self, a, b, threads: int = 1):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.orc_comb_antidiag(a, len(a), b, len(b), out, threads)
        return out

    def comb_batch(self, pairs, threads: int = 1):
        return _batch_call(self.L.orc_comb_batch, pairs, threads)

    def llcs_from_perm(self, perm, m, n) -> int:
        return self.L.orc_llcs_from_perm(_i32(perm), m, n)

    def llcs_dp(self, a, b) -> int:
        a, b = _i32(a), _i32(b)
        return self.L.orc_llcs_dp(a, len(a), b, len(b))

    def kernel_bruteforce(self, a, b):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.orc_kernel_bruteforce(a, len(a), b, len(b), out)
        return out

    def H(self, perm, m, n, i, j) -> int:
        return self.L.orc_semi_local_H(_i32(perm), m, n, i, j)

    def mult_naive(self, p, q):
        p, q = _i32(p), _i32(q)
        out = np.empty(len(p), np.int32)
        self.L.orc_mult_naive(p, q, len(p), out)
        return out

    def steady_ant(self, p, q):
        p, q = _i32(p), _i32(q)
        out = np.empty(len(p), np.int32)
        self.L.orc_steady_ant(p, q, len(p), out)
        return out

    def staggered(self, p, q, k, row_glue: bool):
        p, q = _i32(p), _i32(q)
        out = np.empty(len(p) + len(q) - k, np.int32)
        self.L.orc_staggered(p, len(p), q, len(q), k, int(row_glue), out)
        return out

    def hybrid_iterative(self, a, b, small_m, small_n, threads: int = 1):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.orc_hybrid_iterative(a, len(a), b, len(b), small_m, small_n, out, threads)
        return out

    def pack_bits_reverse(self, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        out = np.zeros((len(bits) + 63) // 64, np.uint64)
        self.L.orc_pack_bits_reverse(bits, len(bits), out)
        return out

    def pack_bits(self, bits):
        bits = np.ascontiguousarray(bits, dtype=np.uint8)
        out = np.zeros((len(bits) + 63) // 64, np.uint64)
        self.L.orc_pack_bits(bits, len(bits), out)
        return out

    def bitpar_llcs(self, a_bits, b_bits, threads: int = 1) -> int:
        ar, bw = self.pack_bits_reverse(a_bits), self.pack_bits(b_bits)
        return self.L.orc_bitpar_llcs(ar, len(ar), bw, len(bw), len(a_bits), threads)

    def max_threads(self) -> int:
        return self.L.orc_openmp_max_threads()


def _batch_layout(pairs):
    a_len = np.array([len(a) for a, _ in pairs], np.int32)
    b_len = np.array([len(b) for _, b in pairs], np.int32)
    a_off = np.zeros(len(pairs), np.int64)
    b_off = np.zeros(len(pairs), np.int64)
    p_off = np.zeros(len(pairs), np.int64)
    if len(pairs) > 1:
        a_off[1:] = np.cumsum(a_len[:-1], dtype=np.int64)
        b_off[1:] = np.cumsum(b_len[:-1], dtype=np.int64)
        p_off[1:] = np.cumsum((a_len + b_len)[:-1], dtype=np.int64)
    a_cat = _i32(np.concatenate([np.asarray(a) for a, _ in pairs])) if pairs else np.zeros(0, np.int32)
    b_cat = _i32(np.concatenate([np.asarray(b) for _, b in pairs])) if pairs else np.zeros(0, np.int32)
    return a_cat, a_off, a_len, b_cat, b_off, b_len, p_off


def _batch_call(fn, pairs, threads):
    a_cat, a_off, a_len, b_cat, b_off, b_len, p_off = _batch_layout(pairs)
    total = int((a_len.astype(np.int64) + b_len).sum())
    out = np.empty(max(total, 1), np.int32)
    fn(a_cat, a_off, a_len, b_cat, b_off, b_len, p_off, out, len(pairs), threads)
    res = []
    for i in range(len(pairs)):
        s = int(p_off[i])
        res.append(out[s:s + int(a_len[i]) + int(b_len[i])].copy())
    return res


class Reference:
    """The reference's own OpenMP implementation, compiled in place."""

    def __init__(self, path: str | None = None):
        path = path or reference_so()
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.path = path
        self.march = "x86-64-v4" if path.endswith("_v4.so") else "x86-64-v3"
        L = C.CDLL(path)
        self.L = L
        L.ref_sticky_braid_mpi.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p, C.c_int]
        L.ref_sticky_braid_sequential.argtypes = [_i32p, C.c_int, _i32p, C.c_int, _i32p]
        L.ref_sticky_braid_batch.argtypes = [_i32p, _i64p, _i32p, _i32p, _i64p, _i32p, _i64p, _i32p, C.c_int, C.c_int]
        L.ref_hybrid_iterative.argtypes = [_i32p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int, _i32p, C.c_int]
        L.ref_steady_ant.argtypes = [_i32p, _i32p, C.c_int, _i32p, C.c_int]
        L.ref_llcs_2symbol_smart.argtypes = [_u64p, C.c_int, _u64p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ref_llcs_2symbol_smart.restype = C.c_int
        L.ref_hyyro.argtypes = [_i32p, C.c_int, _i32p, C.c_int]
        L.ref_hyyro.restype = C.c_int
        L.ref_precalc_warmup.argtypes = []

    @staticmethod
    def available(path: str = REF_SO) -> bool:
        return os.path.exists(path)

    def comb(self, a, b, threads: int = 1):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.ref_sticky_braid_mpi(a, len(a), b, len(b), out, threads)
        return out

    def comb_rowmajor(self, a, b):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.ref_sticky_braid_sequential(a, len(a), b, len(b), out)
        return out

    def comb_batch(self, pairs, threads: int = 1):
        return _batch_call(self.L.ref_sticky_braid_batch, pairs, threads)

    def hybrid_iterative(self, a, b, small_m, small_n, threads: int = 1):
        a, b = _i32(a), _i32(b)
        out = np.empty(len(a) + len(b), np.int32)
        self.L.ref_hybrid_iterative(a, len(a), b, len(b), small_m, small_n, out, threads)
        return out

    def steady_ant(self, p, q, nested: int = 0):
        p, q = _i32(p), _i32(q)
        out = np.empty(len(p), np.int32)
        self.L.ref_steady_ant(p, q, len(p), out, nested)
        return out

    def hyyro(self, a, b) -> int:
        """lllcs_hyyro::preprocess + hyyro_magic<uint64_t> over (a, b)."""
        a, b = _i32(a), _i32(b)
        return self.L.ref_hyyro(a, len(a), b, len(b))

    def llcs_binary(self, a_bits, b_bits, threads: int = 1, formula_one: bool = False) -> int:
        o = Oracle()
        ar, bw = o.pack_bits_reverse(a_bits), o.pack_bits(b_bits)
        return self.L.ref_llcs_2symbol_smart(ar, len(ar), bw, len(bw), len(a_bits), threads, int(formula_one))


AMATCH_SO = os.path.join(HERE, "libamatch.so")
REF_AMATCH_SO = os.path.join(HERE, "_ref", "libref_amatch.so")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _dup_call(fn, variant, p, t, k):
    p, t = _i32(p), _i32(t)
    cap = max(16, len(t))
    s, e, sc = np.empty(cap, np.int32), np.empty(cap, np.int32), np.empty(cap, np.float64)
    cnt = fn(variant, p, len(p), t, len(t), C.c_double(k), s, e, sc, cap)
    if cnt < 0:
        raise ValueError("bad duplicate-search arguments")
    if cnt > cap:
        raise RuntimeError("more intervals than the buffer holds")
    return [(int(s[i]), int(e[i]), float(sc[i])) for i in range(cnt)]


class AMatchOracle:
    """The restatement of the reference's duplicate searches (amatch.cpp):
    window values by direct LCS DP, the sequential phases as the reference
    runs them."""

    def __init__(self, path: str = AMATCH_SO):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        L.orc_am_windows.argtypes = [_i32p, C.c_int, _i32p, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p]
        L.orc_dup_search.argtypes = [C.c_int, _i32p, C.c_int, _i32p, C.c_int, C.c_double, _i32p, _i32p, _f64p,
                                     C.c_int]

    def windows(self, p, t, w, min_len):
        """(window score, best start, best end, best score) per window."""
        p, t = _i32(p), _i32(t)
        nw = len(t) - w + 1
        out = [np.empty(nw, np.int32) for _ in range(4)]
        if self.L.orc_am_windows(p, len(p), t, len(t), w, min_len, *out):
            raise ValueError("bad window arguments")
        return tuple(out)

    def dup_search(self, p, t, k, variant):
        return _dup_call(self.L.orc_dup_search, variant, p, t, k)


class RefAMatch:
    """The reference's own duplicate searches, compiled in place."""

    def __init__(self, path: str = REF_AMATCH_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.L = C.CDLL(path)
        self.L.ref_dup_search.argtypes = [C.c_int, _i32p, C.c_int, _i32p, C.c_int, C.c_double, _i32p, _i32p, _f64p,
                                          C.c_int]
        self.L.ref_semi_local_H.argtypes = [_i32p, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _i64p]
        self.L.ref_semi_local_H.restype = C.c_double

    @staticmethod
    def available(path: str = REF_AMATCH_SO) -> bool:
        return os.path.exists(path)

    def dup_search(self, p, t, k, variant):
        return _dup_call(self.L.ref_dup_search, variant, p, t, k)

    def semi_local_H(self, perm, m, qi, qj):
        """(H(i, j) for every (qi[k], qj[k]) by the reference's O(m+n)
        dominance scan, seconds spent in the queries alone)."""
        perm, qi, qj = _i32(perm), _i32(qi), _i32(qj)
        out = np.empty(len(qi), np.int64)
        sec = self.L.ref_semi_local_H(perm, len(perm), m, len(qi), qi, qj, out)
        return out, sec
