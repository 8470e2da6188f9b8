"""ctypes binding of libsemilocal_hip.so (the C ABI in include/semilocal_hip.h).

The library is required: importing this module on a machine where it has not
been built raises immediately -- there is no CPU fallback anywhere in the
product path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SEMILOCAL_LIB") or os.path.join(PKG, "libsemilocal_hip.so")

SL_OK = 0
SL_PENDING = 6
SL_BATCH_WITH_COL_TO_ROW = 1
SL_QUERY_H = 0
SL_QUERY_STRING_SUBSTRING = 1
SL_QUERY_SUBSTRING_STRING = 2

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class SemiLocalError(RuntimeError):
    """A non-zero status from the C ABI; .code holds it."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"semilocal_hip error {code}: {msg}")
        self.code = code


_lib = None
_lock = threading.Lock()

# Exported entry points and their argtypes/restype (checked by the CPU tests).
SIGNATURES = {
    "sl_version": ([], C.c_int),
    "sl_last_error": ([], C.c_char_p),
    "sl_init": ([C.c_int], C.c_int),
    "sl_device_info": ([C.c_char_p, C.c_int], C.c_int),
    "sl_probe_copy": ([C.c_int64, C.POINTER(C.c_double)], C.c_int),
    "sl_semi_local_comb": ([_vp, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "sl_semi_local_comb_noalloc": ([_vp, C.c_int, _vp, C.c_int, C.c_int, _vp, _vp, _vp], C.c_int),
    "sl_llcs_from_kernel": ([_i32p, C.c_int, C.c_int], C.c_int),
    "sl_batch_create": ([C.POINTER(_vp), C.c_int, _vp, _i64p, _i32p, _vp, _i64p, _i32p, C.c_int, C.c_int, C.c_int],
                        C.c_int),
    "sl_batch_run": ([_vp], C.c_int),
    "sl_batch_run_per_pair": ([_vp], C.c_int),
    "sl_batch_sync": ([_vp], C.c_int),
    "sl_batch_download": ([_vp, _vp, _vp], C.c_int),
    "sl_batch_time": ([_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "sl_batch_set_tiles": ([_vp, C.c_int, C.c_int], C.c_int),
    "sl_batch_get_tiles": ([_vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "sl_batch_autotune": ([_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "sl_batch_set_schedule": ([_vp, C.c_int], C.c_int),
    "sl_batch_set_critical_fraction": ([_vp, C.c_double], C.c_int),
    "sl_batch_get_critical_fraction": ([_vp, C.POINTER(C.c_double)], C.c_int),
    "sl_batch_pair_tiles": ([_vp, _vp, _vp], C.c_int),
    "sl_batch_get_schedule": ([_vp, C.POINTER(C.c_int)], C.c_int),
    "sl_batch_stats": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "sl_batch_cell_modes": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "sl_plan_queue": ([C.c_int, _i32p, _i32p, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp,
                       C.POINTER(C.c_int64)], C.c_int),
    "sl_braid_index_create": ([C.POINTER(_vp), _i32p, C.c_int, C.c_int], C.c_int),
    "sl_braid_index_query": ([_vp, C.c_int, _i32p, _i32p, _i32p, np.ctypeslib.ndpointer(dtype=np.int64,
                                                                                       flags="C_CONTIGUOUS")],
                             C.c_int),
    "sl_braid_index_stats": ([_vp, C.POINTER(C.c_int), C.POINTER(C.c_int64)], C.c_int),
    "sl_braid_index_destroy": ([_vp], None),
    "sl_llcs_hyyro": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "sl_llcs_hyyro_batch": ([C.c_int, _vp, _i64p, _i32p, _vp, _i64p, _i32p, C.c_int, _i32p], C.c_int),
    "sl_hyyro_create": ([C.POINTER(_vp), C.c_int, _vp, _i64p, _i32p, _vp, _i64p, _i32p, C.c_int], C.c_int),
    "sl_hyyro_run": ([_vp], C.c_int),
    "sl_hyyro_result": ([_vp, _i32p], C.c_int),
    "sl_hyyro_time": ([_vp, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "sl_hyyro_stats": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)], C.c_int),
    "sl_hyyro_buckets": ([_vp, _i32p, _i32p, _i32p, _i32p], C.c_int),
    "sl_hyyro_set_words_per_lane": ([_vp, _i32p], C.c_int),
    "sl_hyyro_autotune": ([_vp, C.c_int, _i32p], C.c_int),
    "sl_hyyro_destroy": ([_vp], None),
    "sl_amatch_create": ([C.POINTER(_vp), _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "sl_amatch_stats": ([_vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int64)], C.c_int),
    "sl_amatch_windows": ([_vp, C.c_int, C.c_int, _i32p, _i32p, _i32p, _i32p], C.c_int),
    "sl_amatch_time": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)],
                       C.c_int),
    "sl_duplicate_search": ([_vp, C.c_double, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "sl_duplicate_results": ([_vp, _vp, _vp, _vp], C.c_int),
    "sl_amatch_last_times": ([_vp, C.POINTER(C.c_double), C.POINTER(C.c_double)], C.c_int),
    "sl_amatch_destroy": ([_vp], None),
    "sl_batch_destroy": ([_vp], None),
    "sl_llcs_bitparallel": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int)], C.c_int),
    "sl_bitpar_create": ([C.POINTER(_vp), _vp, C.c_int, _vp, C.c_int, C.c_int], C.c_int),
    "sl_bitpar_run": ([_vp], C.c_int),
    "sl_bitpar_result": ([_vp, C.POINTER(C.c_int)], C.c_int),
    "sl_bitpar_time": ([_vp, C.c_int, C.c_int, C.POINTER(C.c_double)], C.c_int),
    "sl_bitpar_destroy": ([_vp], None),
    "sl_bitpar_set_form": ([_vp, C.c_int], C.c_int),
    "sl_bitpar_get_form": ([_vp, C.POINTER(C.c_int)], C.c_int),
    "sl_comb_batch": ([C.c_int, _vp, _i64p, _i32p, _vp, _i64p, _i32p, C.c_int, _vp, _vp], C.c_int),
    "sl_braid_multiply": ([_vp, _vp, C.c_int, _vp], C.c_int),
    "sl_braid_multiply_batch": ([C.c_int, _vp, _vp, _vp, _vp], C.c_int),
    "sl_braid_multiply_time": ([C.c_int, _vp, _vp, _vp, C.c_int, C.c_int, C.POINTER(C.c_double),
                                C.POINTER(C.c_int64)], C.c_int),
    "sl_staggered_multiply": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _vp], C.c_int),
    "sl_hybrid_iterative": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp], C.c_int),
    "sl_hybrid_create": ([C.POINTER(_vp), _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
    "sl_hybrid_run": ([_vp], C.c_int),
    "sl_hybrid_download": ([_vp, _vp], C.c_int),
    "sl_hybrid_time": ([_vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                        C.POINTER(C.c_double)], C.c_int),
    "sl_hybrid_stats": ([_vp, C.POINTER(C.c_int64), C.POINTER(C.c_int)], C.c_int),
    "sl_hybrid_destroy": ([_vp], None),
    "sl_hybrid_set_graph": ([_vp, C.c_int], C.c_int),
    "sl_hybrid_autotune": ([_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
    "sl_hybrid_autotune_blocks": ([_vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, _i32p, _i32p, C.c_int,
                                   C.POINTER(C.c_int32), C.POINTER(C.c_int32), _f64p], C.c_int),
    "sl_hybrid_bucket_blocks": ([C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int32)], C.c_int),
    "sl_server_create": ([C.POINTER(_vp), C.c_int64, C.c_int, C.c_int, C.c_int64], C.c_int),
    "sl_server_submit_compare": ([_vp, _vp, C.c_int, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)],
                                 C.c_int),
    "sl_server_submit_query": ([_vp, C.c_int64, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64)], C.c_int),
    "sl_server_result": ([_vp, C.c_int64, C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_double)], C.c_int),
    "sl_server_download_braid": ([_vp, C.c_int64, _vp], C.c_int),
    "sl_server_release": ([_vp, C.c_int64], C.c_int),
    "sl_server_stats": ([_vp, _vp], C.c_int),
    "sl_server_set_index": ([_vp, C.c_int64], C.c_int),
    "sl_server_destroy": ([_vp], None),
}


def lib():
    """Load the engine library (once).  Raises if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(
                    f"{LIB_PATH} is missing: build it with `python -m semilocal_amd.build` "
                    "(there is no CPU fallback)")
            L = C.CDLL(LIB_PATH)
            for name, (argtypes, restype) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.argtypes = argtypes
                fn.restype = restype
            _lib = L
        return _lib


def check(status: int) -> None:
    if status != SL_OK:
        raise SemiLocalError(status, lib().sl_last_error().decode(errors="replace"))


def _ptr(x: np.ndarray):
    return x.ctypes.data_as(_vp)


def as_symbols(x) -> tuple[np.ndarray, int]:
    """Contiguous symbol array and its element width (1, 2 or 4 bytes).

    Symbols must be integers (or booleans); wider integer types are accepted
    when every value fits int32, so that no two distinct symbols collapse to
    one.  Anything else raises ValueError."""
    arr = np.asarray(x)
    if arr.dtype == np.bool_:
        arr = arr.astype(np.uint8)
    if arr.dtype == np.uint8:
        return np.ascontiguousarray(arr), 1
    if arr.dtype == np.uint16:
        return np.ascontiguousarray(arr), 2
    if arr.dtype.kind not in "iu":
        raise ValueError(f"symbols must be an integer array, not {arr.dtype}")
    if arr.dtype.itemsize > 4 or arr.dtype == np.uint32:
        if arr.size and (int(arr.min()) < -(1 << 31) or int(arr.max()) >= (1 << 31)):
            raise ValueError(f"symbol values outside the int32 range ({arr.dtype})")
    return np.ascontiguousarray(arr, dtype=np.int32), 4


def device_info() -> dict:
    buf = C.create_string_buffer(512)
    check(lib().sl_device_info(buf, 512))
    name, cus, clk, hbm, arch = buf.value.decode().split(";")
    return {"name": name, "compute_units": int(cus), "clock_khz": int(clk), "hbm_bytes": int(hbm), "arch": arch}


def semi_local_comb(a, b, with_col_to_row: bool = False):
    """Kernel permutation (row->col, int32[m+n]) of strings a (rows), b (cols)."""
    a_arr, wa = as_symbols(a)
    b_arr, wb = as_symbols(b)
    if wa != wb:
        a_arr, wa = np.ascontiguousarray(a_arr, dtype=np.int32), 4
        b_arr, wb = np.ascontiguousarray(b_arr, dtype=np.int32), 4
    m, n = len(a_arr), len(b_arr)
    rc = np.empty(max(m + n, 1), np.int32)
    cr = np.empty(max(m + n, 1), np.int32) if with_col_to_row else None
    check(lib().sl_semi_local_comb(_ptr(a_arr) if m else None, m, _ptr(b_arr) if n else None, n, wa,
                                   _ptr(rc), _ptr(cr) if cr is not None else None))
    rc = rc[:m + n]
    return (rc, cr[:m + n]) if with_col_to_row else rc


def _concat(pairs):
    widths = set()
    arrs = []
    for a, b in pairs:
        a_arr, wa = as_symbols(a)
        b_arr, wb = as_symbols(b)
        widths.update((wa, wb))
        arrs.append((a_arr, b_arr))
    width = 1 if widths <= {1} else (2 if widths <= {1, 2} else 4)
    dt = {1: np.uint8, 2: np.uint16, 4: np.int32}[width]
    a_len = np.array([len(a) for a, _ in arrs], np.int32)
    b_len = np.array([len(b) for _, b in arrs], np.int32)
    a_cat = np.concatenate([a.astype(dt, copy=False) for a, _ in arrs]) if arrs else np.zeros(0, dt)
    b_cat = np.concatenate([b.astype(dt, copy=False) for _, b in arrs]) if arrs else np.zeros(0, dt)
    a_off = np.zeros(len(arrs), np.int64)
    b_off = np.zeros(len(arrs), np.int64)
    if len(arrs) > 1:
        a_off[1:] = np.cumsum(a_len[:-1], dtype=np.int64)
        b_off[1:] = np.cumsum(b_len[:-1], dtype=np.int64)
    if a_cat.size == 0:
        a_cat = np.zeros(1, dt)
    if b_cat.size == 0:
        b_cat = np.zeros(1, dt)
    return a_cat, a_off, a_len, b_cat, b_off, b_len, width


def _split(flat, a_len, b_len):
    out, s = [], 0
    for m, n in zip(a_len.tolist(), b_len.tolist()):
        out.append(flat[s:s + m + n].copy())
        s += m + n
    return out


def plan_queue(a_lens, b_lens, rows_per_lane: int, waves_per_cu: int = 8, cus: int = 256, schedule: int = 2):
    """The (pair, pass) tickets a batch of pairs with these lengths gets, in
    launch order (host-only planning: sl_plan_queue, no device needed)."""
    a_len = np.ascontiguousarray(a_lens, np.int32)
    b_len = np.ascontiguousarray(b_lens, np.int32)
    n = C.c_int64()
    check(lib().sl_plan_queue(len(a_len), a_len, b_len, rows_per_lane, waves_per_cu, cus, schedule, None, None,
                              C.byref(n)))
    pair = np.empty(max(n.value, 1), np.int32)
    pas = np.empty(max(n.value, 1), np.int32)
    check(lib().sl_plan_queue(len(a_len), a_len, b_len, rows_per_lane, waves_per_cu, cus, schedule,
                              pair.ctypes.data, pas.ctypes.data, C.byref(n)))
    return pair[:n.value], pas[:n.value]


class BraidIndex:
    """Wavelet-matrix index over a kernel (sl_braid_index_*): O(log(m+n))
    H(i, j) / string-substring / substring-string queries."""

    H, STRING_SUBSTRING, SUBSTRING_STRING = 0, 1, 2

    def __init__(self, perm, m: int, n: int):
        p = np.ascontiguousarray(perm, np.int32)
        if len(p) != m + n:
            raise ValueError("perm must have m + n entries")
        h = _vp()
        check(lib().sl_braid_index_create(C.byref(h), p if len(p) else np.zeros(1, np.int32), m, n))
        self.h = h

    def query(self, kinds, qi, qj) -> np.ndarray:
        k = np.ascontiguousarray(kinds, np.int32)
        i = np.ascontiguousarray(qi, np.int32)
        j = np.ascontiguousarray(qj, np.int32)
        out = np.empty(max(len(k), 1), np.int64)
        check(lib().sl_braid_index_query(self.h, len(k), k, i, j, out))
        return out[:len(k)]

    def stats(self):
        lv, b = C.c_int(), C.c_int64()
        check(lib().sl_braid_index_stats(self.h, C.byref(lv), C.byref(b)))
        return {"levels": lv.value, "bytes": b.value}

    def close(self):
        if getattr(self, "h", None):
            lib().sl_braid_index_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def llcs_hyyro(a, b) -> int:
    """LLCS of one pair by Hyyro's bit-vector recurrence on the device."""
    a_arr, wa = as_symbols(a)
    b_arr, wb = as_symbols(b)
    if wa != wb:
        a_arr, wa = np.ascontiguousarray(a_arr, dtype=np.int32), 4
        b_arr, wb = np.ascontiguousarray(b_arr, dtype=np.int32), 4
    out = C.c_int()
    check(lib().sl_llcs_hyyro(_ptr(a_arr) if len(a_arr) else None, len(a_arr), _ptr(b_arr) if len(b_arr) else None,
                              len(b_arr), wa, C.byref(out)))
    return out.value


class HyyroBatch:
    """Many pairs' LLCS by Hyyro's recurrence (sl_hyyro_*): symbols recoded
    and match tables built once, then run / time."""

    def __init__(self, pairs):
        self.a_cat, self.a_off, self.a_len, self.b_cat, self.b_off, self.b_len, w = _concat(pairs)
        self.npairs = len(pairs)
        h = _vp()
        check(lib().sl_hyyro_create(C.byref(h), self.npairs, _ptr(self.a_cat), self.a_off, self.a_len,
                                    _ptr(self.b_cat), self.b_off, self.b_len, w))
        self.h = h

    def run(self) -> np.ndarray:
        check(lib().sl_hyyro_run(self.h))
        out = np.empty(max(self.npairs, 1), np.int32)
        check(lib().sl_hyyro_result(self.h, out))
        return out[:self.npairs]

    def time(self, warmup: int = 1, iters: int = 5) -> float:
        ms = C.c_double()
        check(lib().sl_hyyro_time(self.h, warmup, iters, C.byref(ms)))
        return ms.value

    def stats(self):
        wr, tw = C.c_int64(), C.c_int64()
        check(lib().sl_hyyro_stats(self.h, C.byref(wr), C.byref(tw)))
        return {"word_rows": wr.value, "table_words": tw.value}

    HYYRO_BUCKETS = 5

    def buckets(self) -> list:
        """Per length bucket (packed width up to max_words 64-bit words):
        pairs, 32-bit words per lane of the kernel it runs (1, 2 or 4: 32, 64
        or 128 bits), and waves per pair or segment."""
        n = self.HYYRO_BUCKETS
        mw, pr, wpl, wv = (np.zeros(n, np.int32) for _ in range(4))
        check(lib().sl_hyyro_buckets(self.h, mw, pr, wpl, wv))
        return [{"max_words": int(mw[i]), "pairs": int(pr[i]), "words_per_lane": int(wpl[i]),
                 "waves": int(wv[i])} for i in range(n)]

    def set_words_per_lane(self, wpl) -> None:
        """Forces the 32-bit words per lane per bucket (0 keeps a bucket's)."""
        arr = np.ascontiguousarray(wpl, np.int32)
        if arr.shape != (self.HYYRO_BUCKETS,):
            raise ValueError("one words-per-lane value per bucket")
        check(lib().sl_hyyro_set_words_per_lane(self.h, arr))

    def autotune(self, rounds: int = 3) -> list:
        out = np.zeros(self.HYYRO_BUCKETS, np.int32)
        check(lib().sl_hyyro_autotune(self.h, rounds, out))
        return [int(x) for x in out]

    def close(self):
        if getattr(self, "h", None):
            lib().sl_hyyro_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class AMatch:
    """Approximate matching of a pattern against a text (sl_amatch_*): the
    pattern is combed against overlapping chunks of the text in one batched
    launch and the chunk braids stay in HBM; window queries and the
    reference's duplicate searches run on them."""

    SEMI_DENSE, SEMI_SPARSE = 0, 1

    def __init__(self, pattern, text, max_window: int, chunk: int = 0):
        p_arr, wp = as_symbols(pattern)
        t_arr, wt = as_symbols(text)
        if wp != wt:
            p_arr, wp = np.ascontiguousarray(p_arr, dtype=np.int32), 4
            t_arr, wt = np.ascontiguousarray(t_arr, dtype=np.int32), 4
        self.m, self.n = len(p_arr), len(t_arr)
        h = _vp()
        check(lib().sl_amatch_create(C.byref(h), _ptr(p_arr) if self.m else None, self.m,
                                     _ptr(t_arr) if self.n else None, self.n, wp, max_window, chunk))
        self.h = h

    def stats(self):
        c, b, cells = C.c_int(), C.c_int(), C.c_int64()
        check(lib().sl_amatch_stats(self.h, C.byref(c), C.byref(b), C.byref(cells)))
        return {"chunks": c.value, "chunk_windows": b.value, "cells": cells.value}

    def windows(self, w: int, min_len: int):
        """Per window [i, i+w): its score and its best substring of length
        >= min_len -- (window_score, best_start, best_end, best_score)."""
        nw = max(self.n - w + 1, 1)
        out = [np.empty(nw, np.int32) for _ in range(4)]
        check(lib().sl_amatch_windows(self.h, w, min_len, *out))
        return tuple(o[:self.n - w + 1] for o in out)

    def time(self, w: int, min_len: int, warmup: int = 1, iters: int = 5):
        comb, win = C.c_double(), C.c_double()
        check(lib().sl_amatch_time(self.h, w, min_len, warmup, iters, C.byref(comb), C.byref(win)))
        return comb.value, win.value

    def duplicate_search(self, k: float, variant: int = 0):
        """InteractiveDuplicateSearchSemiDense (0) / SemiSparse (1): list of
        (start, end, score)."""
        cnt = C.c_int()
        check(lib().sl_duplicate_search(self.h, float(k), variant, C.byref(cnt)))
        n = cnt.value
        s, e, sc = np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.int32), np.empty(max(n, 1), np.float64)
        check(lib().sl_duplicate_results(self.h, s.ctypes.data, e.ctypes.data, sc.ctypes.data))
        return [(int(s[i]), int(e[i]), float(sc[i])) for i in range(n)]

    def last_times(self):
        """(device ms, host ms) of the last duplicate_search: window kernel +
        copy to host, then the sequential phases."""
        d, h = C.c_double(), C.c_double()
        check(lib().sl_amatch_last_times(self.h, C.byref(d), C.byref(h)))
        return d.value, h.value

    def close(self):
        if getattr(self, "h", None):
            lib().sl_amatch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def duplicate_search(pattern, text, k: float, variant: int = 0):
    """One-shot duplicate search (the reference's
    InteractiveDuplicateSearchSemiDense / SemiSparse ::find)."""
    am = AMatch(pattern, text, int(len(pattern) / k))
    try:
        return am.duplicate_search(k, variant)
    finally:
        am.close()


class Batch:
    """A ragged batch resident in HBM (sl_batch_* entry points)."""

    def __init__(self, pairs, with_col_to_row: bool = False, rows_per_lane: int = 0):
        self.a_cat, self.a_off, self.a_len, self.b_cat, self.b_off, self.b_len, w = _concat(pairs)
        self.with_cr = with_col_to_row
        h = _vp()
        check(lib().sl_batch_create(C.byref(h), len(self.a_len), _ptr(self.a_cat), self.a_off, self.a_len,
                                    _ptr(self.b_cat), self.b_off, self.b_len, w,
                                    SL_BATCH_WITH_COL_TO_ROW if with_col_to_row else 0, rows_per_lane))
        self.h = h
        self.total = int((self.a_len.astype(np.int64) + self.b_len).sum())

    def run(self, per_pair: bool = False):
        check(lib().sl_batch_run_per_pair(self.h) if per_pair else lib().sl_batch_run(self.h))
        check(lib().sl_batch_sync(self.h))
        return self

    def results(self):
        rc = np.empty(max(self.total, 1), np.int32)
        cr = np.empty(max(self.total, 1), np.int32) if self.with_cr else None
        check(lib().sl_batch_download(self.h, _ptr(rc), _ptr(cr) if cr is not None else None))
        rcs = _split(rc, self.a_len, self.b_len)
        if cr is None:
            return rcs
        return rcs, _split(cr, self.a_len, self.b_len)

    def time(self, per_pair: bool = False, warmup: int = 1, iters: int = 3):
        ms = C.c_double()
        kms = C.c_double()
        check(lib().sl_batch_time(self.h, 1 if per_pair else 0, warmup, iters, C.byref(ms), C.byref(kms)))
        return ms.value, kms.value

    def set_tiles(self, rows_per_lane: int, waves_per_cu: int = 16):
        """rows_per_lane 4, 8 or 16 for every pair, or 0 for a height per pair."""
        check(lib().sl_batch_set_tiles(self.h, rows_per_lane, waves_per_cu))

    def tiles(self):
        r, w = C.c_int(), C.c_int()
        check(lib().sl_batch_get_tiles(self.h, C.byref(r), C.byref(w)))
        return r.value, w.value

    def set_schedule(self, schedule: int):
        """Queue order of the batched launch: 0 pair by pair (largest first), 1-8 simulated (lag 128 s)."""
        check(lib().sl_batch_set_schedule(self.h, int(schedule)))

    def schedule(self) -> int:
        out = C.c_int()
        check(lib().sl_batch_get_schedule(self.h, C.byref(out)))
        return out.value

    def set_critical_fraction(self, fraction: float):
        """Per-pair tiles: pairs whose modelled chain exceeds `fraction` x the
        batch's modelled time become critical chains (0 = none)."""
        check(lib().sl_batch_set_critical_fraction(self.h, float(fraction)))

    def critical_fraction(self) -> float:
        out = C.c_double()
        check(lib().sl_batch_get_critical_fraction(self.h, C.byref(out)))
        return out.value

    def pair_tiles(self):
        """(rows_per_lane, priority) per pair, caller order (0 rows for an empty pair)."""
        n = len(self.a_len)
        r = np.zeros(max(n, 1), np.int32)
        p = np.zeros(max(n, 1), np.int32)
        check(lib().sl_batch_pair_tiles(self.h, _ptr(r), _ptr(p)))
        return r[:n], p[:n]

    def autotune(self, iters: int = 1):
        r, w = C.c_int(), C.c_int()
        check(lib().sl_batch_autotune(self.h, iters, C.byref(r), C.byref(w)))
        return r.value, w.value

    def stats(self):
        c, t, p = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().sl_batch_stats(self.h, C.byref(c), C.byref(t), C.byref(p)))
        return {"cells": c.value, "tiles": t.value, "pairs": p.value}

    def cell_modes(self):
        """Pairs per symbol test of the combing cell at the current tile height."""
        c, o, t = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().sl_batch_cell_modes(self.h, C.byref(c), C.byref(o), C.byref(t)))
        return {"code": c.value, "onehot": o.value, "table": t.value}

    def close(self):
        if getattr(self, "h", None):
            lib().sl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def comb_batch(pairs, with_col_to_row: bool = False):
    """One-shot ragged batch: list of row->col arrays (and col->row)."""
    b = Batch(pairs, with_col_to_row)
    try:
        return b.run().results()
    finally:
        b.close()


def pack_bits_reverse(bits) -> np.ndarray:
    """a reversed, MSB-first uint64 words (the reference's encode_reverse layout)."""
    bits = np.asarray(bits, dtype=np.uint8) & 1
    words = (len(bits) + 63) // 64
    padded = np.zeros(words * 64, np.uint8)
    padded[:len(bits)] = bits
    # word w holds bits[64w .. 64w+63] with bits[64w] at bit 63; stored at index words-1-w
    packed = np.packbits(padded.reshape(words, 64), axis=1, bitorder="big").view(">u8").astype(np.uint64).ravel()
    return np.ascontiguousarray(packed[::-1])


def pack_bits(bits) -> np.ndarray:
    """LSB-first uint64 words (the reference's encode layout)."""
    bits = np.asarray(bits, dtype=np.uint8) & 1
    words = (len(bits) + 63) // 64
    padded = np.zeros(words * 64, np.uint8)
    padded[:len(bits)] = bits
    return np.ascontiguousarray(np.packbits(padded.reshape(words, 64), axis=1, bitorder="little").view("<u8").astype(np.uint64).ravel())


BITPAR_CARRY, BITPAR_COMBING = 0, 1   # SL_BITPAR_CARRY, SL_BITPAR_COMBING


class BitParallel:
    """Resident bit-parallel LLCS problem (sl_bitpar_* entry points).

    `form`: BITPAR_CARRY (Hyyro's carry recurrence with the packed bits as
    masks; the default where it fits) or BITPAR_COMBING (sliding-diagonal
    combing).  Both return the same score."""

    def __init__(self, a_bits, b_bits, form=None):
        self.ar = pack_bits_reverse(a_bits)
        self.bw = pack_bits(b_bits)
        h = _vp()
        check(lib().sl_bitpar_create(C.byref(h), _ptr(self.ar), len(self.ar), _ptr(self.bw), len(self.bw),
                                     len(a_bits)))
        self.h = h
        if form is not None:
            self.form = form

    @property
    def form(self) -> int:
        out = C.c_int()
        check(lib().sl_bitpar_get_form(self.h, C.byref(out)))
        return out.value

    @form.setter
    def form(self, value: int):
        check(lib().sl_bitpar_set_form(self.h, int(value)))

    def llcs(self) -> int:
        out = C.c_int()
        check(lib().sl_bitpar_run(self.h))
        check(lib().sl_bitpar_result(self.h, C.byref(out)))
        return out.value

    def time(self, warmup: int = 1, iters: int = 3) -> float:
        ms = C.c_double()
        check(lib().sl_bitpar_time(self.h, warmup, iters, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            lib().sl_bitpar_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def llcs_bitparallel(a_bits, b_bits) -> int:
    """LLCS of two binary strings by bit-parallel combing (one-shot)."""
    ar, bw = pack_bits_reverse(a_bits), pack_bits(b_bits)
    out = C.c_int()
    check(lib().sl_llcs_bitparallel(_ptr(ar), len(ar), _ptr(bw), len(bw), len(a_bits), 0, 0, C.byref(out)))
    return out.value


# ---------------------------------------------------------------------------
# Braid multiplication (kernel (c)), staggered glue and the hybrid scheduler.
# ---------------------------------------------------------------------------
def _perm(x) -> np.ndarray:
    """int32 copy of a permutation (row -> col); integer input only, values
    inside the int32 range (the engine rejects anything not a permutation)."""
    arr = np.asarray(x)
    if arr.dtype.kind not in "iu":
        raise ValueError(f"a permutation must be an integer array, not {arr.dtype}")
    if arr.dtype.itemsize > 4 or arr.dtype == np.uint32:
        if arr.size and (int(arr.min()) < -(1 << 31) or int(arr.max()) >= (1 << 31)):
            raise ValueError(f"permutation values outside the int32 range ({arr.dtype})")
    return np.ascontiguousarray(arr, dtype=np.int32)


def braid_multiply(p, q) -> np.ndarray:
    """R = P (.) Q for full permutations p, q (row -> col) of equal size."""
    p, q = _perm(p), _perm(q)
    if p.shape != q.shape or p.ndim != 1:
        raise ValueError("p and q must be 1-D permutations of the same size")
    n = len(p)
    r = np.empty(max(n, 1), np.int32)
    check(lib().sl_braid_multiply(_ptr(p), _ptr(q), n, _ptr(r)))
    return r[:n]


def _mult_batch_args(pairs):
    sizes = np.array([len(p) for p, _ in pairs], np.int32)
    for p, q in pairs:
        if len(p) != len(q):
            raise ValueError("p and q must have the same size")
    pc = _perm(np.concatenate([_perm(p) for p, _ in pairs])) if len(pairs) else np.zeros(1, np.int32)
    qc = _perm(np.concatenate([_perm(q) for _, q in pairs])) if len(pairs) else np.zeros(1, np.int32)
    if pc.size == 0:
        pc = np.zeros(1, np.int32)
        qc = np.zeros(1, np.int32)
    return sizes, pc, qc


def braid_multiply_batch(pairs) -> list:
    """Products of a list of (p, q) pairs, all in the same level launches."""
    sizes, pc, qc = _mult_batch_args(pairs)
    out = np.empty(max(int(sizes.sum()), 1), np.int32)
    check(lib().sl_braid_multiply_batch(len(pairs), _ptr(pc), _ptr(qc), _ptr(sizes) if len(pairs) else None,
                                        _ptr(out)))
    res, o = [], 0
    for n in sizes:
        res.append(out[o:o + n].copy())
        o += int(n)
    return res


def braid_multiply_time(pairs, warmup: int = 1, iters: int = 3):
    """(ms per run with inputs resident, sum over levels of elements processed)."""
    sizes, pc, qc = _mult_batch_args(pairs)
    ms, el = C.c_double(), C.c_int64()
    check(lib().sl_braid_multiply_time(len(pairs), _ptr(pc), _ptr(qc), _ptr(sizes), warmup, iters, C.byref(ms),
                                       C.byref(el)))
    return ms.value, el.value


def staggered_multiply(p, q, k: int, row_glue: bool) -> np.ndarray:
    """Glue of two adjacent block braids sharing k strands (steady_ant.h:609-737)."""
    p, q = _perm(p), _perm(q)
    pr, qr = len(p), len(q)
    out = np.empty(max(pr + qr - k, 1), np.int32)
    check(lib().sl_staggered_multiply(_ptr(p) if pr else None, pr, _ptr(q) if qr else None, qr, int(k),
                                      1 if row_glue else 0, _ptr(out)))
    return out[:pr + qr - k]


def _pair_symbols(a, b):
    a_arr, wa = as_symbols(a)
    b_arr, wb = as_symbols(b)
    if wa != wb:
        a_arr, wa = np.ascontiguousarray(a_arr, dtype=np.int32), 4
        b_arr, wb = np.ascontiguousarray(b_arr, dtype=np.int32), 4
    return a_arr, b_arr, wa


class Hybrid:
    """Resident hybrid decomposition (sl_hybrid_* entry points)."""

    def __init__(self, a, b, small_m: int, small_n: int):
        self.a, self.b, w = _pair_symbols(a, b)
        self.m, self.n = len(self.a), len(self.b)
        h = _vp()
        check(lib().sl_hybrid_create(C.byref(h), _ptr(self.a) if self.m else None, self.m,
                                     _ptr(self.b) if self.n else None, self.n, w, small_m, small_n))
        self.h = h

    def run(self) -> np.ndarray:
        check(lib().sl_hybrid_run(self.h))
        out = np.empty(max(self.m + self.n, 1), np.int32)
        check(lib().sl_hybrid_download(self.h, _ptr(out)))
        return out[:self.m + self.n]

    def time(self, warmup: int = 1, iters: int = 3):
        tot, comb, merge = C.c_double(), C.c_double(), C.c_double()
        check(lib().sl_hybrid_time(self.h, warmup, iters, C.byref(tot), C.byref(comb), C.byref(merge)))
        return tot.value, comb.value, merge.value

    def set_graph(self, enable: bool = True):
        """Capture the whole run into a HIP graph and replay it (same results)."""
        check(lib().sl_hybrid_set_graph(self.h, 1 if enable else 0))

    def autotune(self, iters: int = 1):
        """Tune the combing launch's tile shape and queue order (same results)."""
        r, w = C.c_int(), C.c_int()
        check(lib().sl_hybrid_autotune(self.h, iters, C.byref(r), C.byref(w)))
        return r.value, w.value

    def stats(self):
        cells, levels = C.c_int64(), C.c_int()
        check(lib().sl_hybrid_stats(self.h, C.byref(cells), C.byref(levels)))
        return cells.value, levels.value

    def close(self):
        if getattr(self, "h", None):
            lib().sl_hybrid_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def hybrid_iterative(a, b, small_m: int, small_n: int) -> np.ndarray:
    """Kernel permutation of (a, b) by the hybrid block decomposition (one-shot)."""
    a_arr, b_arr, w = _pair_symbols(a, b)
    m, n = len(a_arr), len(b_arr)
    out = np.empty(max(m + n, 1), np.int32)
    check(lib().sl_hybrid_iterative(_ptr(a_arr) if m else None, m, _ptr(b_arr) if n else None, n, w, small_m,
                                    small_n, _ptr(out)))
    return out[:m + n]


def hybrid_autotune_blocks(a, b, grids, iters: int = 2):
    """Times the hybrid of (a, b) at every candidate block grid (sm, sn) and
    returns (best grid, list of ms per grid); the choice is kept for the
    pair's length bucket (see hybrid_bucket_blocks)."""
    a_arr, b_arr, w = _pair_symbols(a, b)
    m, n = len(a_arr), len(b_arr)
    cm = np.ascontiguousarray([g[0] for g in grids], np.int32)
    cn = np.ascontiguousarray([g[1] for g in grids], np.int32)
    ms = np.zeros(len(grids), np.float64)
    bm, bn = C.c_int32(), C.c_int32()
    check(lib().sl_hybrid_autotune_blocks(_ptr(a_arr) if m else None, m, _ptr(b_arr) if n else None, n, w,
                                          len(grids), cm, cn, iters, C.byref(bm), C.byref(bn), ms))
    return (bm.value, bn.value), ms.tolist()


def hybrid_bucket_blocks(m: int, n: int):
    """The block grid kept for the length bucket of (m, n); (0, 0) if none."""
    bm, bn = C.c_int32(), C.c_int32()
    check(lib().sl_hybrid_bucket_blocks(m, n, C.byref(bm), C.byref(bn)))
    return bm.value, bn.value


# ---------------------------------------------------------------------------
# Request loop (serving).
# ---------------------------------------------------------------------------
class ServerStats(C.Structure):
    _fields_ = [("compares_done", C.c_int64), ("queries_done", C.c_int64), ("cells_done", C.c_int64),
                ("launches", C.c_int64), ("deferrals", C.c_int64), ("evictions", C.c_int64),
                ("pool_bytes", C.c_int64), ("pool_in_use", C.c_int64), ("pool_peak", C.c_int64),
                ("resident_braids", C.c_int64), ("busy_seconds", C.c_double),
                ("indexed_braids", C.c_int64), ("index_queries", C.c_int64),
                ("peak_inflight", C.c_int64), ("peak_inflight_compares", C.c_int64)]


class Server:
    """Continuous-batching request loop (sl_server_* entry points)."""

    STRING_SUBSTRING = SL_QUERY_STRING_SUBSTRING
    SUBSTRING_STRING = SL_QUERY_SUBSTRING_STRING
    H = SL_QUERY_H

    def __init__(self, pool_bytes: int = 0, max_inflight: int = 0, max_batch: int = 0, max_batch_cells: int = 0):
        h = _vp()
        check(lib().sl_server_create(C.byref(h), int(pool_bytes), int(max_inflight), int(max_batch),
                                     int(max_batch_cells)))
        self.h = h

    def set_index(self, min_size: int):
        """Braids with m + n >= min_size get a wavelet-matrix index (<= 0: none)."""
        check(lib().sl_server_set_index(self.h, int(min_size)))

    def submit_compare(self, a, b, keep: bool = True) -> int:
        a_arr, b_arr, w = _pair_symbols(a, b)
        rid = C.c_int64()
        check(lib().sl_server_submit_compare(self.h, _ptr(a_arr) if len(a_arr) else None, len(a_arr),
                                             _ptr(b_arr) if len(b_arr) else None, len(b_arr), w, 1 if keep else 0,
                                             C.byref(rid)))
        return rid.value

    def submit_query(self, braid: int, kind: int, i: int, j: int) -> int:
        rid = C.c_int64()
        check(lib().sl_server_submit_query(self.h, int(braid), int(kind), int(i), int(j), C.byref(rid)))
        return rid.value

    def result(self, rid: int, wait: bool = True):
        """(value, latency_ms), or None while pending (wait=False)."""
        val, lat = C.c_int64(), C.c_double()
        st = lib().sl_server_result(self.h, int(rid), 1 if wait else 0, C.byref(val), C.byref(lat))
        if st == SL_PENDING:
            return None
        check(st)
        return val.value, lat.value

    def braid(self, rid: int, size: int) -> np.ndarray:
        out = np.empty(max(size, 1), np.int32)
        check(lib().sl_server_download_braid(self.h, int(rid), _ptr(out)))
        return out[:size]

    def release(self, rid: int):
        check(lib().sl_server_release(self.h, int(rid)))

    def stats(self) -> dict:
        st = ServerStats()
        check(lib().sl_server_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().sl_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def probe_copy(nbytes: int) -> float:
    """One plain device copy of nbytes (read nbytes, write nbytes); ms."""
    ms = C.c_double()
    check(lib().sl_probe_copy(int(nbytes), C.byref(ms)))
    return ms.value
