"""The allocation-free entry point, sl_semi_local_comb_noalloc, through the
C ABI: device buffers in and out (allocated here, by the caller, with the
same HIP runtime the engine links), results bit-exact against the existing
entry point sl_semi_local_comb on the same strings.  Two calls in a row, the
second smaller, so a call that read the previous call's scratch contents
would show."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class DeviceMem:
    """hipMalloc'd buffers through the engine's own HIP runtime."""

    def __init__(self):
        from semilocal_amd import _lib

        _lib.lib()                               # loads the engine and, with it, libamdhip64
        self.hip = C.CDLL("libamdhip64.so.7")    # already loaded: the same runtime instance
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.live = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(int(nbytes), 4)) == 0
        self.live.append(p.value)
        return p.value

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        if arr.nbytes:
            assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def garbage(self, count):
        p = self.alloc(4 * count)
        assert self.hip.hipMemset(p, 0xA5, max(4 * count, 4)) == 0   # every entry must be overwritten
        return p

    def download(self, p, count):
        out = np.empty(count, np.int32)
        if count:
            assert self.hip.hipMemcpy(out.ctypes.data, p, 4 * count, D2H) == 0
        return out

    def close(self):
        for p in self.live:
            self.hip.hipFree(p)
        self.live = []


def _noalloc(dev, gpu, a, b):
    from semilocal_amd._lib import check, lib

    m, n = len(a), len(b)
    width = a.dtype.itemsize
    da, db = dev.upload(a), dev.upload(b)
    rc, cr = dev.garbage(m + n), dev.garbage(m + n)
    check(lib().sl_semi_local_comb_noalloc(da if m else None, m, db if n else None, n, width, rc, cr, None))
    return dev.download(rc, m + n), dev.download(cr, m + n)


def _want(gpu, a, b):
    return gpu.semi_local_comb(a, b, with_col_to_row=True)


def test_noalloc_equals_the_allocating_entry_point_on_two_calls(gpu):
    g = np.random.Generator(np.random.PCG64(2718))
    dev = DeviceMem()
    try:
        # first call: m > n (the reverse mapping), alphabet 4, several passes
        a = g.integers(0, 4, 5000).astype(np.uint8)
        b = g.integers(0, 4, 3100).astype(np.uint8)
        # second call, smaller: m < n, alphabet 20
        c = g.integers(0, 20, 700).astype(np.uint8)
        d = g.integers(0, 20, 1300).astype(np.uint8)
        for x, y in ((a, b), (c, d)):
            rc, cr = _noalloc(dev, gpu, x, y)
            want_rc, want_cr = _want(gpu, x, y)
            assert np.array_equal(rc, want_rc)
            assert np.array_equal(cr, want_cr)
    finally:
        dev.close()


def test_noalloc_uint16_symbols_edge_shapes_and_config1_pair(gpu, oracle):
    from semilocal_amd import workloads as W
    from golden_io import configs

    g = np.random.Generator(np.random.PCG64(31415))
    dev = DeviceMem()
    try:
        # 16-bit symbols, compared as given (values above 255)
        a = (g.integers(0, 300, 900) * 211).astype(np.uint16)
        b = (g.integers(0, 300, 1700) * 211).astype(np.uint16)
        rc, cr = _noalloc(dev, gpu, a, b)
        assert np.array_equal(rc, _want(gpu, a, b)[0]) and np.array_equal(cr, _want(gpu, a, b)[1])
        assert np.array_equal(rc, oracle.comb(a.astype(np.int32), b.astype(np.int32)))
        # one row, one column, an empty string
        for m, n in ((1, 513), (513, 1), (0, 9), (9, 0), (64, 64), (257, 255)):
            x = g.integers(0, 2, m).astype(np.uint8)
            y = g.integers(0, 2, n).astype(np.uint8)
            rc, cr = _noalloc(dev, gpu, x, y)
            want_rc, want_cr = _want(gpu, x, y)
            assert np.array_equal(rc, want_rc) and np.array_equal(cr, want_cr), (m, n)
        # the plain 10,000 x 10,000 pair against the reference's fixture
        a, b = W.plain_pair()
        rc, _ = _noalloc(dev, gpu, a, b)
        assert W.sha256(rc) == configs()["plain_pair_10k"]["perm_sha256"]
    finally:
        dev.close()


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    from semilocal_amd._lib import SemiLocalError, check, lib

    dev = DeviceMem()
    try:
        p = dev.alloc(64)
        too_long = lib().sl_semi_local_comb_noalloc(p, 1048577, p, 10, 1, p, None, None)
        assert too_long == 1   # SL_ERR_ARG, before anything is launched
        assert b"scratch" in lib().sl_last_error()
        with pytest.raises(SemiLocalError):
            check(lib().sl_semi_local_comb_noalloc(p, 10, p, 10, 4, p, None, None))   # int32 symbols: not here
    finally:
        dev.close()
