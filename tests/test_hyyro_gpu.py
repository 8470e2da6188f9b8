"""GPU parity of the Hyyro bit-vector LLCS (sl_llcs_hyyro, sl_hyyro_*):
scores equal plain LCS dynamic programming (oracle) across the launch
classes -- the systolic kernel with 32, 64 and 128 bits per lane, one
workgroup of 1, 2 or 4 waves per pair up to 2,048 words and segments of 4
waves beyond,
its mask table in LDS for small alphabets and in global memory for large
ones -- at word, lane, wave and segment boundaries, for 1-, 2- and 4-byte
symbols; at full size they equal the semi-local kernel's LLCS.  Bar: exact
integers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# packed-string lengths straddling the launch-class and segment boundaries
# (64, 128, 256, 2,048, 8,192 words; segments of 256 and 512 words)
BOUNDARY_N = [1, 63, 64, 65, 4095, 4096, 4097, 8192, 8193, 16384, 16385, 32769, 65536, 65537, 131071, 131073,
              262144, 262145, 524288, 524289, 524353]


@pytest.mark.parametrize("n", BOUNDARY_N)
def test_single_pairs_at_class_boundaries_equal_dp(gpu, oracle, n):
    g = np.random.default_rng(n)
    m = int(min(700, max(1, 4_000_000 // n)))
    a = g.integers(0, 4, m).astype(np.uint8)
    b = g.integers(0, 4, n).astype(np.uint8)
    print(f"hyyro m={m} n={n}", flush=True)
    assert gpu.llcs_hyyro(a, b) == oracle.llcs_dp(a, b)
    assert gpu.llcs_hyyro(b, a) == oracle.llcs_dp(a, b)     # the longer string is the packed one


def test_batch_of_ragged_pairs_equals_dp(gpu, oracle):
    g = np.random.default_rng(3)
    pairs = []
    for _ in range(300):
        m, n = int(g.integers(0, 900)), int(g.integers(0, 5000))
        sigma = int(g.choice([2, 4, 20, 200]))
        pairs.append((g.integers(0, sigma, m).astype(np.uint8), g.integers(0, sigma, n).astype(np.uint8)))
    hb = gpu.HyyroBatch(pairs)
    got = hb.run()
    assert got.tolist() == [oracle.llcs_dp(a, b) for a, b in pairs]
    again = hb.run()                                          # re-runs reset their state
    assert np.array_equal(got, again)
    assert hb.time(warmup=0, iters=2) > 0 and hb.stats()["word_rows"] > 0
    hb.close()


@pytest.mark.parametrize("dtype,sigma", [(np.uint16, 256), (np.int32, 37), (np.int32, 2)])
def test_wide_symbols_equal_dp(gpu, oracle, dtype, sigma):
    g = np.random.default_rng(sigma)
    vals = g.choice(np.arange(-100_000, 100_000), sigma, replace=False) if dtype == np.int32 else np.arange(sigma)
    a = vals[g.integers(0, sigma, 1500)].astype(dtype)
    b = vals[g.integers(0, sigma, 9000)].astype(dtype)
    assert gpu.llcs_hyyro(a, b) == oracle.llcs_dp(a.astype(np.int64).astype(np.int32), b.astype(np.int32))


# the segmented form's mask table sits in LDS while (sigma + 1) rows of a
# segment fit 64 KiB: sigma <= 63, 31 and 15 at 32, 64 and 128 bits per
# lane; larger alphabets gather from the global table.  Pairs of 3 to 9
# segments, both sides of every limit, with the segment hand-off in play.
@pytest.mark.parametrize("sigma", [15, 16, 31, 32, 63, 64, 200])
def test_segmented_pairs_both_table_forms_equal_dp(gpu, oracle, sigma):
    g = np.random.default_rng(100 + sigma)
    a = g.integers(0, sigma, 400).astype(np.uint8)
    b = g.integers(0, sigma, 64 * 1024 + 37).astype(np.uint8)   # 1,025 words
    b[:sigma] = np.arange(sigma, dtype=np.uint8)               # every symbol occurs in the packed string
    want = oracle.llcs_dp(a, b)
    hb = gpu.HyyroBatch([(a, b), (b[:5000], a)])
    for r in (1, 2, 4):
        hb.set_words_per_lane([0, 0, 0, r, 0])
        print(f"hyyro segmented sigma={sigma} 32-bit words per lane {r}", flush=True)
        got = hb.run()
        assert got[0] == want
        assert got[1] == oracle.llcs_dp(b[:5000], a)
    hb.close()


def test_large_pairs_equal_the_semi_local_kernel(gpu):
    g = np.random.default_rng(9)
    for m, n in [(3000, 1 << 20), (20_000, 300_000)]:
        a = g.integers(0, 4, m).astype(np.uint8)
        b = a[g.integers(0, m, n)] if n < m else g.integers(0, 4, n).astype(np.uint8)
        print(f"hyyro large m={m} n={n}", flush=True)
        perm = gpu.semi_local_comb(a, b)
        assert gpu.llcs_hyyro(a, b) == gpu.llcs_from_kernel(perm, m, n)


def test_unsupported_inputs_are_rejected(gpu):
    g = np.random.default_rng(1)
    a = np.arange(300, dtype=np.int32)
    with pytest.raises(gpu.SemiLocalError):
        gpu.llcs_hyyro(a, a)                                  # 300 distinct symbols
    with pytest.raises(gpu.SemiLocalError):
        gpu.llcs_hyyro(g.integers(0, 4, 10).astype(np.uint8), np.zeros((1 << 20) + 1, np.uint8))
    assert gpu.llcs_hyyro(np.zeros(0, np.uint8), np.ones(5, np.uint8)) == 0
    assert gpu.llcs_hyyro(np.array([1, 2, 3], np.uint8), np.array([4, 5], np.uint8)) == 0


# offered 32-bit words per lane per bucket (packed width up to 64, 128, 256,
# 2,048 and 16,384 64-bit words), default first
WPL_OPTIONS = [[2, 1, 4], [2, 1, 4], [2, 4], [2, 1, 4], [2, 1, 4]]


def test_every_words_per_lane_form_equals_dp(gpu, oracle):
    g = np.random.default_rng(77)
    widths = [1, 63, 64, 65, 4095, 4096, 4097, 8191, 8192, 8193, 12000, 16384, 16385, 65536, 65537, 131072,
              131073, 200000, 524289, 1 << 20]
    pairs = []
    for n in widths:
        m = int(min(600, max(1, 4_000_000 // n)))
        sigma = int(g.choice([2, 4, 20]))
        pairs.append((g.integers(0, sigma, m).astype(np.uint8), g.integers(0, sigma, n).astype(np.uint8)))
    want = [oracle.llcs_dp(a, b) for a, b in pairs]
    hb = gpu.HyyroBatch(pairs)
    b0 = hb.buckets()
    assert [b["max_words"] for b in b0] == [64, 128, 256, 2048, 16384]
    assert [b["pairs"] for b in b0] == [6, 3, 3, 4, 4]
    assert [b["words_per_lane"] for b in b0] == [o[0] for o in WPL_OPTIONS]
    for k in range(3):
        wpl = [o[min(k, len(o) - 1)] for o in WPL_OPTIONS]   # every form of every bucket
        hb.set_words_per_lane(wpl)
        assert [b["words_per_lane"] for b in hb.buckets()] == wpl
        print(f"hyyro words per lane {wpl}", flush=True)
        assert hb.run().tolist() == want
    with pytest.raises(gpu.SemiLocalError):
        hb.set_words_per_lane([0, 0, 1, 0, 0])                # bucket 2 has no 32-bit form
    chosen = hb.autotune(rounds=1)
    assert all(c in o for c, o in zip(chosen, WPL_OPTIONS))
    assert [b["words_per_lane"] for b in hb.buckets()] == chosen
    assert hb.run().tolist() == want
    hb.close()
