"""GPU parity of approximate matching (sl_amatch_*, sl_duplicate_*): the
per-window arrays against the oracle's direct-DP windows (many chunk sizes,
min_len from 0 to w, window triangles from 1 to ~2,100 columns), the
duplicate searches against the oracle and the reference compiled in place,
and large texts against the oracle on slices (a window's values depend on
its own symbols only).  Bar: exact integers and identical intervals."""
import numpy as np
import pytest

from oracle.binding import AMatchOracle, Oracle, RefAMatch

pytestmark = pytest.mark.gpu


def planted(g, m, n, sigma, copies):
    p = g.integers(0, sigma, m)
    t = g.integers(0, sigma, n)
    for _ in range(copies):
        pos = int(g.integers(0, max(1, n - m)))
        c = p.copy()
        c[g.integers(0, m, max(1, m // 8))] = g.integers(0, sigma)
        t[pos:pos + m] = c[:len(t[pos:pos + m])]
    return p.astype(np.int32), t.astype(np.int32)


@pytest.mark.parametrize("m,n,W,w,L,chunk,sigma", [
    (1, 50, 5, 5, 0, 0, 2), (10, 400, 30, 12, 6, 16, 4), (20, 3000, 25, 25, 18, 0, 4),
    (20, 3000, 40, 33, 0, 100, 2), (8, 700, 64, 64, 64, 64, 3), (33, 2500, 200, 150, 20, 257, 4),
    (50, 1200, 300, 300, 1, 0, 26), (13, 900, 70, 70, 35, 1, 4)])
def test_windows_equal_the_oracle(gpu, m, n, W, w, L, chunk, sigma):
    g = np.random.default_rng(m * 31 + n + w)
    p, t = planted(g, m, n, sigma, 3)
    print(f"windows m={m} n={n} W={W} w={w} L={L} chunk={chunk}", flush=True)
    am = gpu.AMatch(p, t, W, chunk)
    got = am.windows(w, L)
    want = AMatchOracle().windows(p, t, w, L)
    for a, b, name in zip(got, want, ("window_score", "best_start", "best_end", "best_score")):
        assert np.array_equal(a, b), name
    am.close()


def test_wide_window_triangle_uses_64_columns_per_lane(gpu):
    # w - L + 1 = 2,091 columns: the widest register layout
    g = np.random.default_rng(11)
    m, w, L = 20, 2100, 10
    p, t = planted(g, m, w + 40, 4, 2)
    am = gpu.AMatch(p, t, w)
    got = am.windows(w, L)
    want = AMatchOracle().windows(p, t, w, L)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    am.close()


@pytest.mark.parametrize("m,n,k,sigma", [(12, 400, 0.8, 4), (30, 2000, 0.9, 4), (25, 1500, 0.6, 2),
                                         (40, 6000, 0.95, 8), (7, 300, 1.0, 2), (64, 8000, 0.75, 4)])
def test_duplicate_searches_equal_the_oracle_and_the_reference(gpu, m, n, k, sigma):
    g = np.random.default_rng(m + n)
    p, t = planted(g, m, n, sigma, 8)
    ao = AMatchOracle()
    ref = RefAMatch() if RefAMatch.available() else None
    am = gpu.AMatch(p, t, int(m / k))
    for v in (0, 1):
        print(f"duplicate search variant {v} m={m} n={n} k={k}", flush=True)
        got = am.duplicate_search(k, v)
        assert got == ao.dup_search(p, t, k, v)
        if ref is not None:
            assert got == ref.dup_search(p, t, k, v)
    am.close()
    assert gpu.duplicate_search(p, t, k, 0) == ao.dup_search(p, t, k, 0)


def test_large_text_against_the_oracle_on_slices(gpu):
    g = np.random.default_rng(5)
    m, n, k = 200, 400_000, 0.9
    p, t = planted(g, m, n, 4, 200)
    w, L = int(m / k), int(m * k)
    am = gpu.AMatch(p, t, w)
    assert am.stats()["chunks"] > 1
    ws, bs, be, bsc = am.windows(w, L)
    assert len(ws) == n - w + 1
    # every best substring lies in its window, is long enough, and scores at
    # least the window itself
    i = np.arange(n - w + 1)
    assert np.all(bs >= i) and np.all(be <= i + w) and np.all(be - bs >= L) and np.all(bsc >= ws)
    ao = AMatchOracle()
    for a in (0, 123_457, n - 2_500):
        sl = t[a:a + 2_500]
        want = ao.windows(p, sl, w, L)
        cnt = len(sl) - w + 1
        assert np.array_equal(ws[a:a + cnt], want[0])
        assert np.array_equal(bs[a:a + cnt] - a, want[1])
        assert np.array_equal(be[a:a + cnt] - a, want[2])
        assert np.array_equal(bsc[a:a + cnt], want[3])
    o = Oracle()
    for j in g.integers(0, n - w + 1, 10):
        assert bsc[j] == 2 * o.llcs_dp(p, t[bs[j]:be[j]]) - m - (be[j] - bs[j])
    comb_ms, win_ms = am.time(w, L, warmup=1, iters=2)
    assert comb_ms > 0 and win_ms > 0
    am.close()


def test_bad_arguments_are_rejected(gpu):
    p = np.array([1, 2, 3], np.int32)
    t = np.arange(50, dtype=np.int32) % 4
    am = gpu.AMatch(p, t, 10)
    with pytest.raises(gpu.SemiLocalError):
        am.windows(11, 2)          # longer than max_window
    with pytest.raises(gpu.SemiLocalError):
        am.windows(5, 6)           # min_len > w
    with pytest.raises(gpu.SemiLocalError):
        am.duplicate_search(0.0)   # k outside (0, 1]
    with pytest.raises(gpu.SemiLocalError):
        am.duplicate_search(0.2)   # int(m / k) = 15 > max_window
    with pytest.raises(gpu.SemiLocalError):
        am.duplicate_search(0.9, variant=2)
    am.close()
    with pytest.raises(gpu.SemiLocalError):
        gpu.AMatch(p, t[:0], 10)   # empty text
