"""CPU pinning of the approximate-matching oracle (oracle/amatch.cpp): its
duplicate searches against the reference's own InteractiveDuplicateSearch
SemiDense / SemiSparse (approximate_matching/ApproximateMatching.h:252-432)
compiled in place, and its window maxima against the definition (LCS by
plain DP over every substring of small windows).  Bar: identical intervals
and scores."""
import numpy as np
import pytest

from oracle.binding import AMatchOracle, Oracle, RefAMatch

KS = [0.5, 0.6, 0.7, 0.75, 0.8, 0.85, 0.9, 0.95, 1.0]


def planted(g, m, n, sigma, copies):
    p = g.integers(0, sigma, m)
    t = g.integers(0, sigma, n)
    for _ in range(copies):
        pos = int(g.integers(0, max(1, n - m)))
        c = p.copy()
        c[g.integers(0, m, max(1, m // 8))] = g.integers(0, sigma)
        t[pos:pos + m] = c[:len(t[pos:pos + m])]
    return p, t


@pytest.mark.skipif(not RefAMatch.available(), reason="oracle/_ref needs the reference tree to build")
@pytest.mark.parametrize("seed", range(4))
def test_oracle_duplicate_searches_equal_the_reference(seed):
    o, r = AMatchOracle(), RefAMatch()
    g = np.random.default_rng(100 + seed)
    nonempty = 0
    for _ in range(120):
        m = int(g.integers(1, 40))
        k = float(g.choice(KS))
        sigma = int(g.choice([2, 4, 8]))
        n = int(g.integers(int(m / k), int(m / k) + 400))
        p, t = planted(g, m, n, sigma, int(g.integers(0, 4)))
        for v in (0, 1):
            want = r.dup_search(p, t, k, v)
            assert o.dup_search(p, t, k, v) == want, (m, n, k, sigma, v)
            nonempty += bool(want)
    assert nonempty > 100


def test_oracle_windows_equal_the_definition():
    # window maximum = max over substrings [s, e) of the window with e - s >= L
    # of (score, length, end); score = 2 LCS - m - (e - s)
    o, ao = Oracle(), AMatchOracle()
    g = np.random.default_rng(7)
    for m, n, w, L in [(5, 30, 8, 3), (7, 25, 7, 0), (3, 20, 12, 12), (6, 40, 10, 5)]:
        p = g.integers(0, 3, m).astype(np.int32)
        t = g.integers(0, 3, n).astype(np.int32)
        ws, bs, be, bsc = ao.windows(p, t, w, L)
        for i in range(n - w + 1):
            def sc(s, e):
                return 2 * o.llcs_dp(p, t[s:e]) - m - (e - s)
            assert ws[i] == sc(i, i + w)
            cands = [(sc(s, e), e - s, e) for e in range(i + L, i + w + 1) for s in range(i, e - L + 1)]
            top = max(cands)
            assert (bsc[i], be[i] - bs[i], be[i]) == top


def test_oracle_rejects_bad_windows():
    ao = AMatchOracle()
    with pytest.raises(ValueError):
        ao.windows([1, 2], [1, 2, 3], 4, 1)       # window longer than the text
    with pytest.raises(ValueError):
        ao.windows([1, 2], [1, 2, 3], 2, 3)       # min_len > w
