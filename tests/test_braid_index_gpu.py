"""GPU parity of the resident-braid index (wavelet matrix, sl_braid_index_*):
every query kind answered from the index equals the dominance count taken
directly from the permutation, and, on samples, the LCS of the actual
substrings from the oracle's DP.  Bar: exact integers."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _count(perm, r0, j):
    # #{r >= r0 : P[r] < j}
    r0 = max(0, min(len(perm), r0))
    return int(np.count_nonzero(perm[r0:] < j))


def _expected(perm, m, n, kind, i, j):
    if kind == 2:                                   # LCS(a[i:j), b) = H(-i, n+m-j) - i - (m-j)
        hi_, hj = -i, n + m - j
        return hj - hi_ - _count(perm, hi_ + m, hj) - i - (m - j)
    return j - i - _count(perm, i + m, j)           # H(i, j); LCS(a, b[i:j)) = H(i, j)


def _queries(g, m, n, count):
    kinds, qi, qj = [], [], []
    for _ in range(count):
        k = int(g.integers(0, 3))
        if k == 0:
            i, j = int(g.integers(-m, n + 1)), int(g.integers(0, m + n + 1))
        elif k == 1:
            i = int(g.integers(0, n + 1)); j = int(g.integers(i, n + 1))
        else:
            i = int(g.integers(0, m + 1)); j = int(g.integers(i, m + 1))
        kinds.append(k); qi.append(i); qj.append(j)
    return np.array(kinds), np.array(qi), np.array(qj)


@pytest.mark.parametrize("m,n", [(1, 1), (0, 7), (5, 0), (31, 33), (64, 64), (65, 1000), (1000, 3000),
                                 (5000, 20000)])
def test_index_answers_equal_direct_counts(gpu, oracle, m, n):
    g = np.random.default_rng(m * 7 + n)
    a = g.integers(0, 4, m).astype(np.uint8)
    b = g.integers(0, 4, n).astype(np.uint8)
    perm = gpu.semi_local_comb(a, b) if m and n else np.arange(m + n, dtype=np.int32)
    ix = gpu.BraidIndex(perm, m, n)
    kinds, qi, qj = _queries(g, m, n, 400)
    got = ix.query(kinds, qi, qj)
    want = [_expected(perm, m, n, k, i, j) for k, i, j in zip(kinds, qi, qj)]
    assert np.array_equal(got, np.array(want, np.int64))
    # the formulas themselves, on samples: LCS of the actual substrings
    for k, i, j, v in list(zip(kinds, qi, qj, got))[:25]:
        if k == 1:
            assert v == oracle.llcs_dp(a, b[i:j])
        elif k == 2:
            assert v == oracle.llcs_dp(a[i:j], b)
        elif m and n:
            assert v == oracle.H(perm, m, n, int(i), int(j))
    ix.close()


def test_large_random_permutation_and_index_size(gpu):
    # 2^20 entries: 20 levels; counts against numpy on a random permutation
    g = np.random.default_rng(3)
    N = 1 << 20
    perm = g.permutation(N).astype(np.int32)
    m = N // 3
    n = N - m
    ix = gpu.BraidIndex(perm, m, n)
    st = ix.stats()
    assert st["levels"] == 20 and st["bytes"] < 0.2 * N * 20 + 4096
    i = g.integers(-m, n + 1, 200)
    j = g.integers(0, N + 1, 200)
    got = ix.query(np.zeros(200, np.int32), i, j)
    want = [int(jj) - int(ii) - _count(perm, int(ii) + m, int(jj)) for ii, jj in zip(i, j)]
    assert np.array_equal(got, np.array(want, np.int64))
    ix.close()


def test_bad_query_kind_is_rejected(gpu):
    ix = gpu.BraidIndex(np.arange(10, dtype=np.int32), 4, 6)
    with pytest.raises(gpu.SemiLocalError):
        ix.query([3], [0], [1])
    ix.close()
