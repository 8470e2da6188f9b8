"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted:
golden vectors from the reference's own build, brute force from the explicit
semi-local matrix, DP LLCS, the O(n^3) tropical product and algebraic
invariants.  A planted mistake in the oracle is shown to fail these checks."""
import numpy as np
import pytest

import golden_io as G
from semilocal_amd import workloads as W


def test_comb_matches_reference_golden(oracle):
    for a, b, perm, _ in G.comb_sweep():
        assert np.array_equal(oracle.comb(a, b), perm)
        assert np.array_equal(oracle.comb_rowmajor(a, b), perm)


def test_comb_matches_bruteforce_kernel(oracle):
    # DummyPrefixLCSStrategy-style kernel from all window LCS values: pins
    # orientation, strand numbering and the reverse mapping for m > n
    g = np.random.default_rng(5)
    for _ in range(120):
        m, n = int(g.integers(1, 14)), int(g.integers(1, 14))
        sigma = int(g.integers(1, 5))
        a, b = g.integers(0, sigma, m), g.integers(0, sigma, n)
        assert np.array_equal(oracle.comb(a, b), oracle.kernel_bruteforce(a, b)), (m, n, sigma)


def test_llcs_readout_matches_dp(oracle):
    g = np.random.default_rng(6)
    for _ in range(100):
        m, n = int(g.integers(1, 200)), int(g.integers(1, 200))
        sigma = int(g.integers(2, 6))
        a, b = g.integers(0, sigma, m), g.integers(0, sigma, n)
        p = oracle.comb(a, b)
        assert oracle.llcs_from_perm(p, m, n) == oracle.llcs_dp(a, b)


def test_semi_local_H_matches_dp_on_windows(oracle):
    g = np.random.default_rng(7)
    for _ in range(40):
        m, n = int(g.integers(1, 40)), int(g.integers(1, 40))
        a, b = g.integers(0, 3, m), g.integers(0, 3, n)
        p = oracle.comb(a, b)
        for _ in range(10):
            lo = int(g.integers(0, n + 1))
            hi = int(g.integers(lo, n + 1))
            assert oracle.H(p, m, n, lo, hi) == oracle.llcs_dp(a, b[lo:hi])      # string-substring
            s = int(g.integers(0, m + 1))
            e = int(g.integers(s, m + 1))
            assert oracle.H(p, m, n, -s, n + m - e) - s - (m - e) == oracle.llcs_dp(a[s:e], b)   # substring-string


def test_steady_ant_matches_reference_and_naive(oracle):
    for p, q, r in G.steady_ant():
        assert np.array_equal(oracle.steady_ant(p, q), r)
        if len(p) <= 130:
            assert np.array_equal(oracle.mult_naive(p, q), r)


def test_steady_ant_invariants(oracle):
    g = np.random.default_rng(8)
    for n in [1, 2, 3, 10, 57, 200]:
        p, q, s = (g.permutation(n).astype(np.int32) for _ in range(3))
        ident = np.arange(n, dtype=np.int32)
        assert np.array_equal(oracle.steady_ant(p, ident), p)   # identity is neutral
        assert np.array_equal(oracle.steady_ant(ident, q), q)
        lhs = oracle.steady_ant(oracle.steady_ant(p, q), s)      # associativity
        rhs = oracle.steady_ant(p, oracle.steady_ant(q, s))
        assert np.array_equal(lhs, rhs)
        r = oracle.steady_ant(p, q)
        assert np.array_equal(np.sort(r), ident)                 # product is a permutation


def test_staggered_glue_equals_combing_concatenation(oracle):
    # multiplying the braids of adjacent blocks equals combing the concatenation
    g = np.random.default_rng(9)
    for _ in range(30):
        m, n1, n2 = (int(x) for x in g.integers(1, 60, 3))
        a = g.integers(0, 4, m)
        b1, b2 = g.integers(0, 4, n1), g.integers(0, 4, n2)
        p, q = oracle.comb(a, b1), oracle.comb(a, b2)
        assert np.array_equal(oracle.staggered(p, q, m, False), oracle.comb(a, np.concatenate([b1, b2])))
        a2 = g.integers(0, 4, n1)
        p2, q2 = oracle.comb(a, b1), oracle.comb(a2, b1)
        assert np.array_equal(oracle.staggered(p2, q2, n1, True), oracle.comb(np.concatenate([a, a2]), b1))


def test_hybrid_matches_reference_golden(oracle):
    for a, b, perm, sm, sn in G.hybrid_small():
        assert np.array_equal(oracle.hybrid_iterative(a, b, sm, sn), perm)
        assert np.array_equal(oracle.comb(a, b), perm)


def test_bitparallel_matches_reference_and_dp(oracle):
    for a, b, llcs in G.bitpar():
        assert oracle.bitpar_llcs(a, b) == llcs
        if len(a) * len(b) <= 4096 * 4096:
            assert oracle.llcs_dp(a, b) == llcs


def test_config_fixtures_reproduce(oracle):
    cfg = G.configs()
    a, b = W.plain_pair()
    c = cfg["plain_pair_10k"]
    assert W.sha256(a, b) == c["input_sha256"]
    p = oracle.comb(a, b, threads=2)
    assert W.sha256(p) == c["perm_sha256"]
    assert oracle.llcs_from_perm(p, len(a), len(b)) == c["llcs"]


def test_ragged_batch_fixture_subset(oracle):
    cfg = G.configs()["ragged_batch_4096"]
    pairs = W.ragged_batch()
    assert W.sha256(*[np.concatenate([x, y]) for x, y in pairs]) == cfg["input_sha256"]
    idx = [i for i, (x, y) in enumerate(pairs) if len(x) * len(y) < 2_000_000][:40]
    perms = oracle.comb_batch([pairs[i] for i in idx], threads=2)
    for i, p in zip(idx, perms):
        assert W.sha256(p) == cfg["perm_sha256_each"][i]


def test_planted_mistake_is_caught(oracle):
    # an oracle that forgot the reverse mapping for m > n must fail the pins
    g = np.random.default_rng(10)
    a, b = g.integers(0, 3, 9), g.integers(0, 3, 5)
    good = oracle.comb(a, b)
    size = len(a) + len(b)
    wrong = np.empty_like(good)
    pba = oracle.comb(b, a)      # kernel of (b, a) used as if it were (a, b)
    wrong[:] = pba
    assert not np.array_equal(wrong, oracle.kernel_bruteforce(a, b))
    assert np.array_equal(good, oracle.kernel_bruteforce(a, b))
    # and the correct reverse mapping recovers it: P_ab[x] = size-1 - P_ba[size-1-x]
    fixed = size - 1 - pba[::-1]
    assert np.array_equal(fixed, good)


@pytest.mark.parametrize("sigma", [2, 4, 20])
def test_generators_are_deterministic(sigma):
    x = W.uniform_string(W.rng(3), 1000, sigma)
    y = W.uniform_string(W.rng(3), 1000, sigma)
    assert np.array_equal(x, y) and x.max() < sigma
    m = W.mutate(W.rng(4), x, sigma, 0.1, 0.02)
    assert abs(len(m) - len(x)) <= 1
