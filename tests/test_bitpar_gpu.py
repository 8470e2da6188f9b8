"""GPU parity of kernel (b) (bit-parallel LLCS, formula 2) against the
reference's golden vectors, DP, and the integer permutation of kernel (a)
(0-1 principle: the 1-bit strands are the integer strands thresholded at m)."""
import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu


def test_bitparallel_golden_vectors(gpu):
    for a, b, llcs in G.bitpar():
        assert gpu.llcs_bitparallel(a, b) == llcs


def test_bitparallel_vs_dp_random(gpu, oracle):
    g = np.random.default_rng(21)
    for _ in range(25):
        la, lb = 64 * int(g.integers(1, 40)), 64 * int(g.integers(1, 40))
        a = g.integers(0, 2, la).astype(np.uint8)
        b = g.integers(0, 2, lb).astype(np.uint8)
        assert gpu.llcs_bitparallel(a, b) == oracle.llcs_dp(a, b), (la, lb)


def test_bitparallel_matches_integer_permutation(gpu):
    g = np.random.default_rng(22)
    for la, lb in [(1 << 14, 1 << 15), (1 << 15, 1 << 14), (4160, 4160 * 3)]:
        a = g.integers(0, 2, la).astype(np.uint8)
        b = g.integers(0, 2, lb).astype(np.uint8)
        perm = gpu.semi_local_comb(a, b)
        assert gpu.llcs_bitparallel(a, b) == gpu.llcs_from_kernel(perm, la, lb)


def test_resident_plan_repeats(gpu, oracle):
    g = np.random.default_rng(23)
    a = g.integers(0, 2, 64 * 300).astype(np.uint8)
    b = g.integers(0, 2, 64 * 500).astype(np.uint8)
    plan = gpu.BitParallel(a, b)
    want = oracle.bitpar_llcs(a, b)
    assert [plan.llcs() for _ in range(3)] == [want] * 3
    plan.close()
