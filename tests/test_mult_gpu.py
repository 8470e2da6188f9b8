"""Kernel (c) on the GPU: braid multiplication (steady ant, level by level),
staggered glue and the hybrid scheduler, against the reference's outputs
(golden fixtures written by the compiled reference), the oracle, and the
uniqueness of the semi-local kernel."""
import numpy as np
import pytest

import golden_io as G

pytestmark = pytest.mark.gpu


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def test_steady_ant_golden_each(gpu):
    # steady_ant_parallel_wrapper outputs of the reference (n = 1 .. 20000)
    for p, q, r in G.steady_ant():
        assert np.array_equal(gpu.braid_multiply(p, q), r), len(p)


def test_steady_ant_golden_as_one_batch(gpu):
    cases = G.steady_ant()
    got = gpu.braid_multiply_batch([(p, q) for p, q, _ in cases])
    for (p, q, r), g in zip(cases, got):
        assert np.array_equal(g, r), len(p)


@pytest.mark.parametrize("n", [1023, 1024, 1025, 2047, 2048, 2049, 3000, 4097, 12345, 65536, 65537, 200003])
def test_random_sizes_vs_oracle(gpu, oracle, n):
    g = _rng(n)
    p, q = g.permutation(n).astype(np.int32), g.permutation(n).astype(np.int32)
    assert np.array_equal(gpu.braid_multiply(p, q), oracle.steady_ant(p, q))


def test_small_sizes_vs_naive_product(gpu, oracle):
    g = _rng(11)
    pairs = []
    for n in list(range(0, 70)) + [127, 128, 129, 200]:
        pairs.append((g.permutation(n).astype(np.int32), g.permutation(n).astype(np.int32)))
    got = gpu.braid_multiply_batch(pairs)
    for (p, q), r in zip(pairs, got):
        want = oracle.mult_naive(p, q) if len(p) else np.zeros(0, np.int32)
        assert np.array_equal(r, want), len(p)


def test_identity_is_neutral_and_product_is_permutation_at_2p21(gpu):
    n = (1 << 21) + 5
    g = _rng(21)
    p = g.permutation(n).astype(np.int32)
    ident = np.arange(n, dtype=np.int32)
    assert np.array_equal(gpu.braid_multiply(p, ident), p)
    assert np.array_equal(gpu.braid_multiply(ident, p), p)
    q = g.permutation(n).astype(np.int32)
    r = gpu.braid_multiply(p, q)
    assert np.array_equal(np.sort(r), ident)


def test_associativity_at_300k(gpu):
    n = 300_001
    g = _rng(3)
    p, q, s = (g.permutation(n).astype(np.int32) for _ in range(3))
    left = gpu.braid_multiply(gpu.braid_multiply(p, q), s)
    right = gpu.braid_multiply(p, gpu.braid_multiply(q, s))
    assert np.array_equal(left, right)


def test_mixed_batch_equals_individual_products(gpu, oracle):
    g = _rng(5)
    sizes = [0, 1, 5, 1024, 1500, 70000, 3, 40000, 2049]
    pairs = [(g.permutation(n).astype(np.int32), g.permutation(n).astype(np.int32)) for n in sizes]
    got = gpu.braid_multiply_batch(pairs)
    for (p, q), r in zip(pairs, got):
        assert np.array_equal(r, oracle.steady_ant(p, q) if len(p) else np.zeros(0, np.int32)), len(p)


def test_rejects_non_permutations(gpu):
    with pytest.raises(gpu.SemiLocalError):
        gpu.braid_multiply(np.array([0, 0, 1], np.int32), np.array([0, 1, 2], np.int32))
    with pytest.raises(gpu.SemiLocalError):
        gpu.braid_multiply(np.array([0, 1, 3], np.int32), np.array([0, 1, 2], np.int32))


@pytest.mark.parametrize("row_glue", [False, True])
def test_staggered_glue_matches_oracle_and_whole_combing(gpu, oracle, row_glue):
    # two adjacent blocks of one grid: glued braid == braid of the whole grid
    g = _rng(17 + row_glue)
    for (m, n1, n2) in [(300, 200, 500), (1500, 2500, 1200), (7, 3, 9), (2048, 1024, 3000)]:
        if row_glue:   # vertical neighbours: a = a1 + a2 over one b; p = upper block (a1), q = lower
            a = g.integers(0, 4, size=n1 + n2).astype(np.uint8)
            b = g.integers(0, 4, size=m).astype(np.uint8)
            p = gpu.semi_local_comb(a[:n1], b)
            q = gpu.semi_local_comb(a[n1:], b)
            k = m   # shared strands: the columns
            whole = gpu.semi_local_comb(a, b)
            got = gpu.staggered_multiply(p, q, k, True)
            assert np.array_equal(got, oracle.staggered(p, q, k, True))
        else:          # horizontal neighbours: b = b1 + b2 for one a
            a = g.integers(0, 4, size=m).astype(np.uint8)
            b = g.integers(0, 4, size=n1 + n2).astype(np.uint8)
            p = gpu.semi_local_comb(a, b[:n1])
            q = gpu.semi_local_comb(a, b[n1:])
            k = m   # shared strands: the rows
            whole = gpu.semi_local_comb(a, b)
            got = gpu.staggered_multiply(p, q, k, False)
            assert np.array_equal(got, oracle.staggered(p, q, k, False))
        assert np.array_equal(got, whole)


def test_staggered_random_permutations_vs_oracle(gpu, oracle):
    g = _rng(23)
    for pr, qr, k, rg in [(10, 12, 4, 0), (10, 12, 4, 1), (3000, 2500, 1700, 0), (3000, 2500, 1700, 1),
                          (5, 5, 0, 0), (5, 5, 0, 1), (6, 9, 6, 0), (9, 6, 6, 1)]:
        p, q = g.permutation(pr).astype(np.int32), g.permutation(qr).astype(np.int32)
        assert np.array_equal(gpu.staggered_multiply(p, q, k, bool(rg)), oracle.staggered(p, q, k, bool(rg)))
    # k == pr == qr: the reference returns without filling the product
    with pytest.raises(gpu.SemiLocalError):
        gpu.staggered_multiply(np.arange(6, dtype=np.int32), np.arange(6, dtype=np.int32), 6, False)


def test_hybrid_small_golden(gpu):
    # hybrid_iterative outputs of the compiled reference
    for a, b, perm, sm, sn in G.hybrid_small():
        assert np.array_equal(gpu.hybrid_iterative(a, b, sm, sn), perm), (len(a), len(b), sm, sn)


def test_hybrid_shapes_equal_direct_combing(gpu):
    g = _rng(29)
    a = g.integers(0, 4, size=20000).astype(np.uint8)
    b = g.integers(0, 4, size=13000).astype(np.uint8)
    whole = gpu.semi_local_comb(a, b)
    for sm, sn in [(1, 1), (2, 3), (4, 4), (1, 8), (8, 1), (5, 7)]:
        assert np.array_equal(gpu.hybrid_iterative(a, b, sm, sn), whole), (sm, sn)


def test_resident_hybrid_repeats(gpu):
    g = _rng(31)
    a = g.integers(0, 20, size=5000).astype(np.uint8)
    b = g.integers(0, 20, size=7000).astype(np.uint8)
    h = gpu.Hybrid(a, b, 4, 4)
    first = h.run()
    assert np.array_equal(first, gpu.semi_local_comb(a, b))
    tot, comb, merge = h.time(warmup=1, iters=2)
    assert tot > 0 and comb >= 0 and merge >= 0
    assert np.array_equal(h.run(), first)
    cells, levels = h.stats()
    assert cells == len(a) * len(b) and levels == 4
    h.close()


def test_hybrid_graph_replays_equal_the_eager_run(gpu):
    # the whole pipeline captured into a HIP graph and replayed: every replay,
    # with eager work and synchronisation in between, equals the eager run
    g = _rng(37)
    a = g.integers(0, 4, size=30000).astype(np.uint8)
    b = g.integers(0, 4, size=26000).astype(np.uint8)
    eager = gpu.Hybrid(a, b, 4, 4).run()
    assert np.array_equal(eager, gpu.semi_local_comb(a, b))
    h = gpu.Hybrid(a, b, 4, 4)
    h.set_graph(True)
    for k in range(4):
        assert np.array_equal(h.run(), eager), k
        x = gpu.semi_local_comb(a[:3000], b[:2000])            # unrelated eager work between replays
        assert gpu.llcs_from_kernel(x, 3000, 2000) >= 0
    tot, comb, merge = h.time(warmup=1, iters=3)
    assert tot > 0 and comb == -1 and merge == -1
    assert np.array_equal(h.run(), eager)
    h.set_graph(False)                                        # back to eager launches on the same handle
    assert np.array_equal(h.run(), eager)
    h.close()


def test_hybrid_autotune_keeps_results_and_recaptures_the_graph(gpu):
    # tuning the combing launch changes tile shape and queue order (and can
    # move the queue's device buffers): eager runs and a graph captured before
    # the tuning must both still give the eager result
    g = _rng(41)
    a = g.integers(0, 4, size=20000).astype(np.uint8)
    b = g.integers(0, 4, size=17000).astype(np.uint8)
    want = gpu.semi_local_comb(a, b)
    h = gpu.Hybrid(a, b, 4, 2)
    h.set_graph(True)
    assert np.array_equal(h.run(), want)                      # graph captured at the default shape
    r, w = h.autotune(iters=1)
    assert r in (0, 4, 8, 16) and 1 <= w <= 64
    for _ in range(2):
        assert np.array_equal(h.run(), want)                  # captured again at the tuned shape
    h.set_graph(False)
    assert np.array_equal(h.run(), want)
    h.close()


def test_block_grid_autotuner_picks_a_candidate_per_length_bucket(gpu):
    # the grid autotuner times each candidate block split of the pair and keeps
    # the fastest for the pair's length bucket (floor(log2(max(m, n)))); the
    # hybrid at that grid gives the same permutation as direct combing
    g = _rng(43)
    a = g.integers(0, 4, size=40000).astype(np.uint8)
    b = a.copy()
    b[g.integers(0, len(b), 2000)] = g.integers(0, 4, 2000).astype(np.uint8)
    grids = [(1, 1), (2, 2), (4, 4), (2, 8)]
    assert gpu.hybrid_bucket_blocks(1 << 29, 5) == (0, 0)     # a bucket never tuned
    best, ms = gpu.hybrid_autotune_blocks(a, b, grids, iters=1)
    assert best in grids and len(ms) == len(grids) and all(t > 0 for t in ms)
    assert ms[grids.index(best)] == min(ms)
    assert gpu.hybrid_bucket_blocks(len(a), len(b)) == best
    assert gpu.hybrid_bucket_blocks(33000, 20000) == best       # the same bucket (2^15 .. 2^16 - 1)
    assert np.array_equal(gpu.hybrid_iterative(a, b, *best), gpu.semi_local_comb(a, b))
    with pytest.raises(gpu.SemiLocalError):
        gpu.hybrid_autotune_blocks(a, b, [(0, 2)])
