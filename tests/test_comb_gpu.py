"""GPU parity of kernel (a) (anti-diagonal sticky-braid combing) against the
reference's golden vectors and the oracle, through the C ABI.  Bar: bit-exact
row->col permutations (and col->row where requested)."""
import numpy as np
import pytest

import golden_io as G
from semilocal_amd import workloads as W

pytestmark = pytest.mark.gpu


def test_golden_sweep_single_pair(gpu):
    for a, b, perm, _ in G.comb_sweep():
        got = gpu.semi_local_comb(a, b)
        assert np.array_equal(got, perm), (len(a), len(b))


def test_golden_sweep_as_one_batch(gpu):
    cases = G.comb_sweep()
    res = gpu.comb_batch([(a.astype(np.uint8), b.astype(np.uint8)) for a, b, _, _ in cases])
    for (a, b, perm, _), got in zip(cases, res):
        assert np.array_equal(got, perm), (len(a), len(b))


@pytest.mark.parametrize("rows_per_lane", [4, 8, 16])
def test_tile_height_does_not_change_result(gpu, oracle, rows_per_lane):
    g = np.random.default_rng(rows_per_lane)
    pairs = [(g.integers(0, 4, int(g.integers(1, 3000))).astype(np.uint8),
              g.integers(0, 4, int(g.integers(1, 3000))).astype(np.uint8)) for _ in range(12)]
    b = gpu.Batch(pairs, rows_per_lane=rows_per_lane)
    res = b.run().results()
    for (x, y), got in zip(pairs, res):
        assert np.array_equal(got, oracle.comb(x, y))
    res2 = b.run(per_pair=True).results()
    for r1, r2 in zip(res, res2):
        assert np.array_equal(r1, r2)
    b.close()


def test_edge_shapes(gpu, oracle):
    g = np.random.default_rng(11)
    shapes = [(0, 0), (0, 5), (5, 0), (1, 1), (1, 1000), (1000, 1), (63, 64), (64, 63), (1023, 1025),
              (1025, 1023), (2049, 100), (100, 2049)]
    for m, n in shapes:
        a = g.integers(0, 3, m).astype(np.uint8)
        b = g.integers(0, 3, n).astype(np.uint8)
        got = gpu.semi_local_comb(a, b)
        assert np.array_equal(got, oracle.comb(a, b)) if m and n else np.array_equal(got, np.arange(m + n))


def test_symbol_widths_and_remapping(gpu, oracle):
    g = np.random.default_rng(12)
    a = g.integers(-1000, 1000, 700).astype(np.int32)      # > 256 distinct symbols -> uint16 pool
    b = np.concatenate([a[100:500], g.integers(-1000, 1000, 300).astype(np.int32)])
    assert np.array_equal(gpu.semi_local_comb(a, b), oracle.comb(a, b))
    a8 = g.integers(0, 256, 900).astype(np.uint8)          # all byte values incl. 255
    b8 = g.integers(0, 256, 1300).astype(np.uint8)
    assert np.array_equal(gpu.semi_local_comb(a8, b8), oracle.comb(a8, b8))


def test_col_to_row_is_inverse(gpu):
    g = np.random.default_rng(13)
    a, b = g.integers(0, 4, 1500).astype(np.uint8), g.integers(0, 4, 900).astype(np.uint8)
    rc, cr = gpu.semi_local_comb(a, b, with_col_to_row=True)
    assert np.array_equal(cr[rc], np.arange(len(rc)))


def test_reference_sweep_naive_vs_openmp_shapes(gpu, oracle):
    # the shapes of SemiLocalLCSTest.NaiveVsOpenMPCorrectness (4900..4999 step 11, sigma 11)
    g = np.random.default_rng(14)
    pairs = []
    for ma in range(4900, 5000, 33):
        for nb in range(4900, 5000, 33):
            pairs.append((g.integers(0, 11, ma).astype(np.uint8), g.integers(0, 11, nb).astype(np.uint8)))
    res = gpu.comb_batch(pairs)
    for (x, y), got in zip(pairs, res):
        assert np.array_equal(got, oracle.comb(x, y))


def test_plain_pair_10k_matches_reference_fixture(gpu):
    c = G.configs()["plain_pair_10k"]
    a, b = W.plain_pair()
    assert W.sha256(a, b) == c["input_sha256"]
    p = gpu.semi_local_comb(a, b)
    assert W.sha256(p) == c["perm_sha256"]
    assert gpu.llcs_from_kernel(p, len(a), len(b)) == c["llcs"]


def test_ragged_batch_4096_matches_reference_fixture(gpu):
    c = G.configs()["ragged_batch_4096"]
    pairs = W.ragged_batch()
    b = gpu.Batch(pairs)
    # the default plan: a tile height per pair, the largest pairs as
    # critical chains at raised priority
    assert b.tiles()[0] == 0
    rows, prio = b.pair_tiles()
    assert set(rows.tolist()) == {4, 8, 16} and 1 <= prio.sum() <= 64
    res = b.run().results()
    bad = [i for i, p in enumerate(res) if W.sha256(p) != c["perm_sha256_each"][i]]
    assert not bad, bad[:10]
    # the same pairs, one launch per pair: identical
    res2 = b.run(per_pair=True).results()
    assert all(np.array_equal(x, y) for x, y in zip(res, res2))
    # per-pair heights with no critical chains, long queue lags, 3 waves per SIMD
    for crit, sched, w in [(0.0, 6, 8), (0.8, 4, 12), (0.4, 0, 8)]:
        b.set_critical_fraction(crit)
        b.set_schedule(sched)
        b.set_tiles(0, w)
        res3 = b.run().results()
        assert all(np.array_equal(x, y) for x, y in zip(res, res3)), (crit, sched, w)
    b.close()


def test_last_pass_heights_match_full_height_passes(gpu, oracle):
    """Per-pair heights run a pair's last pass at the smallest height that
    covers its rows.  1,024 pairs whose row counts sit just past a pass
    boundary at each height (64 R k + r) -- enough passes for the per-pair
    plan -- give the same permutations as every pass at R = 16, and a sample
    matches the oracle."""
    g = np.random.default_rng(31)
    rests = [1, 2, 63, 64, 65, 127, 255, 256, 257, 300, 511, 512, 513, 700, 1023]
    pairs = []
    for i in range(1024):
        m = 1024 * int(g.integers(1, 3)) + rests[i % len(rests)]
        n = m + int(g.integers(0, 200))
        a = g.integers(0, 20, m).astype(np.uint8)
        b = g.integers(0, 20, n).astype(np.uint8)
        pairs.append((a, b) if i % 2 else (b, a))   # either string as the rows' side
    per_pair = gpu.Batch(pairs)
    assert per_pair.tiles()[0] == 0
    rows, _ = per_pair.pair_tiles()
    assert (rows == 16).sum() >= 100, np.bincount(rows)   # pairs whose last pass can be smaller
    got = per_pair.run().results()
    full = gpu.Batch(pairs, rows_per_lane=16)
    want = full.run().results()
    bad = [i for i, (x, y) in enumerate(zip(got, want)) if not np.array_equal(x, y)]
    assert not bad, bad[:10]
    for i in g.choice(len(pairs), 8, replace=False):
        assert np.array_equal(got[i], oracle.comb(*pairs[i])), i
    per_pair.close()
    full.close()


def test_autotune_and_tiles_do_not_change_results(gpu, oracle):
    g = np.random.default_rng(15)
    pairs = [(g.integers(0, 20, int(g.integers(1, 4000))).astype(np.uint8),
              g.integers(0, 20, int(g.integers(1, 9000))).astype(np.uint8)) for _ in range(30)]
    want = [oracle.comb(x, y) for x, y in pairs]
    b = gpu.Batch(pairs)
    r, w = b.autotune(iters=1)
    assert r in (0, 4, 8, 16) and 1 <= w <= 64
    assert b.tiles() == (r, w)
    for got, exp in zip(b.run().results(), want):
        assert np.array_equal(got, exp)
    for sched in (0, 1, 2, 3, 6, 8):          # queue order: pair by pair, or the simulated schedules
        b.set_schedule(sched)
        assert b.schedule() == sched
        for r2, w2 in [(4, 8), (8, 12), (16, 16), (4, 32), (0, 8), (0, 12)]:
            b.set_tiles(r2, w2)
            for got, exp in zip(b.run().results(), want):
                assert np.array_equal(got, exp)
    b.close()


def test_large_alphabet_pair_uses_code_mode(gpu, oracle):
    # 40 symbols and 60k strands exceed the one-hot tag budget: the code-compare cell path
    g = np.random.default_rng(16)
    a = g.integers(0, 40, 20000).astype(np.uint8)
    b = g.integers(0, 40, 40000).astype(np.uint8)
    assert np.array_equal(gpu.semi_local_comb(a, b), oracle.comb(a, b))


def test_cell_modes_every_tile_height(gpu, oracle):
    # One batch that takes all three symbol tests of the cell: one-hot tags
    # (bits(m+n-1) + sigma <= 32), the per-lane row-match table (sigma <= 32,
    # bits(m+n-1) <= 32 - R) and symbol codes.  The 24000 x 42000 pair has
    # 17 id bits, so it uses the table at R = 4 and 8 but codes at R = 16.
    g = np.random.default_rng(17)

    def pair(sigma, m, n):
        x = g.integers(0, sigma, m).astype(np.uint8)
        y = g.integers(0, sigma, n).astype(np.uint8)
        return x, y

    pairs = [pair(20, 3000, 7000),     # table
             pair(32, 2500, 6000),     # table, the largest alphabet it takes
             pair(33, 2000, 5000),     # codes
             pair(20, 24000, 42000),   # table at R = 4, 8; codes at R = 16
             pair(4, 3000, 3000),      # one-hot
             pair(20, 200, 300),       # one-hot (9 id bits + 20 tags)
             pair(20, 1, 5000),        # table, a single row
             pair(20, 0, 700)]         # empty
    y = pairs[3][1]
    pairs[3] = (pairs[3][0], np.concatenate([y[:20000], pairs[3][0][:12000], y[32000:]]))   # similar stretch
    # the table is used by launches with at least one wave per CU: 300 more table pairs
    pairs += [pair(20, 300, 4000) for _ in range(300)]
    want = [oracle.comb(x, y, threads=8) if len(x) and len(y) else np.arange(len(x) + len(y), dtype=np.int32)
            for x, y in pairs]
    expect_modes = {4: {"code": 1, "onehot": 2, "table": 304}, 8: {"code": 1, "onehot": 2, "table": 304},
                    16: {"code": 2, "onehot": 2, "table": 303}}
    b = gpu.Batch(pairs, rows_per_lane=4)
    for R in (4, 8, 16):
        for w in (8, 24):
            b.set_tiles(R, w)
            assert b.cell_modes() == expect_modes[R], R
            for sched in (0, 2):
                b.set_schedule(sched)
                for k, (got, exp) in enumerate(zip(b.run().results(), want)):
                    assert np.array_equal(got, exp), (R, w, sched, k)
        for k, (got, exp) in enumerate(zip(b.run(per_pair=True).results(), want)):
            assert np.array_equal(got, exp), (R, "per pair", k)
    b.close()


def test_cell_table_with_16bit_symbols(gpu, oracle):
    # 25 distinct symbols spread over a 16-bit range: dense codes 0..24, table mode
    g = np.random.default_rng(18)
    alphabet = g.choice(np.arange(300, 60000), 25, replace=False).astype(np.int32)
    pairs = [(alphabet[g.integers(0, 25, m)], alphabet[g.integers(0, 25, n)])
             for m, n in [(4000, 9000), (700, 12000), (5000, 5000)] + [(300, 4500)] * 300]
    b = gpu.Batch(pairs, rows_per_lane=8)
    assert b.cell_modes()["table"] == 303
    want = [oracle.comb(x, y, threads=8) for x, y in pairs]
    for per_pair in (False, True):
        for k, (got, exp) in enumerate(zip(b.run(per_pair=per_pair).results(), want)):
            assert np.array_equal(got, exp), (per_pair, k)
    b.close()


def test_new_batches_in_freed_memory_match_the_oracle(gpu, oracle):
    # The strand buffer of a new batch lands in memory freed by earlier
    # batches, which holds their tagged hand-off entries.  Run tags are unique
    # across the process and a new buffer is cleared on the batch's own
    # stream, so no entry left behind is taken for a new run's.
    g = np.random.default_rng(19)
    big = [(g.integers(0, 4, 20000).astype(np.uint8), g.integers(0, 4, 30000).astype(np.uint8)),
           (g.integers(0, 4, 15000).astype(np.uint8), g.integers(0, 4, 40000).astype(np.uint8))]
    b = gpu.Batch(big, rows_per_lane=16)
    for _ in range(3):
        b.run()
    b.close()
    for k, R in enumerate((4, 8, 16, 4, 8, 16)):
        pairs = [(g.integers(0, 4, int(g.integers(2000, 9000))).astype(np.uint8),
                  g.integers(0, 4, int(g.integers(2000, 12000))).astype(np.uint8)) for _ in range(6)]
        b = gpu.Batch(pairs, rows_per_lane=R)
        for i, ((x, y), got) in enumerate(zip(pairs, b.run().results())):
            assert np.array_equal(got, oracle.comb(x, y, threads=8)), (k, R, i)
        b.close()


def test_timing_before_any_run_reports_no_stale_error(gpu, oracle):
    # A new batch's control blocks may sit in memory another handle just
    # freed (here a bit-parallel handle's scratch words).  They are cleared at
    # creation, so a timing call without warm-up as the batch's first call
    # (what the autotuner does) reports only its own runs.
    g = np.random.default_rng(20)
    for _ in range(3):
        x = g.integers(0, 2, 1 << 16).astype(np.uint8)
        y = g.integers(0, 2, 1 << 16).astype(np.uint8)
        bp = gpu.BitParallel(x, y)
        bp.llcs()
        bp.time(warmup=1, iters=2)
        bp.close()
        pairs = [(g.integers(0, 4, 3000).astype(np.uint8), g.integers(0, 4, 5000).astype(np.uint8)) for _ in range(4)]
        b = gpu.Batch(pairs)
        b.set_tiles(4, 4)
        b.time(warmup=0, iters=1)
        for (u, v), got in zip(pairs, b.results()):
            assert np.array_equal(got, oracle.comb(u, v, threads=8))
        b.close()
