"""Host-side queue planning (sl_plan_queue, no device needed): the tickets a
batched launch hands out.  Every pass of every pair exactly once, the passes
of a pair in pass order (a pass only ever waits on a pass handed out before
it, so the persistent launch cannot deadlock), and the pair-by-pair order
largest pairs first."""
import numpy as np
import pytest

import semilocal_amd as S


def _shapes(seed, n):
    g = np.random.default_rng(seed)
    a = np.exp(g.normal(np.log(2000), 1.0, n)).clip(0, 50000).astype(np.int32)
    b = np.exp(g.normal(np.log(2000), 1.0, n)).clip(0, 50000).astype(np.int32)
    a[::17] = 0                                  # empty pairs get no tickets
    return a, b


@pytest.mark.parametrize("rows_per_lane", [4, 8, 16])
@pytest.mark.parametrize("schedule", [0, 1, 2, 3, 6, 8])
@pytest.mark.parametrize("waves_per_cu", [4, 32])
def test_every_pass_once_in_pass_order(rows_per_lane, schedule, waves_per_cu):
    a, b = _shapes(rows_per_lane * 10 + schedule, 300)
    pair, pas = S.plan_queue(a, b, rows_per_lane, waves_per_cu, 256, schedule)
    m = np.minimum(a, b).astype(np.int64)
    npasses = np.where((a > 0) & (b > 0), (m + 64 * rows_per_lane - 1) // (64 * rows_per_lane), 0)
    assert len(pair) == npasses.sum()
    want = {(i, p) for i in range(len(a)) for p in range(npasses[i])}
    got = list(zip(pair.tolist(), pas.tolist()))
    assert len(set(got)) == len(got) and set(got) == want
    last = {}
    for i, p in got:                             # passes of a pair in order
        assert p == last.get(i, -1) + 1
        last[i] = p


def _per_pair_passes(a, b, pair, pas):
    """Passes per pair of a per-pair plan, checked to be one of the three tile heights' counts."""
    m = np.minimum(a, b).astype(np.int64)
    got = np.zeros(len(a), np.int64)
    last = {}
    for i, p in zip(pair.tolist(), pas.tolist()):   # every pass once, in pass order
        assert p == last.get(i, -1) + 1
        last[i] = p
        got[i] += 1
    for i in range(len(a)):
        if a[i] == 0 or b[i] == 0:
            assert got[i] == 0
        else:
            assert got[i] in {(m[i] + 64 * R - 1) // (64 * R) for R in (4, 8, 16)}, (i, m[i], got[i])
    return got


@pytest.mark.parametrize("schedule", [0, 2, 3, 6])
@pytest.mark.parametrize("waves_per_cu", [8, 12])
def test_per_pair_tiles_every_pass_once_in_pass_order(schedule, waves_per_cu):
    a, b = _shapes(40 + schedule, 3000)
    pair, pas = S.plan_queue(a, b, 0, waves_per_cu, 256, schedule)
    _per_pair_passes(a, b, pair, pas)


def test_per_pair_tiles_follow_padding_and_critical_chains():
    # A large batch: each pair takes the tile height with the fewest
    # instructions for its padded rows, and the pairs whose chain of passes
    # would outlast the batch take the shortest chain (R = 4, most passes).
    a, b = _shapes(7, 4096)
    a[:3] = 50000
    b[:3] = 50000                                # three 50,000^2 pairs: the critical chains
    pair, pas = S.plan_queue(a, b, 0, 8, 256, 3)
    got = _per_pair_passes(a, b, pair, pas)
    assert list(got[:3]) == [(50000 + 255) // 256] * 3
    m = np.minimum(a, b).astype(np.int64)
    # rows just above a tall tile boundary are not padded to 1,024 rows
    small = [i for i in range(3, len(a)) if 1025 <= m[i] <= 1100 and max(a[i], b[i]) < 4000]
    assert small and all(got[i] > 1 for i in small)
    # a pair of exactly 1,024 rows, short columns, takes one R = 16 pass
    a2, b2 = a.copy(), b.copy()
    a2[3], b2[3] = 1024, 1500
    pair2, pas2 = S.plan_queue(a2, b2, 0, 8, 256, 3)
    assert _per_pair_passes(a2, b2, pair2, pas2)[3] == 1


def test_small_per_pair_batch_takes_one_height():
    # too few passes to give every SIMD two waves: one height for every pair
    a = np.array([3000, 5000, 200, 9000], np.int32)
    b = np.array([4000, 5000, 7000, 9000], np.int32)
    pair, pas = S.plan_queue(a, b, 0, 8, 256, 2)
    got = _per_pair_passes(a, b, pair, pas)
    m = np.minimum(a, b)
    heights = [R for R in (4, 8, 16) if all(got[i] == (m[i] + 64 * R - 1) // (64 * R) for i in range(len(a)))]
    assert heights, got


def test_pair_by_pair_order_is_largest_first():
    a, b = _shapes(5, 200)
    pair, pas = S.plan_queue(a, b, 8, 8, 256, 0)
    cells = a.astype(np.int64) * b
    starts = pair[pas == 0]
    assert np.all(np.diff(cells[starts]) <= 0)
    # consecutive passes of one pair
    for i in np.unique(pair):
        idx = np.nonzero(pair == i)[0]
        assert np.all(np.diff(idx) == 1)


def test_simulated_order_starts_with_the_longest_chain_and_interleaves():
    # critical-path list scheduling: the first ticket is the first pass of the
    # pair whose chain (n + 64 steps for its last pass, plus the lag of 256
    # steps at schedule 2 before each later pass) is the longest
    a, b = _shapes(6, 200)
    m = np.minimum(a, b).astype(np.int64)
    n = np.maximum(a, b).astype(np.int64)
    passes = np.where((a > 0) & (b > 0), (m + 255) // 256, 0)
    chain = np.where(passes > 0, n + 64 + (passes - 1) * 256, -1)
    pair, pas = S.plan_queue(a, b, 4, 8, 256, 2)
    assert pair[0] == int(np.argmax(chain)) and pas[0] == 0
    # a large pair's passes are spread over the queue, not handed out back to back
    big = int(np.argmax(chain))
    idx = np.nonzero(pair == big)[0]
    assert len(idx) > 1 and np.any(np.diff(idx) > 1)


def test_a_skinny_pair_of_one_long_pass_is_not_left_for_the_end():
    # 600 x 40,000 has few cells but one 40,000-step pass: it starts before
    # the many short pairs of more cells
    a = np.full(2000, 3000, np.int32)
    b = np.full(2000, 3000, np.int32)
    a[1000], b[1000] = 600, 40000
    pair, pas = S.plan_queue(a, b, 8, 8, 256, 2)
    assert int(np.nonzero(pair == 1000)[0][0]) == 0


def test_bad_arguments_are_rejected():
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([10], [10], 5)
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([10], [10], 4, schedule=9)
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([-1], [10], 4)
