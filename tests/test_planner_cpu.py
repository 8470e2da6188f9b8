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
@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
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


def test_simulated_order_starts_with_the_largest_pair_and_interleaves():
    a, b = _shapes(6, 200)
    cells = a.astype(np.int64) * b
    pair, pas = S.plan_queue(a, b, 4, 8, 256, 2)
    assert pair[0] == int(np.argmax(cells)) and pas[0] == 0
    # a large pair's passes are spread over the queue, not handed out back to back
    big = int(np.argmax(cells))
    idx = np.nonzero(pair == big)[0]
    assert len(idx) > 1 and np.any(np.diff(idx) > 1)


def test_bad_arguments_are_rejected():
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([10], [10], 5)
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([10], [10], 4, schedule=4)
    with pytest.raises(S.SemiLocalError):
        S.plan_queue([-1], [10], 4)
