"""CPU checks of two formulations behind kernel (c) (steady_ant.hip), each
against the sequential definition it replaces, on small random inputs.

* The ant passage walked by a whole wave, 64 x 64 positions per round
  (wave_walk): a line-for-line Python port of its window formulas -- row
  terms by a ballot and a prefix XOR of one-hot row bits, the turning row of
  each column by a bit select, the path as a prefix maximum, good points
  where delta first reaches 0, stops, budgets and window exits -- against
  the sequential walk of the reference's ant_passage
  (unit_monge_mult/steady_ant.h:186-265: move right while delta < 0, up
  otherwise; a good point at (r, c-1) when delta(r, c-1) < 0 < delta(r+1,
  c)).  Whole walks and segments of 1 to 10^6 steps, on labelled
  permutations with few, half and many hi points.
* Two split levels in one pass (split_count2 / split_scatter2): positions by
  the top two bits of the value within the segment, against two single
  stable partitions (get_vertical_slice / get_horizontal_slice,
  steady_ant.h:100-136).

The GPU kernels themselves are checked bit for bit against the reference's
products in tests/test_mult_gpu.py; this file pins the derivations.
"""
import numpy as np
import pytest

W = 64
M64 = (1 << 64) - 1


def _colterm(Ui_row, Ui_hi, c, r):   # a right move past column c at row r
    return int((Ui_row[c] >= r) ^ Ui_hi[c])


def _rowterm(U_col, U_hi, r, c):     # an up move past row r at column c
    return int((U_col[r] >= c) ^ U_hi[r])


def walk_sequential(U_col, U_hi, Ui_row, Ui_hi, s, r, c, d, steps):
    """The reference's walk, step by step: (good points, end state, evaluations, stopped)."""
    puts = []
    for t in range(steps):
        if c > 0 and d - _colterm(Ui_row, Ui_hi, c - 1, r) < 0 and d + _rowterm(U_col, U_hi, r, c) > 0:
            puts.append((r, c - 1))
        right = d < 0
        if (right and c >= s) or (not right and r == 0):
            return puts, None, t + 1, True
        if right:
            d += _colterm(Ui_row, Ui_hi, c, r)
            c += 1
        else:
            d -= _rowterm(U_col, U_hi, r - 1, c)
            r -= 1
    return puts, (r, c, d), steps, False


def _popc(x):
    return bin(x & M64).count("1")


def _ctz64(x):
    x &= M64
    return 64 if x == 0 else (x & -x).bit_length() - 1


def _select64(x, k):
    """As the kernel: index of the k-th set bit by halving bit counts, 64 if fewer."""
    lo, hi = x & 0xFFFFFFFF, (x >> 32) & 0xFFFFFFFF
    c = _popc(lo)
    base, w = (32, hi) if k > c else (0, lo)
    k -= c if k > c else 0
    half = 16
    while half >= 1:
        c = _popc(w & ((1 << half) - 1))
        if k > c:
            base, w, k = base + half, w >> half, k - c
        half >>= 1
    return base if (k == 1 and (w & 1)) else 64


def walk_wave(U_col, U_hi, Ui_row, Ui_hi, s, r0, c0, d0, B):
    """Port of wave_walk: lists stand for the 64 lanes."""
    puts, evals = [], 0
    while True:
        L = range(W)
        u = [(U_col[r0 - 1 - l], U_hi[r0 - 1 - l]) if r0 - 1 - l >= 0 else (0, 0) for l in L]
        v = [(Ui_row[c0 + l], Ui_hi[c0 + l]) if c0 + l < s else (0, 0) for l in L]
        u_r0 = (U_col[r0], U_hi[r0])
        v_left = (Ui_row[c0 - 1], Ui_hi[c0 - 1]) if c0 > 0 else (Ui_row[0], Ui_hi[0])
        RT0 = sum(1 << l for l in L if (u[l][0] >= c0) != (u[l][1] != 0))
        oh = [(1 << (r0 - 1 - v[l][0])) if (c0 + l < s and 0 <= r0 - 1 - v[l][0] < W) else 0 for l in L]
        RT, acc = [], 0
        for l in L:                       # exclusive prefix XOR
            RT.append(RT0 ^ acc)
            acc ^= oh[l]
        CT = sum(1 << l for l in L if c0 + l < s and ((v[l][0] >= r0) != (v[l][1] != 0)))
        D = [d0 + _popc(CT & ((1 << l) - 1)) for l in L]
        istar, tau = [], []
        for l in L:
            p = _select64(RT[l], D[l] if D[l] > 1 else 1)
            ist = 64 if D[l] < 0 else (0 if D[l] == 0 else (p + 1 if p < 64 else 64))
            tn = ist + 1 + _ctz64(RT[l] >> ist) if ist < 64 else 64
            istar.append(ist)
            tau.append(0 if D[l] < 0 else min(tn, 64))
        b, m = [], 0
        for t in tau:                     # inclusive prefix maximum
            m = max(m, t)
            b.append(m)
        a = [0] + b[:-1]
        jt = _ctz64(sum(1 << l for l in L if b[l] >= 64))
        X = 64 + jt if jt < 64 else b[63] + 64
        kstop = 1 << 30
        js = s - c0
        if 0 <= js < 64 and b[js] < 64 and jt > js:
            kstop = b[js] + js
        if r0 < 64:
            rm = sum(1 << l for l in L if b[l] > r0)
            if rm:
                kstop = min(kstop, r0 + _ctz64(rm))
        stops = kstop + 1 <= B
        lim = kstop + 1 if stops else min(B, X)
        for l in L:
            vl = v_left if l == 0 else v[l - 1]
            cc = c0 + l
            rt_r0 = (u_r0[0] >= cc) != (u_r0[1] != 0)
            ct_l = (vl[0] >= r0 - istar[l]) != (vl[1] != 0)
            if istar[l] < 64 and a[l] <= istar[l] and istar[l] + l < lim and cc >= 1 and \
                    (istar[l] > 0 or rt_r0) and ct_l:
                puts.append((r0 - istar[l], cc - 1))
        evals += lim
        if stops:
            return puts, None, evals, True
        if B < X:
            j = next(l for l in L if a[l] + l <= B <= b[l] + l)
            i = B - j
            return puts, (r0 - i, c0 + j, D[j] - _popc(RT[j] & ((1 << i) - 1))), evals, False
        if jt < 64:
            ie, je, dl = 64, jt, D[jt] - _popc(RT[jt])
        else:
            ie, je = b[63], 64
            ct = 1 if ((v[63][0] >= r0 - b[63]) != (v[63][1] != 0)) else 0
            dl = D[63] - _popc(RT[63] & ((1 << b[63]) - 1)) + ct
        r0, c0, d0, B = r0 - ie, c0 + je, dl, B - X
        if B == 0:
            return puts, (r0, c0, d0), evals, False


def _labelled(g, s, hi_frac):
    perm = [int(x) for x in g.permutation(s)]
    hi = [int(x) for x in (g.random(s) < hi_frac)]
    Ui_row, Ui_hi = [0] * s, [0] * s
    for r in range(s):
        Ui_row[perm[r]], Ui_hi[perm[r]] = r, hi[r]
    return perm, hi, Ui_row, Ui_hi


def test_select64_equals_the_kth_set_bit():
    g = np.random.default_rng(5)
    for _ in range(2000):
        x = int(g.integers(0, 1 << 62)) * int(g.integers(1, 4)) & M64
        k = int(g.integers(1, 70))
        bits = [i for i in range(64) if (x >> i) & 1]
        assert _select64(x, k) == (bits[k - 1] if k <= len(bits) else 64)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_wave_walk_equals_the_sequential_walk(seed):
    g = np.random.default_rng(seed)
    for _ in range(150):
        s = int(g.choice([1, 2, 5, 63, 64, 65, 130, 300, 700]))
        U_col, U_hi, Ui_row, Ui_hi = _labelled(g, s, float(g.choice([0.1, 0.5, 0.9])))
        d0 = 0 if U_hi[s - 1] else -1
        lo_cols = [U_col[r] for r in range(s) if not U_hi[r]]
        total = s + (max(lo_cols) + 1 if lo_cols else 0)   # the walk's evaluations (t_end + 1)
        L = int(g.choice([1, 7, 64, 127, 128, 512, 10 ** 6]))
        r, c, d, left = s - 1, 0, d0, total
        while left > 0:                   # segments start where the sequential walk stands
            st = min(L, left)
            want = walk_sequential(U_col, U_hi, Ui_row, Ui_hi, s, r, c, d, st)
            got = walk_wave(U_col, U_hi, Ui_row, Ui_hi, s, r, c, d, st)
            assert sorted(got[0]) == sorted(want[0]) and got[1:] == want[1:], (s, L, r, c, d)
            if want[3]:
                break
            left -= st
            r, c, d = want[1]


def _split_once(P, S):
    out, pr, h = np.empty_like(P), np.empty_like(P), S // 2
    for s0 in range(0, len(P), S):
        seg = P[s0:s0 + S]
        lo = [i for i, x in enumerate(seg) if x < h]
        hi = [i for i, x in enumerate(seg) if x >= h]
        for j, i in enumerate(lo):
            out[s0 + j], pr[s0 + j] = seg[i], i
        for j, i in enumerate(hi):
            out[s0 + h + j], pr[s0 + h + j] = seg[i] - h, i
    return out, pr


def _split_twice_by_keys(P, S):
    n, h, q = len(P), S // 2, S // 4
    P2, PR1, PR2 = (np.full(n, -1) for _ in range(3))
    for s0 in range(0, n, S):
        cnt = [0, 0, 0, 0]
        for r, x in enumerate(P[s0:s0 + S]):
            k, k1 = x // q, x // q >> 1
            rank1, rank2 = cnt[2 * k1] + cnt[2 * k1 + 1], cnt[k]
            pos1, pos2 = s0 + k1 * h + rank1, s0 + k * q + rank2
            assert s0 <= pos1 < min(n, s0 + S) and s0 <= pos2 < min(n, s0 + S)
            PR1[pos1], PR2[pos2], P2[pos2] = r, rank1, x - k * q
            cnt[k] += 1
    return P2, PR1, PR2


def test_two_split_levels_by_key_equal_two_single_splits():
    g = np.random.default_rng(9)
    for _ in range(200):
        S = int(g.choice([16, 32, 64]))
        n = int(g.integers(1, 5 * S))
        P = np.concatenate([g.permutation(min(S, n - s0)) for s0 in range(0, n, S)])
        a1, pr1 = _split_once(P, S)
        a2, pr2 = _split_once(a1, S // 2)
        b2, qr1, qr2 = _split_twice_by_keys(P, S)
        assert np.array_equal(a2, b2) and np.array_equal(pr1, qr1) and np.array_equal(pr2, qr2)
