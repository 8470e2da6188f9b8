// steady_ant.hip -- kernel (c): unit-Monge (sticky braid) multiplication
// R = P (.) Q of n x n permutations, level by level, for a batch of
// independent products at once.
//
// Reference: distance_unit_monge_product::steady_ant::steady_ant_optimized_seq
// (unit_monge_mult/steady_ant.h:378-453) with get_vertical_slice (:100-112),
// get_horizontal_slice (:123-136), inverse_mapping (:146-160) and ant_passage
// (:186-265).  The product is unique, so the recursion may split the shared
// dimension anywhere; this engine splits a subproblem of aligned size S at
// h = S/2.  All problems of a batch share one sequence of segment sizes
// S = 2^K, 2^(K-1), ..., 1: at level S, problem p (elements [base_p,
// base_p + n_p) of every level array, base_p a multiple of kTile) consists of
// the subproblems [base_p + j S, base_p + min((j+1) S, n_p)).  A subproblem of
// size <= S/2 is not split at that level (all of it is "lo").
//
// Top-down (split): P's rows are stably partitioned by col < h (vertical
// slices) and Q's columns by row < h (horizontal slices): segmented prefix
// sums.  Each level keeps, per element, the parent row (PR) and parent column
// (QC) it came from.
//
// Bottom-up (merge): the children's results are mapped back to the parent
// (U = union of r_lo and r_hi with a label per row), then the ant passage is
// evaluated in parallel.  With
//     delta(r, c) = #{hi points: row < r, col < c} - #{lo points: row >= r, col >= c},
// the reference's walk is a monotone staircase: at (r, c) it moves right if
// delta < 0 and up otherwise, and reports a "good point" at (r, c-1) when
// delta(r, c-1) < 0 < delta(r+1, c).  The walk is cut into segments of
// kSegSteps evaluations.  Where it crosses anti-diagonal t (t = steps taken)
// is the first c on that diagonal with delta(r(c)+1, c) >= 0, r(c) = s-1-t+c,
// found by binary search (the predicate is monotone along a diagonal) from a
// per-subproblem table of block counts plus two short scans.  Each segment is
// then walked by one lane from LDS.  The product is U with the good points
// written over it, which is how the reference merges r_hi into r_lo and adds
// the good points.
//
// Subproblems of at most kTile elements are solved entirely inside one block
// (split, base case and merges) in LDS.
#include "steady_ant.h"

namespace slhip {
namespace sa {

constexpr uint32_t kHiBit = 0x80000000u;
constexpr uint32_t kValMask = 0x7FFFFFFFu;

__device__ __forceinline__ int block_excl_scan(int v, int *tmp /* blockDim/64 + 1 ints */, int &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            const int t = tmp[w];
            tmp[w] = acc;
            acc += t;
        }
        tmp[blockDim.x >> 6] = acc;
    }
    __syncthreads();
    total = tmp[blockDim.x >> 6];
    const int r = tmp[wid] + x - v;
    __syncthreads();
    return r;
}

__device__ __forceinline__ int wave_sum(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(v, off);
        if (lane >= off) v += y;
    }
    return v;
}

// ---------------------------------------------------------------------------
// Loading: P row->col as given; Q, given row->col, is stored inverted.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void load_pq(const int32_t *__restrict__ p_cat, const int32_t *__restrict__ q_cat,
                                                 const int64_t *__restrict__ src_off,
                                                 const int32_t *__restrict__ tile_prob,
                                                 const SaProb *__restrict__ probs, int32_t *__restrict__ P,
                                                 int32_t *__restrict__ Qi) {
    const int t = blockIdx.x;
    const int pid = tile_prob[t];
    const SaProb pb = probs[pid];
    const int64_t x = (int64_t)t * kTile + threadIdx.x;
    const int64_t lx = x - pb.base;
    if (lx >= pb.n) return;
    const int64_t so = src_off[pid];
    P[x] = p_cat[so + lx];
    Qi[pb.base + q_cat[so + lx]] = (int32_t)lx;
}

__global__ __launch_bounds__(kTile) void store_r(const int32_t *__restrict__ R, const int32_t *__restrict__ tile_prob,
                                                 const SaProb *__restrict__ probs,
                                                 const int64_t *__restrict__ dst_off, int32_t *__restrict__ r_cat) {
    const int t = blockIdx.x;
    const int pid = tile_prob[t];
    const SaProb pb = probs[pid];
    const int64_t x = (int64_t)t * kTile + threadIdx.x;
    const int64_t lx = x - pb.base;
    if (lx >= pb.n) return;
    r_cat[dst_off[pid] + lx] = R[x];
}

// ---------------------------------------------------------------------------
// Split, large levels (segment size S > kTile; a tile lies in one segment).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kTile) void split_count(const int32_t *__restrict__ P, const int32_t *__restrict__ Qi,
                                                     const int32_t *__restrict__ tile_prob,
                                                     const SaProb *__restrict__ probs, int S,
                                                     int32_t *__restrict__ cntP, int32_t *__restrict__ cntQ) {
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x = (int64_t)t * kTile + threadIdx.x;
    const bool in = x - pb.base < pb.n;
    const int h = S >> 1;
    int fp = 0, fq = 0;
    if (in) {
        fp = P[x] < h;
        fq = Qi[x] < h;
    }
    int tp, tq;
    (void)block_excl_scan(fp, tmp, tp);
    (void)block_excl_scan(fq, tmp, tq);
    if (threadIdx.x == 0) {
        cntP[t] = tp;
        cntQ[t] = tq;
    }
}

// Exclusive scan of the tile counts of one segment (one block per segment).
__global__ __launch_bounds__(kTile) void seg_tile_scan(const SaSeg *__restrict__ segs, int32_t *__restrict__ cntP,
                                                       int32_t *__restrict__ cntQ) {
    __shared__ int tmp[kTile / 64 + 1];
    const SaSeg sg = segs[blockIdx.x];
    for (int which = 0; which < 2; ++which) {
        int32_t *cnt = which == 0 ? cntP : cntQ;
        int carry = 0;
        for (int b0 = 0; b0 < sg.ntiles; b0 += kTile) {
            const int k = b0 + (int)threadIdx.x;
            const bool ok = k < sg.ntiles;
            const int v = ok ? cnt[sg.tile0 + k] : 0;
            int tot;
            const int ex = block_excl_scan(v, tmp, tot);
            if (ok) cnt[sg.tile0 + k] = carry + ex;
            carry += tot;
        }
    }
}

__global__ __launch_bounds__(kTile) void split_scatter(const int32_t *__restrict__ P, const int32_t *__restrict__ Qi,
                                                       const int32_t *__restrict__ tile_prob,
                                                       const SaProb *__restrict__ probs, int S,
                                                       const int32_t *__restrict__ offP,
                                                       const int32_t *__restrict__ offQ, int32_t *__restrict__ P1,
                                                       int32_t *__restrict__ Qi1, int32_t *__restrict__ PR1,
                                                       int32_t *__restrict__ QC1) {
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x = (int64_t)t * kTile + threadIdx.x;
    const int64_t lx = x - pb.base;
    const bool in = lx < pb.n;
    const int h = S >> 1;
    int fp = 0, fq = 0, p = 0, q = 0;
    if (in) {
        p = P[x];
        q = Qi[x];
        fp = p < h;
        fq = q < h;
    }
    int tp, tq;
    const int exP = block_excl_scan(fp, tmp, tp) + offP[t];
    const int exQ = block_excl_scan(fq, tmp, tq) + offQ[t];
    if (in) {
        const int64_t seg0 = pb.base + (lx / S) * S;
        const int r = (int)(lx % S);
        // lo elements keep their order at the front, hi elements from position h
        const int64_t dp = fp ? seg0 + exP : seg0 + h + (r - exP);
        const int64_t dq = fq ? seg0 + exQ : seg0 + h + (r - exQ);
        P1[dp] = fp ? p : p - h;
        PR1[dp] = r;
        Qi1[dq] = fq ? q : q - h;
        QC1[dq] = r;
    }
}

// ---------------------------------------------------------------------------
// Small subproblems: one block solves every subproblem of size <= S0 <= kTile
// inside its tile, from the split down to size 1 and back up, in LDS.
// ---------------------------------------------------------------------------
constexpr int kSmallLevels = 10;   // log2(kTile)

// The reference's ant passage over one subproblem held in LDS (size s),
// tracking delta only; good points are written over R.
__device__ void walk_small(const int16_t *U, const uint8_t *hi, const int16_t *Ui, int s, int c_end, int16_t *R) {
    int r = s - 1, c = 0;
    int d = hi[r] ? 0 : -1;
    const int t_end = s - 1 + c_end;
    for (int t = 0; t <= t_end; ++t) {
        const int urow = U[r];
        const bool hrow = hi[r] != 0;
        if (c > 0) {
            const int pr = Ui[c - 1];
            const int wcol = hi[pr] ? (pr < r) : (pr >= r);
            const int wrow = hrow ? (urow < c) : (urow >= c);
            if (d - wcol < 0 && d + wrow > 0) R[r] = (int16_t)(c - 1);
        }
        if (d < 0) {
            if (c >= s) break;
            const int pr = Ui[c];
            d += hi[pr] ? (pr < r) : (pr >= r);
            ++c;
        } else {
            if (r == 0) break;
            --r;
            const int ur = U[r];
            d -= hi[r] ? (ur < c) : (ur >= c);
        }
    }
}

__global__ __launch_bounds__(kTile) void solve_small(const int32_t *__restrict__ P0, const int32_t *__restrict__ Qi0,
                                                     const int32_t *__restrict__ tile_prob,
                                                     const SaProb *__restrict__ probs, int S0,
                                                     int32_t *__restrict__ Rout) {
    __shared__ int16_t sP[2][kTile], sQ[2][kTile];
    __shared__ int16_t sPR[kSmallLevels][kTile], sQC[kSmallLevels][kTile];
    __shared__ int16_t sR[kTile], sU[kTile], sUi[kTile];
    __shared__ int16_t segP[kTile], segQ[kTile];
    __shared__ uint8_t sHi[kTile];
    __shared__ int sCend[kTile / 2];
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t base = (int64_t)t * kTile;
    const int len = (int)min((int64_t)kTile, pb.n - (base - pb.base));
    const int x = threadIdx.x;
    const bool in = x < len;
    if (in) {
        sP[0][x] = (int16_t)P0[base + x];
        sQ[0][x] = (int16_t)Qi0[base + x];
    }
    __syncthreads();
    int cur = 0, lev = 0;
    for (int S = S0; S > 1; S >>= 1, ++lev) {
        const int h = S >> 1;
        const int seg0 = (x / S) * S;
        const int r = x - seg0;
        const int p = in ? sP[cur][x] : 0, q = in ? sQ[cur][x] : 0;
        const int fp = in && p < h, fq = in && q < h;
        int tot;
        // segmented exclusive scans: the block scan minus its value at the segment start
        const int exP_all = block_excl_scan(fp, tmp, tot);
        if (in && r == 0) segP[x] = (int16_t)exP_all;
        const int exQ_all = block_excl_scan(fq, tmp, tot);
        if (in && r == 0) segQ[x] = (int16_t)exQ_all;
        __syncthreads();
        if (in) {
            const int exP = exP_all - segP[seg0];
            const int exQ = exQ_all - segQ[seg0];
            const int dp = fp ? seg0 + exP : seg0 + h + (r - exP);
            const int dq = fq ? seg0 + exQ : seg0 + h + (r - exQ);
            sP[cur ^ 1][dp] = (int16_t)(fp ? p : p - h);
            sPR[lev][dp] = (int16_t)r;
            sQ[cur ^ 1][dq] = (int16_t)(fq ? q : q - h);
            sQC[lev][dq] = (int16_t)r;
        }
        __syncthreads();
        cur ^= 1;
    }
    if (in) sR[x] = 0;   // every subproblem now has size 1; its product is [0]
    __syncthreads();
    for (int S = 2; S <= S0; S <<= 1) {
        --lev;
        const int h = S >> 1;
        const int seg0 = (x / S) * S;
        if (in && x == seg0) sCend[x / S] = 0;
        __syncthreads();
        if (in) {
            const int cid = x / h;
            const int is_hi = cid & 1;
            const int prow = sPR[lev][x];
            const int pcol = sQC[lev][cid * h + sR[x]];
            sU[seg0 + prow] = (int16_t)pcol;
            sHi[seg0 + prow] = (uint8_t)is_hi;
            sUi[seg0 + pcol] = (int16_t)prow;
            if (!is_hi) atomicMax(&sCend[x / S], pcol + 1);
        }
        __syncthreads();
        if (in) sR[x] = sU[x];
        __syncthreads();
        if (in && x == seg0) {
            const int s = min(S, len - x);
            if (s > h) walk_small(sU + x, sHi + x, sUi + x, s, sCend[x / S], sR + x);
        }
        __syncthreads();
    }
    if (in) Rout[base + x] = sR[x];
}

// ---------------------------------------------------------------------------
// Merge, large levels (segment size S >= 2 kTile).
// ---------------------------------------------------------------------------
// Children (level S/2) -> parent (level S): U[parent row] = parent col | hi
// bit, Ui[parent col] = parent row | hi bit, R = U without the bit, and
// c_end[seg] = 1 + largest column of a lo point.
__global__ void merge_map(const int32_t *__restrict__ Rc, const int32_t *__restrict__ PRc,
                          const int32_t *__restrict__ QCc, const int32_t *__restrict__ tile_prob,
                          const SaProb *__restrict__ probs, const int32_t *__restrict__ seg_first, int S,
                          int64_t ntot, uint32_t *__restrict__ U, uint32_t *__restrict__ Ui,
                          int32_t *__restrict__ R, int32_t *__restrict__ c_end) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ntot) return;
    const int pid = tile_prob[x / kTile];
    const SaProb pb = probs[pid];
    const int64_t lx = x - pb.base;
    if (lx >= pb.n) return;
    const int h = S >> 1;
    const int64_t cid = lx / h;
    const uint32_t bit = (cid & 1) ? kHiBit : 0u;
    const int64_t seg0 = pb.base + (lx / S) * S;
    const int prow = PRc[x];
    const int pcol = QCc[pb.base + cid * h + Rc[x]];
    U[seg0 + prow] = (uint32_t)pcol | bit;
    Ui[seg0 + pcol] = (uint32_t)prow | bit;
    R[seg0 + prow] = pcol;
    if (!bit) atomicMax(&c_end[seg_first[pid] + lx / S], pcol + 1);
}

// hist*[seg][bi * nb + bj] counts the points with row in block bi and column
// in block bj (split segments only: nb > 0).
__global__ void table_hist(const uint32_t *__restrict__ U, const int32_t *__restrict__ tile_prob,
                           const SaProb *__restrict__ probs, const int32_t *__restrict__ seg_first,
                           const SaSeg *__restrict__ segs, int S, int64_t ntot, int32_t *__restrict__ histHi,
                           int32_t *__restrict__ histLo) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ntot) return;
    const int pid = tile_prob[x / kTile];
    const SaProb pb = probs[pid];
    const int64_t lx = x - pb.base;
    if (lx >= pb.n) return;
    const SaSeg sg = segs[seg_first[pid] + lx / S];
    if (sg.nb == 0) return;
    const int r = (int)(lx % S);
    const uint32_t u = U[x];
    const int c = (int)(u & kValMask);
    const int64_t cell = sg.tab_off + (int64_t)(r / kBlk) * sg.nb + c / kBlk;
    atomicAdd((u & kHiBit) ? &histHi[cell] : &histLo[cell], 1);
}

// Row pass, one wave per (segment, block-row i): along block-columns j, hi
// gets an exclusive prefix and lo an inclusive suffix.  Written transposed:
// T[seg][j * (nb+1) + i] for j in [0, nb].
__global__ void table_rows(const int32_t *__restrict__ histHi, const int32_t *__restrict__ histLo,
                           const SaSeg *__restrict__ segs, const int2 *__restrict__ rows, int nrows,
                           int32_t *__restrict__ cntHi, int32_t *__restrict__ cntLo) {
    const int w = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= nrows) return;
    const int lane = threadIdx.x & 63;
    const int2 job = rows[w];   // (segment, block-row)
    const SaSeg sg = segs[job.x];
    const int nb = sg.nb, W = nb + 1, i = job.y;
    const int32_t *hh = histHi + sg.tab_off + (int64_t)i * nb;
    const int32_t *hl = histLo + sg.tab_off + (int64_t)i * nb;
    int32_t *ch = cntHi + sg.tab_off2;
    int32_t *cl = cntLo + sg.tab_off2;
    int carry = 0;
    for (int j0 = 0; j0 <= nb; j0 += 64) {
        const int j = j0 + lane;
        const int v = j < nb ? hh[j] : 0;
        const int inc = wave_incl_scan(v);
        if (j <= nb) ch[(int64_t)j * W + i] = carry + inc - v;
        carry += __shfl(inc, 63);
    }
    int total = 0;
    for (int j0 = 0; j0 < nb; j0 += 64) {
        const int j = j0 + lane;
        total += wave_sum(j < nb ? hl[j] : 0);
    }
    carry = 0;
    for (int j0 = 0; j0 <= nb; j0 += 64) {
        const int j = j0 + lane;
        const int v = j < nb ? hl[j] : 0;
        const int inc = wave_incl_scan(v);
        if (j <= nb) cl[(int64_t)j * W + i] = total - (carry + inc - v);
        carry += __shfl(inc, 63);
    }
}

// Column pass, one wave per (segment, block-column j): along block-rows i of
// T[j], hi gets an exclusive prefix and lo an inclusive suffix, in place.
// Entry i = nb (not written by the row pass) is defined here.  Afterwards
// cntHi[seg][j][i] = #hi with row-block < i and col-block < j, and
// cntLo[seg][j][i] = #lo with row-block >= i and col-block >= j.
__global__ void table_cols(const SaSeg *__restrict__ segs, const int2 *__restrict__ cols, int ncols,
                           int32_t *__restrict__ cntHi, int32_t *__restrict__ cntLo) {
    const int w = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= ncols) return;
    const int lane = threadIdx.x & 63;
    const int2 job = cols[w];   // (segment, block-column)
    const SaSeg sg = segs[job.x];
    const int nb = sg.nb, W = nb + 1, j = job.y;
    int32_t *ch = cntHi + sg.tab_off2 + (int64_t)j * W;
    int32_t *cl = cntLo + sg.tab_off2 + (int64_t)j * W;
    int carry = 0;
    for (int i0 = 0; i0 <= nb; i0 += 64) {
        const int i = i0 + lane;
        const int v = i < nb ? ch[i] : 0;
        const int inc = wave_incl_scan(v);
        if (i <= nb) ch[i] = carry + inc - v;
        carry += __shfl(inc, 63);
    }
    int total = 0;
    for (int i0 = 0; i0 < nb; i0 += 64) {
        const int i = i0 + lane;
        total += wave_sum(i < nb ? cl[i] : 0);
    }
    carry = 0;
    for (int i0 = 0; i0 <= nb; i0 += 64) {
        const int i = i0 + lane;
        const int v = i < nb ? cl[i] : 0;
        const int inc = wave_incl_scan(v);
        if (i <= nb) cl[i] = total - (carry + inc - v);
        carry += __shfl(inc, 63);
    }
}

// delta(r, c) of one subproblem, evaluated by a whole wave (uniform result);
// r, c in [0, s].
__device__ int wave_delta(const uint32_t *__restrict__ U, const uint32_t *__restrict__ Ui,
                          const int32_t *__restrict__ cntHi, const int32_t *__restrict__ cntLo, const SaSeg &sg,
                          int r, int c) {
    const int lane = threadIdx.x & 63;
    const int i = r / kBlk, j = c / kBlk;
    const int r0 = i * kBlk, c0 = j * kBlk;
    const int64_t x0 = sg.x0;
    int acc = 0;
    // rows [r0, r): hi with col < c, lo with col >= c
#pragma unroll
    for (int k = 0; k < kBlk / 64; ++k) {
        const int rr = r0 + k * 64 + lane;
        if (rr < r) {
            const uint32_t u = U[x0 + rr];
            const int col = (int)(u & kValMask);
            acc += (u & kHiBit) ? (col < c) : (col >= c);
        }
    }
    // cols [c0, c): hi with row < r0, lo with row >= r0
#pragma unroll
    for (int k = 0; k < kBlk / 64; ++k) {
        const int cc = c0 + k * 64 + lane;
        if (cc < c) {
            const uint32_t v = Ui[x0 + cc];
            const int row = (int)(v & kValMask);
            acc += (v & kHiBit) ? (row < r0) : (row >= r0);
        }
    }
    acc = wave_sum(acc);
    const int64_t tix = sg.tab_off2 + (int64_t)j * (sg.nb + 1) + i;
    return __builtin_amdgcn_readfirstlane(cntHi[tix] - cntLo[tix] + acc);
}

// Start point of each walk segment: one wave per entry (subproblem, k); walk
// segment k covers steps [k kSegSteps, (k+1) kSegSteps).  Output
// (r, c, delta(r, c), t0), or r = -1 when the walk ends before step t0.
__global__ void seg_search(const uint32_t *__restrict__ U, const uint32_t *__restrict__ Ui,
                           const int32_t *__restrict__ cntHi, const int32_t *__restrict__ cntLo,
                           const SaSeg *__restrict__ segs, const int2 *__restrict__ walks, int nwalks,
                           const int32_t *__restrict__ c_end, int4 *__restrict__ seg_start) {
    const int w = (int)((blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6);
    if (w >= nwalks) return;
    const int2 job = walks[w];
    const SaSeg sg = segs[job.x];
    const int s = sg.s;
    const int t_end = s - 1 + c_end[job.x];
    const int t0 = job.y * kSegSteps;
    int4 out = make_int4(-1, -1, 0, 0);
    if (t0 <= t_end) {
        int a = max(0, t0 - s + 1), b = min(t0, s);
        while (a < b) {
            const int mid = (a + b) >> 1;
            const int rr = s - 1 - t0 + mid;
            if (wave_delta(U, Ui, cntHi, cntLo, sg, rr + 1, mid) >= 0) b = mid;
            else a = mid + 1;
        }
        const int c = a, r = s - 1 - t0 + c;
        out = make_int4(r, c, wave_delta(U, Ui, cntHi, cntLo, sg, r, c), t0);
    }
    if ((threadIdx.x & 63) == 0) seg_start[w] = out;
}

// Walk of one segment: the wave stages the rows and columns the walk can
// reach in LDS, lane 0 walks and writes good points over R.
__global__ __launch_bounds__(64) void seg_walk(const uint32_t *__restrict__ U, const uint32_t *__restrict__ Ui,
                                               const SaSeg *__restrict__ segs, const int2 *__restrict__ walks,
                                               const int32_t *__restrict__ c_end,
                                               const int4 *__restrict__ seg_start, int32_t *__restrict__ R) {
    __shared__ uint32_t rowsU[kSegSteps + 1], colsU[kSegSteps + 1];
    const int w = blockIdx.x;
    const int lane = threadIdx.x;
    const int4 st = seg_start[w];
    if (st.x < 0) return;   // uniform over the block (one wave)
    const int2 job = walks[w];
    const SaSeg sg = segs[job.x];
    const int64_t x0 = sg.x0;
    const int s = sg.s;
    const int t_end = s - 1 + c_end[job.x];
    // kSegSteps steps move up at most kSegSteps rows and right at most
    // kSegSteps columns; column c-1 is read before the first move.
    const int r_lo = max(0, st.x - kSegSteps);
    const int nrow = st.x - r_lo + 1;
    const int c_lo = max(0, st.y - 1);
    const int ncol = max(0, min(s, st.y + kSegSteps) - c_lo);
    for (int k = lane; k < nrow; k += 64) rowsU[k] = U[x0 + r_lo + k];
    for (int k = lane; k < ncol; k += 64) colsU[k] = Ui[x0 + c_lo + k];
    __syncthreads();
    if (lane != 0) return;
    int r = st.x, c = st.y, d = st.z;
    const int t_stop = min(t_end, st.w + kSegSteps - 1);
    for (int t = st.w; t <= t_stop; ++t) {
        const uint32_t ur = rowsU[r - r_lo];
        const int urow = (int)(ur & kValMask);
        const bool hrow = (ur & kHiBit) != 0;
        if (c > 0) {
            const uint32_t vc = colsU[c - 1 - c_lo];
            const int pr = (int)(vc & kValMask);
            const bool hp = (vc & kHiBit) != 0;
            const int wcol = hp ? (pr < r) : (pr >= r);
            const int wrow = hrow ? (urow < c) : (urow >= c);
            if (d - wcol < 0 && d + wrow > 0) R[x0 + r] = c - 1;
        }
        if (d < 0) {
            if (c >= s) break;
            const uint32_t vc = colsU[c - c_lo];
            const int pr = (int)(vc & kValMask);
            const bool hp = (vc & kHiBit) != 0;
            d += hp ? (pr < r) : (pr >= r);
            ++c;
        } else {
            if (r == 0) break;
            --r;
            const uint32_t u2 = rowsU[r - r_lo];
            const int ur2 = (int)(u2 & kValMask);
            const bool h2 = (u2 & kHiBit) != 0;
            d -= h2 ? (ur2 < c) : (ur2 >= c);
        }
    }
}

__global__ void fill_i32(int32_t *__restrict__ p, int32_t v, int64_t count) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// Staggered multiplication (steady_ant.h:609-737): slice the k shared strands
// of two adjacent block braids into the multiply layout, and glue afterwards.
// Job tiles: [tileP0, tileP0 + ntilesP) cover p's rows, [tileQ0, tileQ0 +
// ntilesQ) cover q's columns (tileQ0 = tileP0 + ntilesP).
// ---------------------------------------------------------------------------
// Inverse of every job's q (col -> row).
__global__ __launch_bounds__(kTile) void stag_qinv(const int32_t *__restrict__ pool, const StagJob *__restrict__ jobs,
                                                   const int32_t *__restrict__ tile_job,
                                                   int32_t *__restrict__ qcr) {
    const int t = blockIdx.x;
    const StagJob jb = jobs[tile_job[t]];
    if (t < jb.tileQ0) return;
    const int i = (t - jb.tileQ0) * kTile + threadIdx.x;
    if (i >= jb.qr) return;
    qcr[jb.qcr_off + pool[jb.q_off + i]] = i;
}

// Flag of element e of tile t: p's rows with col in the shared range, q's
// columns with row in the shared range (get_vertical_slice /
// get_horizontal_slice, steady_ant.h:100-136).
__device__ __forceinline__ bool stag_flag(const int32_t *__restrict__ pool, const int32_t *__restrict__ qcr,
                                          const StagJob &jb, int t, int e, int &val, int &src) {
    if (t < jb.tileQ0) {
        const int x = (t - jb.tileP0) * kTile + e;
        if (x >= jb.pr) return false;
        const int c0 = jb.row_glue ? 0 : jb.pr - jb.k;
        const int v = pool[jb.p_off + x];
        val = v - c0;
        src = x;
        return v >= c0 && v < c0 + jb.k;
    }
    const int c = (t - jb.tileQ0) * kTile + e;
    if (c >= jb.qr) return false;
    const int r0 = jb.row_glue ? jb.qr - jb.k : 0;
    const int v = qcr[jb.qcr_off + c];
    val = v - r0;
    src = c;
    return v >= r0 && v < r0 + jb.k;
}

__global__ __launch_bounds__(kTile) void stag_count(const int32_t *__restrict__ pool,
                                                    const int32_t *__restrict__ qcr,
                                                    const StagJob *__restrict__ jobs,
                                                    const int32_t *__restrict__ tile_job,
                                                    int32_t *__restrict__ cnt) {
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const StagJob jb = jobs[tile_job[t]];
    int val = 0, src = 0;
    const int f = stag_flag(pool, qcr, jb, t, threadIdx.x, val, src) ? 1 : 0;
    int tot;
    (void)block_excl_scan(f, tmp, tot);
    if (threadIdx.x == 0) cnt[t] = tot;
}

// Exclusive scan of the tile counts of each (job, p or q) range: block 2i
// scans job i's p tiles, block 2i+1 its q tiles.
__global__ __launch_bounds__(kTile) void stag_scan(const StagJob *__restrict__ jobs, int32_t *__restrict__ cnt) {
    __shared__ int tmp[kTile / 64 + 1];
    const StagJob jb = jobs[blockIdx.x >> 1];
    const int first = (blockIdx.x & 1) ? jb.tileQ0 : jb.tileP0;
    const int n = (blockIdx.x & 1) ? jb.ntilesQ : jb.ntilesP;
    int carry = 0;
    for (int b0 = 0; b0 < n; b0 += kTile) {
        const int k = b0 + (int)threadIdx.x;
        const bool ok = k < n;
        const int v = ok ? cnt[first + k] : 0;
        int tot;
        const int ex = block_excl_scan(v, tmp, tot);
        if (ok) cnt[first + k] = carry + ex;
        carry += tot;
    }
}

// Writes the k x k core into the multiply layout: P (row -> col) and Qi
// (col -> row) at the job's problem base, with the maps back to the blocks.
__global__ __launch_bounds__(kTile) void stag_scatter(const int32_t *__restrict__ pool,
                                                      const int32_t *__restrict__ qcr,
                                                      const StagJob *__restrict__ jobs,
                                                      const int32_t *__restrict__ tile_job,
                                                      const int32_t *__restrict__ off,
                                                      const SaProb *__restrict__ probs, int32_t *__restrict__ P,
                                                      int32_t *__restrict__ Qi, int32_t *__restrict__ map_row,
                                                      int32_t *__restrict__ map_col) {
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const StagJob jb = jobs[tile_job[t]];
    int val = 0, src = 0;
    const bool f = stag_flag(pool, qcr, jb, t, threadIdx.x, val, src);
    int tot;
    const int ex = block_excl_scan(f ? 1 : 0, tmp, tot) + off[t];
    if (!f) return;
    const int64_t base = probs[jb.prob].base;
    if (t < jb.tileQ0) {
        P[base + ex] = val;
        map_row[base + ex] = src;
    } else {
        Qi[base + ex] = val;
        map_col[base + ex] = src;
    }
}

// Glue (steady_ant.h:695-729): one thread per element of [0, k + pr + qr).
__global__ void stag_glue(const int32_t *__restrict__ pool, const StagJob *__restrict__ jobs,
                          const SaProb *__restrict__ probs, const int32_t *__restrict__ R,
                          const int32_t *__restrict__ map_row, const int32_t *__restrict__ map_col,
                          int32_t *__restrict__ out) {
    const StagJob jb = jobs[blockIdx.y];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = jb.k, pr = jb.pr, qr = jb.qr;
    if (e >= k + pr + qr) return;
    int32_t *prod = out + jb.out_off;
    if (e < k) {
        const int64_t base = probs[jb.prob].base;
        const int row = map_row[base + e];
        const int col = map_col[base + R[base + e]];
        if (jb.row_glue) prod[qr - k + row] = col;
        else prod[row] = col + pr - k;
    } else if (e < k + pr) {
        const int x = e - k;
        const int px = pool[jb.p_off + x];
        if (jb.row_glue) {
            if (px >= k) prod[x + qr - k] = px + qr - k;
        } else if (px < pr - k) {
            prod[x] = px;
        }
    } else {
        const int i = e - k - pr;
        const int qi = pool[jb.q_off + i];
        if (jb.row_glue) {
            if (i < qr - k) prod[i] = qi;
        } else if (i >= k) {
            prod[i - k + pr] = pr - k + qi;
        }
    }
}

// Copies whole blocks that pass through a merge level unmerged; one
// blockIdx.y per move (src offset, dst offset, length).
__global__ void copy_blocks(const int32_t *__restrict__ src, const int64_t *__restrict__ moves,
                            int32_t *__restrict__ dst) {
    const int64_t so = moves[3 * blockIdx.y], dO = moves[3 * blockIdx.y + 1], len = moves[3 * blockIdx.y + 2];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x)
        dst[dO + e] = src[so + e];
}

}  // namespace sa
}  // namespace slhip
