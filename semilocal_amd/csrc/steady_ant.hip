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
// sums, two levels per pass (by the top two bits of the value within the
// segment).  Each level keeps, per element, the parent row (PR) and parent
// column (QC) it came from.
//
// Bottom-up (merge): the children's results are mapped back to the parent
// (U = union of r_lo and r_hi with a label per row), then the ant passage is
// evaluated in parallel.  With
//     delta(r, c) = #{hi points: row < r, col < c} - #{lo points: row >= r, col >= c},
// the reference's walk is a monotone staircase: at (r, c) it moves right if
// delta < 0 and up otherwise, and reports a "good point" at (r, c-1) when
// delta(r, c-1) < 0 < delta(r+1, c).  The walk is cut into segments of
// kSegSteps evaluations, one wave each.  Where it crosses anti-diagonal t
// (t = steps taken) is the first c on that diagonal with
// delta(r(c)+1, c) >= 0, r(c) = s-1-t+c (the predicate is monotone along a
// diagonal): 64 probes per round bounded by a per-subproblem table of block
// counts narrow the range to about a thousand candidates, which one pass of
// prefix sums over the diagonal's own rows and columns settles.  The wave
// then walks its segment, 64 x 64 positions per round (wave_walk).  The
// product is U with the good points written over it, which is how the
// reference merges r_hi into r_lo and adds the good points.
//
// Subproblems of at most kTile elements are solved entirely inside one block
// (split, base case and merges) in LDS.
#include "steady_ant.h"

namespace slhip {
namespace sa {

constexpr uint32_t kHiBit = 0x80000000u;
constexpr uint32_t kValMask = 0x7FFFFFFFu;

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

__device__ __forceinline__ int block_excl_scan(int v, int *tmp /* blockDim/64 + 1 ints */, int &total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int x = wave_incl_scan(v);
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (wid == 0) {   // the wave totals (at most 64) scanned by wave 0
        const int nw = (int)(blockDim.x >> 6);
        const int t = lane < nw ? tmp[lane] : 0;
        const int inc = wave_incl_scan(t);
        if (lane < nw) tmp[lane] = inc - t;
        if (lane == nw - 1) tmp[nw] = inc;
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
// A tile of kTile elements per block of kSplitThreads threads, kSplitPer
// consecutive elements per thread (one 16-byte load of P and of Qi each):
// a quarter of the waves of one element per thread, for kernels whose time
// is mostly the dispatch of their waves.  The level arrays hold whole tiles,
// so the loads stay inside them; elements past a problem's end are masked.
static_assert(kSplitPer == 4, "one int4 load per array and thread");

__global__ __launch_bounds__(kSplitThreads) void split_count(const int32_t *__restrict__ P,
                                                             const int32_t *__restrict__ Qi,
                                                             const int32_t *__restrict__ tile_prob,
                                                             const SaProb *__restrict__ probs, int S,
                                                             int32_t *__restrict__ cntP, int32_t *__restrict__ cntQ) {
    __shared__ int red[kSplitThreads / 64];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x0 = (int64_t)t * kTile + kSplitPer * threadIdx.x;
    const int64_t lx0 = x0 - pb.base;
    const int h = S >> 1;
    const int4 p4 = *reinterpret_cast<const int4 *>(P + x0);
    const int4 q4 = *reinterpret_cast<const int4 *>(Qi + x0);
    const int pv[kSplitPer] = {p4.x, p4.y, p4.z, p4.w}, qv[kSplitPer] = {q4.x, q4.y, q4.z, q4.w};
    int v = 0;   // both counts in one word: each is at most kTile < 2^16
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e)
        if (lx0 + e < pb.n) v += (pv[e] < h) | ((qv[e] < h) << 16);
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tpq = 0;
#pragma unroll
        for (int w = 0; w < kSplitThreads / 64; ++w) tpq += red[w];
        cntP[t] = tpq & 0xFFFF;
        cntQ[t] = tpq >> 16;
    }
}

// Each block first sums the counts of its segment's earlier tiles (one
// coalesced read of at most S / kTile counts and a block reduction): the
// tile's exclusive offsets, without a separate scan launch.  Segments start
// on tile boundaries (problem bases and S are multiples of kTile).
__global__ __launch_bounds__(kSplitThreads) void split_scatter(const int32_t *__restrict__ P,
                                                               const int32_t *__restrict__ Qi,
                                                               const int32_t *__restrict__ tile_prob,
                                                               const SaProb *__restrict__ probs, int S,
                                                               const int32_t *__restrict__ cntP,
                                                               const int32_t *__restrict__ cntQ,
                                                               int32_t *__restrict__ P1, int32_t *__restrict__ Qi1,
                                                               int32_t *__restrict__ PR1, int32_t *__restrict__ QC1) {
    __shared__ int tmp[kSplitThreads / 64 + 1];
    __shared__ int red[2 * (kSplitThreads / 64)];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x0 = (int64_t)t * kTile + kSplitPer * threadIdx.x;
    const int lx0 = (int)(x0 - pb.base);   // local index (problems are below 2^31 elements)
    const int lg = __builtin_ctz((unsigned)S);   // S is a power of two: shifts, not divisions
    int offp = 0, offq = 0;
    {
        const int lt0 = (int)((int64_t)t * kTile - pb.base);
        const int64_t tile0 = (pb.base + ((lt0 >> lg) << lg)) / kTile;   // base is a multiple of kTile
        int vp = 0, vq = 0;
        for (int64_t u = tile0 + threadIdx.x; u < t; u += kSplitThreads) {
            vp += cntP[u];
            vq += cntQ[u];
        }
        vp = wave_sum(vp);
        vq = wave_sum(vq);
        if ((threadIdx.x & 63) == 0) {
            red[2 * (threadIdx.x >> 6)] = vp;
            red[2 * (threadIdx.x >> 6) + 1] = vq;
        }
        __syncthreads();
#pragma unroll
        for (int w = 0; w < kSplitThreads / 64; ++w) {
            offp += red[2 * w];
            offq += red[2 * w + 1];
        }
    }
    const int h = S >> 1;
    const int4 p4 = *reinterpret_cast<const int4 *>(P + x0);
    const int4 q4 = *reinterpret_cast<const int4 *>(Qi + x0);
    const int pv[kSplitPer] = {p4.x, p4.y, p4.z, p4.w}, qv[kSplitPer] = {q4.x, q4.y, q4.z, q4.w};
    int f[kSplitPer];
    int mine = 0;   // this thread's flags, both counts in one word (each < 2^16)
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {
        f[e] = lx0 + e < pb.n ? (pv[e] < h) | ((qv[e] < h) << 16) : 0;
        mine += f[e];
    }
    int tpq;
    int ex = block_excl_scan(mine, tmp, tpq);
    const int64_t seg0 = pb.base + ((lx0 >> lg) << lg);   // a tile, so a thread's elements, lie in one segment
    const int r0 = lx0 & (S - 1);
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {
        if (lx0 + e < pb.n) {
            const int exP = (ex & 0xFFFF) + offp, exQ = (ex >> 16) + offq;
            const int fp = f[e] & 1, fq = f[e] >> 16;
            const int r = r0 + e;
            // lo elements keep their order at the front, hi elements from position h
            const int64_t dp = fp ? seg0 + exP : seg0 + h + (r - exP);
            const int64_t dq = fq ? seg0 + exQ : seg0 + h + (r - exQ);
            P1[dp] = fp ? pv[e] : pv[e] - h;
            PR1[dp] = r;
            Qi1[dq] = fq ? qv[e] : qv[e] - h;
            QC1[dq] = r;
        }
        ex += f[e];
    }
}

// Two split levels in one pass (S and S/2, S >= 4 kTile).  Both partitions
// are stable, so the elements of an S-segment end, at level S/4, ordered by
// the top two bits of their value within the segment (key k = v / (S/4))
// and then by their order at level S; at level S/2 by the top bit (k >> 1).
// An element's position at level S/2 is seg0 + (k >> 1) h + its rank among
// the segment's elements with the same top bit, at S/4 seg0 + k (S/4) + its
// rank among those with the same key: earlier tiles' counts of the key(s)
// plus the in-tile prefix.  PR at S/2 is the element's row in its
// S-segment, PR at S/4 its row in its S/2-segment (its rank by the top bit);
// P at S/4 is its value less k (S/4).  The same for Q's columns.  Counts:
// per tile, 8 words (P's keys 0-3, Q's keys 0-3).
__global__ __launch_bounds__(kSplitThreads) void split_count2(const int32_t *__restrict__ P,
                                                              const int32_t *__restrict__ Qi,
                                                              const int32_t *__restrict__ tile_prob,
                                                              const SaProb *__restrict__ probs, int S,
                                                              int32_t *__restrict__ cnt8) {
    __shared__ int red[4][kSplitThreads / 64];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x0 = (int64_t)t * kTile + kSplitPer * threadIdx.x;
    const int lx0 = (int)(x0 - pb.base);
    const int sh = __builtin_ctz((unsigned)S) - 2;   // key = value >> sh
    const int4 p4 = *reinterpret_cast<const int4 *>(P + x0);
    const int4 q4 = *reinterpret_cast<const int4 *>(Qi + x0);
    const int pv[kSplitPer] = {p4.x, p4.y, p4.z, p4.w}, qv[kSplitPer] = {q4.x, q4.y, q4.z, q4.w};
    int w[4] = {0, 0, 0, 0};   // key pairs packed 16 bits each (a tile holds at most kTile < 2^16 per key)
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e)
        if (lx0 + e < pb.n) {
            const int kp = pv[e] >> sh, kq = qv[e] >> sh;
            w[kp >> 1] += 1 << (16 * (kp & 1));
            w[2 + (kq >> 1)] += 1 << (16 * (kq & 1));
        }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int v = wave_sum(w[i]);
        if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int v = 0;
#pragma unroll
        for (int j = 0; j < kSplitThreads / 64; ++j) v += red[threadIdx.x][j];
        cnt8[(int64_t)t * 8 + 2 * threadIdx.x] = v & 0xFFFF;
        cnt8[(int64_t)t * 8 + 2 * threadIdx.x + 1] = v >> 16;
    }
}

__global__ __launch_bounds__(kSplitThreads) void split_scatter2(const int32_t *__restrict__ P,
                                                                const int32_t *__restrict__ Qi,
                                                                const int32_t *__restrict__ tile_prob,
                                                                const SaProb *__restrict__ probs, int S,
                                                                const int32_t *__restrict__ cnt8,
                                                                int32_t *__restrict__ P2, int32_t *__restrict__ Qi2,
                                                                int32_t *__restrict__ PR1, int32_t *__restrict__ QC1,
                                                                int32_t *__restrict__ PR2,
                                                                int32_t *__restrict__ QC2) {
    __shared__ int tmp[kSplitThreads / 64 + 1];
    __shared__ int red[8][kSplitThreads / 64];
    __shared__ int offs[8];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t x0 = (int64_t)t * kTile + kSplitPer * threadIdx.x;
    const int lx0 = (int)(x0 - pb.base);
    const int lg = __builtin_ctz((unsigned)S), sh = lg - 2;
    const int h = S >> 1, q = S >> 2;
    {   // counts of the segment's earlier tiles, per key (the segment starts on a tile)
        const int lt0 = (int)((int64_t)t * kTile - pb.base);
        const int64_t tile0 = (pb.base + ((lt0 >> lg) << lg)) / kTile;
        int v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (int64_t u = tile0 + threadIdx.x; u < t; u += kSplitThreads) {
            const int4 a = *reinterpret_cast<const int4 *>(cnt8 + u * 8);
            const int4 b = *reinterpret_cast<const int4 *>(cnt8 + u * 8 + 4);
            v[0] += a.x; v[1] += a.y; v[2] += a.z; v[3] += a.w;
            v[4] += b.x; v[5] += b.y; v[6] += b.z; v[7] += b.w;
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int x = wave_sum(v[i]);
            if ((threadIdx.x & 63) == 0) red[i][threadIdx.x >> 6] = x;
        }
        __syncthreads();
        if (threadIdx.x < 8) {
            int x = 0;
#pragma unroll
            for (int j = 0; j < kSplitThreads / 64; ++j) x += red[threadIdx.x][j];
            offs[threadIdx.x] = x;
        }
        __syncthreads();
    }
    const int4 p4 = *reinterpret_cast<const int4 *>(P + x0);
    const int4 q4 = *reinterpret_cast<const int4 *>(Qi + x0);
    const int pv[kSplitPer] = {p4.x, p4.y, p4.z, p4.w}, qv[kSplitPer] = {q4.x, q4.y, q4.z, q4.w};
    int kp[kSplitPer], kq[kSplitPer];
    int w[4] = {0, 0, 0, 0};
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {
        const bool in = lx0 + e < pb.n;
        kp[e] = pv[e] >> sh;
        kq[e] = qv[e] >> sh;
        if (in) {
            w[kp[e] >> 1] += 1 << (16 * (kp[e] & 1));
            w[2 + (kq[e] >> 1)] += 1 << (16 * (kq[e] & 1));
        }
    }
    int ex[4];   // this thread's in-tile exclusive prefixes, packed as w
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        int tot;
        ex[i] = block_excl_scan(w[i], tmp, tot);
    }
    const int64_t seg0 = pb.base + ((lx0 >> lg) << lg);   // a tile lies in one segment
    const int r0 = lx0 & (S - 1);
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {
        if (lx0 + e < pb.n) {
            int cp[4], cq[4];   // ranks so far by key
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cp[k] = offs[k] + ((ex[k >> 1] >> (16 * (k & 1))) & 0xFFFF);
                cq[k] = offs[4 + k] + ((ex[2 + (k >> 1)] >> (16 * (k & 1))) & 0xFFFF);
            }
            const int r = r0 + e;
            // cp[k]: the segment's elements before this one with key k, so
            // its rank by the top bit is the sum over the two keys of its pair
            const int k = kp[e], k1 = k >> 1;
            const int rank1p = cp[2 * k1] + cp[2 * k1 + 1];
            const int rank2p = cp[k];
            const int64_t pos1p = seg0 + (int64_t)k1 * h + rank1p;
            const int64_t pos2p = seg0 + (int64_t)k * q + rank2p;
            PR1[pos1p] = r;
            PR2[pos2p] = rank1p;
            P2[pos2p] = pv[e] - k * q;
            const int kk = kq[e], kk1 = kk >> 1;
            const int rank1q = cq[2 * kk1] + cq[2 * kk1 + 1];
            const int rank2q = cq[kk];
            const int64_t pos1q = seg0 + (int64_t)kk1 * h + rank1q;
            const int64_t pos2q = seg0 + (int64_t)kk * q + rank2q;
            QC1[pos1q] = r;
            QC2[pos2q] = rank1q;
            Qi2[pos2q] = qv[e] - kk * q;
        }
        // this element counted for the next ones
        if (lx0 + e < pb.n) {
            ex[kp[e] >> 1] += 1 << (16 * (kp[e] & 1));
            ex[2 + (kq[e] >> 1)] += 1 << (16 * (kq[e] & 1));
        }
    }
}

// ---------------------------------------------------------------------------
// The ant passage walk, shared by the small solver and the large levels.
// ---------------------------------------------------------------------------
// Runs at most `steps` evaluations of the reference's walk (steady_ant.h:
// 186-265, tracking delta only) from (r, c) with d = delta(r, c), in a
// subproblem of size s.  rowv(r) gives row r's point (col | hi bit 31),
// colv(c) column c's point (row | hi bit 31); good points go to put(row, col).
// Both candidates of the next move are read together, so a step costs one
// memory round trip; the point of the current row and of column c-1 are
// carried in registers.  rowv must accept r in [r_min, r] and colv c in
// [c_start - 1, min(c_start + steps, s) - 1] (c_start >= 1 or s >= 1).
//
// One step from (r, c): ur / left are the points of row r / column c-1, up /
// nxt those of row r-1 / column c (clamped reads where out of range; such a
// value is never used).  Tests (r, c-1) for a good point, then moves right
// (delta < 0) or up, updating delta by the moved-past point.  Returns false
// where the walk leaves the square.  Branch-free apart from the store of a
// good point and the exit: lanes walking different segments take different
// directions, and divergent if/else arms cost every lane both arms plus the
// exec-mask bookkeeping.
template <class PutF>
__device__ __forceinline__ bool ant_step(int &r, int &c, int &d, int s, uint32_t &ur, uint32_t &left, uint32_t up,
                                         uint32_t nxt, bool &right, PutF put) {
    {
        const int wcol = ((int)(left & kValMask) >= r) ^ (int)(left >> 31);   // hi: row < r; lo: row >= r
        const int wrow = ((int)(ur & kValMask) >= c) ^ (int)(ur >> 31);       // hi: col < c; lo: col >= c
        if (c > 0 && d - wcol < 0 && d + wrow > 0) put(r, c - 1);
    }
    // written with masks, not selects, so the compiler keeps it straight-line
    const int rt = (int)((uint32_t)d >> 31);                 // 1: move right (delta < 0), 0: up
    const uint32_t rm = 0u - (uint32_t)rt;                   // all ones when moving right
    const int stop = (rt & (c >= s)) | ((rt ^ 1) & (r == 0));
    right = rt != 0;
    if (stop) return false;
    const uint32_t v = (nxt & rm) | (up & ~rm);              // the point the move passes
    const int thr = c + ((r - c) & (int)rm);                 // right: compare its row with r; up: its column with c
    const int dd = ((int)(v & kValMask) >= thr) ^ (int)(v >> 31);
    const int up1 = rt ^ 1;
    d += (dd ^ -up1) + up1;                                  // right: +dd, up: -dd
    left = (v & rm) | (left & ~rm);
    ur = (ur & rm) | (v & ~rm);
    r -= up1;
    c += rt;
    return true;
}

template <class RowF, class ColF, class PutF>
__device__ __forceinline__ void ant_walk(int r, int c, int d, int s, int steps, RowF rowv, ColF colv, PutF put) {
    uint32_t ur = rowv(r);
    uint32_t left = c > 0 ? colv(c - 1) : 0u;
    for (int t = 0; t < steps; ++t) {
        // unconditional (clamped) reads, so both are in flight together
        const uint32_t up = rowv(r > 0 ? r - 1 : r);
        const uint32_t nxt = colv(c < s ? c : c - 1);
        bool right;
        if (!ant_step(r, c, d, s, ur, left, up, nxt, right, put)) break;
    }
}

// ---------------------------------------------------------------------------
// The walk by a whole wave, 64 x 64 positions per round.
// ---------------------------------------------------------------------------
// A window is anchored at the walk's position (r0, c0) with d0 = delta(r0,
// c0).  Position (i, j) is (r0 - i, c0 + j); lane j holds column c0 + j's
// point and bit i of a 64-bit mask stands for row r0 - 1 - i, the row an up
// move from (i, j) passes.  Then
//     delta(i, j) = D_j - popcount(RT_j & (2^i - 1)),
// D_j = delta(r0, c0 + j) (d0 plus the column terms of columns c0 .. c0+j-1
// at row r0: one ballot and a masked bit count) and RT_j the row terms of
// the window rows at column c0 + j.  A row's term changes with the column
// only at the row's own point, so RT_j is RT_0 with the bits of the rows
// whose points lie in columns [c0, c0 + j) flipped: a prefix XOR across the
// lanes of one-hot row bits.
//
// delta is nonincreasing in i and nondecreasing in j.  The walk goes up while
// delta >= 0 and right otherwise, so with tau_j the first i with
// delta(i, j) < 0 it leaves column j at row b_j = max(b_{j-1}, tau_j)
// (b_{-1} = 0): a prefix maximum across lanes, and it enters column j at
// a_j = b_{j-1}.  The position of step k of the window is (k - j, j) for the
// j with a_j + j <= k <= b_j + j.
//
// A good point of column j (delta(i, j) = 0, delta(i - 1, j) = 1 and
// delta(i, j - 1) = -1, the reference's test) can only be at i* = the first
// i with delta(i, j) <= 0 (delta moves by at most 1 per row): the D_j-th set
// bit of RT_j plus 1 (0 when D_j = 0); then tau_j is one past the next set
// bit from i*.  It is reported when the walk evaluates (i*, j): a_j <= i* and
// i* + j inside the steps this window evaluates.
//
// The window ends the segment at its step budget or where the walk leaves the
// square (moving right at column s, or up at row 0: positions evaluated and
// then the walk stops), else the walk continues from where it leaves the
// window, at the top (row i = 64 of the first column with b_j >= 64) or the
// right side (column 64 at row b_63).  Lanes outside the square (rows below
// 0, columns past s - 1) hold the value 0 and only affect positions after
// the stop there.  Every value here is the sequential walk's (ant_walk):
// the same positions, budget and good points.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp_or0(uint32_t v) {   // lanes without a source read 0
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppRowBcast15 = 0x142, kDppRowBcast31 = 0x143, kDppWaveShr1 = 0x138;

// Inclusive scans across the wave: within rows of 16 lanes by shifts of 1, 2,
// 4 and 8, then row 0's total into row 1 (and row 2's into row 3), then rows
// 0-1's total into rows 2 and 3.
__device__ __forceinline__ uint64_t wave_xor_scan64(uint64_t x) {
    uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    lo ^= dpp_or0<kDppRowShr1, 0xf>(lo);
    hi ^= dpp_or0<kDppRowShr1, 0xf>(hi);
    lo ^= dpp_or0<kDppRowShr2, 0xf>(lo);
    hi ^= dpp_or0<kDppRowShr2, 0xf>(hi);
    lo ^= dpp_or0<kDppRowShr4, 0xf>(lo);
    hi ^= dpp_or0<kDppRowShr4, 0xf>(hi);
    lo ^= dpp_or0<kDppRowShr8, 0xf>(lo);
    hi ^= dpp_or0<kDppRowShr8, 0xf>(hi);
    lo ^= dpp_or0<kDppRowBcast15, 0xa>(lo);
    hi ^= dpp_or0<kDppRowBcast15, 0xa>(hi);
    lo ^= dpp_or0<kDppRowBcast31, 0xc>(lo);
    hi ^= dpp_or0<kDppRowBcast31, 0xc>(hi);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int wave_max_scan(int x) {   // x >= 0
    x = max(x, (int)dpp_or0<kDppRowShr1, 0xf>((uint32_t)x));
    x = max(x, (int)dpp_or0<kDppRowShr2, 0xf>((uint32_t)x));
    x = max(x, (int)dpp_or0<kDppRowShr4, 0xf>((uint32_t)x));
    x = max(x, (int)dpp_or0<kDppRowShr8, 0xf>((uint32_t)x));
    x = max(x, (int)dpp_or0<kDppRowBcast15, 0xa>((uint32_t)x));
    x = max(x, (int)dpp_or0<kDppRowBcast31, 0xc>((uint32_t)x));
    return x;
}

__device__ __forceinline__ int ctz64(uint64_t x) {   // 64 when x = 0
    const int f = __ffsll((unsigned long long)x);
    return f ? f - 1 : 64;
}

// Index of the k-th (k >= 1) set bit of x, 64 when x has fewer: halves,
// quarters, ... chosen by bit counts.
__device__ __forceinline__ int select64(uint64_t x, int k) {
    const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
    int c = __popc(lo);
    int base = k > c ? 32 : 0;
    uint32_t w = k > c ? hi : lo;
    k -= k > c ? c : 0;
#pragma unroll
    for (int half = 16; half >= 1; half >>= 1) {
        c = __popc(w & ((1u << half) - 1u));
        const bool up = k > c;
        base += up ? half : 0;
        w = up ? w >> half : w;
        k -= up ? c : 0;
    }
    return (k == 1 && (w & 1u)) ? base : 64;
}

// The walk from (r0, c0) with d0 = delta(r0, c0) for at most B evaluations
// (B >= 1), by the whole wave; every argument is uniform.  rowv(r) gives row
// r's point (col | hi bit 31), colv(c) column c's point (row | hi bit 31) for
// r, c in [0, s); good points go to put(row, col).  Returns where the walk
// stands after its B evaluations, (r, c, delta(r, c)) -- the next segment's
// start -- or r = -1 where it left the square before.
template <class RowF, class ColF, class PutF>
__device__ __forceinline__ int3 wave_walk(int r0, int c0, int d0, int s, int B, RowF rowv, ColF colv, PutF put) {
    const int lane = threadIdx.x & 63;
    int3 end = make_int3(-1, 0, 0);
    for (bool more = true; more;) {
        const int rr = r0 - 1 - lane, cc = c0 + lane;
        const uint32_t u_ld = rowv(rr >= 0 ? rr : 0), v_ld = colv(cc < s ? cc : s - 1);
        const uint32_t u_r0 = rowv(r0);
        const uint32_t v_left = colv(c0 > 0 ? c0 - 1 : 0);
        const uint32_t u = rr >= 0 ? u_ld : 0u, v = cc < s ? v_ld : 0u;
        const int ucol = (int)(u & kValMask), vrow = (int)(v & kValMask);
        const uint64_t RT0 = __ballot((ucol >= c0) != (u >> 31 != 0));
        const int ih = r0 - 1 - vrow;
        const uint64_t oh = (cc < s && ih >= 0 && ih < 64) ? (1ull << ih) : 0ull;
        const uint64_t RT = RT0 ^ wave_xor_scan64(oh) ^ oh;
        const uint64_t CT = __ballot(cc < s && ((vrow >= r0) != (v >> 31 != 0)));
        const int D = d0 + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(CT >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)CT, 0u));
        const int p = select64(RT, D > 1 ? D : 1);
        const int istar = D < 0 ? 64 : (D == 0 ? 0 : (p < 64 ? p + 1 : 64));
        const int tnext = istar < 64 ? istar + 1 + ctz64(RT >> istar) : 64;
        const int tau = D < 0 ? 0 : min(tnext, 64);
        const int b = wave_max_scan(tau);
        const int a = (int)dpp_or0<kDppWaveShr1, 0xf>((uint32_t)b);
        // events, in step order
        const int jt = ctz64(__ballot(b >= 64));   // first column left at the top (64: none)
        const int b63 = __builtin_amdgcn_readlane(b, 63);
        const int X = jt < 64 ? 64 + jt : b63 + 64;   // step at which the walk leaves the window
        int kstop = 1 << 30;                          // step of a stop (evaluated, then the walk ends)
        const int js = s - c0;
        if (js >= 0 && js < 64) {
            const int bjs = __builtin_amdgcn_readlane(b, js);
            if (bjs < 64 && jt > js) kstop = bjs + js;
        }
        if (r0 < 64) {
            const uint64_t rm = __ballot(b > r0);
            if (rm) kstop = min(kstop, r0 + ctz64(rm));
        }
        const bool stops = kstop + 1 <= B;
        const int lim = stops ? kstop + 1 : min(B, X);   // steps [0, lim) are evaluated here
        // the good point of column j, at (i*, j)
        const uint32_t vl_sh = dpp_or0<kDppWaveShr1, 0xf>(v);
        const uint32_t vl = lane == 0 ? v_left : vl_sh;
        const bool rt_r0 = ((int)(u_r0 & kValMask) >= cc) != (u_r0 >> 31 != 0);
        const bool ct_l = ((int)(vl & kValMask) >= r0 - istar) != (vl >> 31 != 0);
        if (istar < 64 && a <= istar && istar + lane < lim && cc >= 1 && (istar > 0 || rt_r0) && ct_l)
            put(r0 - istar, cc - 1);
        if (stops) break;
        if (B < X) {   // the budget ends inside the window: the position of step B
            const int j = ctz64(__ballot(a + lane <= B && B <= b + lane));
            const int i = B - j;   // <= 63: step B precedes the window's exit
            const int dj = __builtin_amdgcn_readlane(D - __popcll(RT & ((1ull << i) - 1ull)), j);
            end = make_int3(r0 - i, c0 + j, dj);
            break;
        }
        // continue from the exit: the top of column jt, or the right side at row b_63
        int ie, je, dl;
        if (jt < 64) {
            ie = 64;
            je = jt;
            dl = __builtin_amdgcn_readlane(D - __popcll(RT), jt);
        } else {
            ie = b63;
            je = 64;
            const int ct = ((vrow >= r0 - b63) != (v >> 31 != 0)) ? 1 : 0;
            dl = __builtin_amdgcn_readlane(D - __popcll(RT & ((1ull << b63) - 1ull)) + ct, 63);
        }
        r0 = __builtin_amdgcn_readfirstlane(r0 - ie);
        c0 = __builtin_amdgcn_readfirstlane(c0 + je);
        d0 = __builtin_amdgcn_readfirstlane(dl);
        B = __builtin_amdgcn_readfirstlane(B - X);
        if (B == 0) {   // the budget ends exactly where the walk leaves the window
            end = make_int3(r0, c0, d0);
            more = false;
        }
    }
    return end;
}

// ---------------------------------------------------------------------------
// Small subproblems: one block solves every subproblem of size <= S0 <= kTile
// inside its tile, from the split down to size 1 and back up, in LDS.
// ---------------------------------------------------------------------------
constexpr int kSmallLevels = 10;   // log2(kTile)
constexpr int kSmallSeg = 512;     // walk steps per segment inside the small solver (one wave each)
constexpr int kSmallBoxS = 32;     // from this subproblem size on, walks are walked by whole waves
constexpr uint32_t kHi16 = 0x8000u;

// LDS points of the small solver: value | hi bit 15 (values < kTile).
__device__ __forceinline__ uint32_t widen16(uint16_t x) {
    return (uint32_t)(x & 0x7FFFu) | ((uint32_t)(x & kHi16) << 16);
}

// delta(r, c) of a subproblem held in LDS (s <= kTile rows), by one wave.
__device__ __forceinline__ int wave_delta_lds(const uint16_t *U16, int s, int r, int c) {
    const int lane = threadIdx.x & 63;
    int acc = 0;
    const int rounds = (s + 63) >> 6;   // uniform: the subproblem's own rows only
    for (int k = 0; k < rounds; ++k) {
        const int rho = lane + 64 * k;
        const uint32_t u = U16[rho < s ? rho : 0];
        const int col = (int)(u & 0x7FFFu);
        if (rho < s) acc += (u & kHi16) ? (rho < r && col < c) : -(rho >= r && col >= c);
    }
    return __builtin_amdgcn_readfirstlane(wave_sum(acc));
}

// Crossing of an anti-diagonal by the walk of an LDS subproblem: with
// R(c) = c + off, the first c in [a, b] with D(c) = delta(R(c), c) >= 0
// (b when none below b), given D(a) = Da and b - a <= kTile; returns
// (c, D(c)).  D(c+1) - D(c) reads only row R(c)'s and column c's points (the
// identity above wave_window_search), so one pass over positions
// c = a + 64 e + lane finds it: per row of 64 positions the terms (0, 1 or
// 2) go into two ballots, a lane's prefix is two masked bit counts, and the
// uniform running total carries D across rows.  D is monotone, so the pass
// stops at the first row holding a non-negative D.
__device__ __forceinline__ int2 wave_window_lds(const uint16_t *U16, const uint16_t *Ui16, int off, int a, int b,
                                                int Da) {
    const int lane = threadIdx.x & 63;
    int carry = Da;   // D at the current row's first position
    int c_ans = a;
    for (int e = 0; e < kTile / 64; ++e) {
        const int c0 = a + 64 * e;
        if (c0 >= b) break;   // uniform
        const int c = c0 + lane, R = c + off;
        const bool valid = c < b;
        const uint32_t uu = U16[valid ? R : 0], vv = Ui16[valid ? c : 0];
        const int col = (int)(uu & 0x7FFFu), row = (int)(vv & 0x7FFFu);
        const int tu = (uu & kHi16) ? (col <= c) : (col >= c);
        const int tv = (vv & kHi16) ? (row < R) : (row > R);
        const int d = valid ? tu + tv : 0;
        const uint64_t m1 = __ballot(d & 1), m2 = __ballot(d & 2);
        const int pre = (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m1 >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m1, 0u)) +
                        2 * (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m2 >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)m2, 0u));
        const uint64_t neg = __ballot(valid && carry + pre < 0);   // a prefix of the row's lanes
        c_ans += __popcll(neg);
        if (neg != __ballot(valid)) {   // the crossing is in this row
            carry += __popcll(m1 & neg) + 2 * __popcll(m2 & neg);
            break;
        }
        carry += __popcll(m1) + 2 * __popcll(m2);
    }
    return make_int2(c_ans, carry);
}

// Row r's term of delta in LDS form: delta(r+1, c) - delta(r, c).
__device__ __forceinline__ int row_term16(uint32_t u, int c) {
    return ((int)(u & 0x7FFFu) >= c) ^ (int)((u & kHi16) != 0);
}

__global__ __launch_bounds__(kTile, 8) void solve_small(const int32_t *__restrict__ P0, const int32_t *__restrict__ Qi0,
                                                     const int32_t *__restrict__ tile_prob,
                                                     const SaProb *__restrict__ probs, int S0,
                                                     int32_t *__restrict__ Rout) {
    __shared__ int16_t sP[2][kTile], sQ[2][kTile];
    __shared__ int16_t sPR[kSmallLevels][kTile], sQC[kSmallLevels][kTile];
    __shared__ int16_t sR[kTile];
    __shared__ uint16_t sU[kTile], sUi[kTile];
    __shared__ int16_t segP[kTile], segQ[kTile];
    __shared__ int sCend[kTile / 2];
    __shared__ int tmp[kTile / 64 + 1];
    const int t = blockIdx.x;
    const SaProb pb = probs[tile_prob[t]];
    const int64_t base = (int64_t)t * kTile;
    const int len = (int)min((int64_t)kTile, pb.n - (base - pb.base));
    const int x = threadIdx.x;
    const int wid = x >> 6;
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
        // segmented exclusive scans: the block scan minus its value at the
        // segment start; both flags in one scan (each count < 2^16)
        const int ex_all = block_excl_scan(fp | (fq << 16), tmp, tot);
        const int exP_all = ex_all & 0xFFFF, exQ_all = ex_all >> 16;
        if (in && r == 0) {
            segP[x] = (int16_t)exP_all;
            segQ[x] = (int16_t)exQ_all;
        }
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
            const uint16_t hb = (cid & 1) ? (uint16_t)kHi16 : (uint16_t)0;
            const int prow = sPR[lev][x];
            const int pcol = sQC[lev][cid * h + sR[x]];
            sU[seg0 + prow] = (uint16_t)pcol | hb;
            sUi[seg0 + pcol] = (uint16_t)prow | hb;
            if (!hb) atomicMax(&sCend[x / S], pcol + 1);
        }
        __syncthreads();
        if (in) sR[x] = (int16_t)(sU[x] & 0x7FFFu);
        __syncthreads();
        if (S < kSmallBoxS) {
            // short walks (< 2 kSmallBoxS steps): one lane per subproblem,
            // from the start; lane x takes subproblem x, so the walks fill
            // the first waves instead of a few lanes of every wave
            const int nsub = (len + S - 1) / S;
            if (x < nsub) {
                const int s0 = x * S, s = min(S, len - s0);
                if (s > h) {
                    const uint16_t *U16 = sU + s0, *Ui16 = sUi + s0;
                    int16_t *R = sR + s0;
                    ant_walk(s - 1, 0, (U16[s - 1] & kHi16) ? 0 : -1, s, s + sCend[x],
                             [&](int rr) { return widen16(U16[rr]); }, [&](int cc) { return widen16(Ui16[cc]); },
                             [&](int rr, int cc) { R[rr] = (int16_t)cc; });
                }
            }
        } else {
            // segments of kSmallSeg evaluations, each walked by one wave
            // (wave_walk from LDS).  A wave takes a run of consecutive
            // segments: one that continues the walk its previous segment
            // ended starts where that one ended; a subproblem's first segment
            // starts at the walk's origin (s - 1, 0); any other is found on its
            // anti-diagonal t0: D at the diagonal's lower end by an LDS scan,
            // then one window pass over the diagonal.
            const int nsps = (2 * S + kSmallSeg - 1) / kSmallSeg;   // covers the walk's <= 2s evaluations
            const int nsub = (len + S - 1) / S;
            const int nseg = nsub * nsps;
            const int per = (nseg + kTile / 64 - 1) / (kTile / 64);
            int3 cont = make_int3(-1, 0, 0);   // where this wave's previous segment ended (-1: none)
            for (int w = wid * per; w < min(nseg, (wid + 1) * per); ++w) {   // uniform over the wave
                const int sub = w / nsps, k = w - sub * nsps;
                const int s0 = sub * S, s = min(S, len - s0);
                const int t_end = s - 1 + sCend[sub];
                const int t0 = k * kSmallSeg;
                if (s <= h || t0 > t_end) {   // not split at this level, or the walk ended before t0
                    cont.x = -1;
                    continue;
                }
                const uint16_t *U16 = sU + s0, *Ui16 = sUi + s0;
                int r, c, d;
                if (k == 0) {
                    r = s - 1;
                    c = 0;
                    d = (U16[s - 1] & kHi16) ? 0 : -1;
                } else if (w > wid * per && cont.x >= 0) {
                    r = cont.x;
                    c = cont.y;
                    d = cont.z;
                } else {
                    const int off = s - t0;
                    const int lo = max(0, t0 - s + 1), hi = min(t0, s);
                    const int2 cd = wave_window_lds(U16, Ui16, off, lo, hi, wave_delta_lds(U16, s, lo + off, lo));
                    c = cd.x;
                    r = cd.x + off - 1;
                    d = cd.y - row_term16(U16[r], cd.x);
                }
                int16_t *R = sR + s0;
                cont = wave_walk(__builtin_amdgcn_readfirstlane(r), __builtin_amdgcn_readfirstlane(c),
                                 __builtin_amdgcn_readfirstlane(d), s, min(kSmallSeg, t_end - t0 + 1),
                                 [&](int rr) { return widen16(U16[rr]); }, [&](int cc) { return widen16(Ui16[cc]); },
                                 [&](int rr, int cc) { R[rr] = (int16_t)cc; });
            }
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
// c_end[seg] = 1 + largest column of a lo point (one atomic per block: a
// block's 1,024 elements, a tile, lie in one subproblem).  kSplitPer
// elements per thread, kSplitThreads apart: a store instruction then covers
// a run of consecutive child rows, whose parent rows increase (each child's
// rows come from one stable partition), so the U and R stores of a wave
// fill whole cache lines.
__global__ __launch_bounds__(kSplitThreads) void merge_map(const int32_t *__restrict__ Rc,
                                                           const int32_t *__restrict__ PRc,
                                                           const int32_t *__restrict__ QCc,
                                                           const int32_t *__restrict__ tile_prob,
                                                           const SaProb *__restrict__ probs,
                                                           const int32_t *__restrict__ seg_first, int S, int64_t ntot,
                                                           uint32_t *__restrict__ U, uint32_t *__restrict__ Ui,
                                                           int32_t *__restrict__ R, int32_t *__restrict__ c_end) {
    __shared__ int wmax[kSplitThreads / 64];
    (void)ntot;
    const int t = blockIdx.x;
    const int pid = tile_prob[t];
    const SaProb pb = probs[pid];
    const int64_t x0 = (int64_t)t * kTile + threadIdx.x;
    const int lx0 = (int)(x0 - pb.base);   // local index (problems are below 2^31 elements)
    const int lg = __builtin_ctz((unsigned)S);   // S is a power of two: shifts, not divisions
    int rv[kSplitPer], prv[kSplitPer];
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {   // whole tiles: in bounds past a problem's end too
        rv[e] = Rc[x0 + e * kSplitThreads];
        prv[e] = PRc[x0 + e * kSplitThreads];
    }
    int cand = 0;
#pragma unroll
    for (int e = 0; e < kSplitPer; ++e) {
        const int lx = lx0 + e * kSplitThreads;
        if (lx < pb.n) {
            const int cid = lx >> (lg - 1);   // child: lx / h
            const uint32_t bit = (cid & 1) ? kHiBit : 0u;
            const int64_t seg0 = pb.base + ((lx >> lg) << lg);
            const int prow = prv[e];
            const int pcol = QCc[pb.base + ((int64_t)cid << (lg - 1)) + rv[e]];
            U[seg0 + prow] = (uint32_t)pcol | bit;
            Ui[seg0 + pcol] = (uint32_t)prow | bit;
            R[seg0 + prow] = pcol;
            if (!bit) cand = max(cand, pcol + 1);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cand = max(cand, __shfl_xor(cand, off));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = cand;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = 0;
#pragma unroll
        for (int w = 0; w < kSplitThreads / 64; ++w) m = max(m, wmax[w]);
        if (m > 0) atomicMax(&c_end[seg_first[pid] + ((int)((int64_t)t * kTile - pb.base) >> lg)], m);
    }
}

__global__ __launch_bounds__(256) void table_rows(const uint32_t *__restrict__ U, const SaSeg *__restrict__ segs,
                                                  const int2 *__restrict__ rows, int32_t *__restrict__ cntHi,
                                                  int32_t *__restrict__ cntLo) {
    __shared__ int hh[kMaxNb], hl[kMaxNb];
    __shared__ int tmp[256 / 64 + 1];
    const int2 job = rows[blockIdx.x];   // (segment, block-row)
    const SaSeg sg = segs[job.x];
    const int nb = sg.nb, W = nb + 1, i = job.y;
    for (int j = threadIdx.x; j < nb; j += 256) hh[j] = hl[j] = 0;
    __syncthreads();
    const int r0 = i * kBlk, r1 = min(sg.s, r0 + kBlk);
    for (int rr = r0 + (int)threadIdx.x; rr < r1; rr += 256) {
        const uint32_t u = U[sg.x0 + rr];
        const int cb = (int)(u & kValMask) / kBlk;
        atomicAdd((u & kHiBit) ? &hh[cb] : &hl[cb], 1);
    }
    __syncthreads();
    int32_t *ch = cntHi + sg.tab_off2;
    int32_t *cl = cntLo + sg.tab_off2;
    int carryH = 0, carryL = 0, totalL = 0;
    for (int j = threadIdx.x; j < nb; j += 256) totalL += hl[j];
    {
        int tt;
        (void)block_excl_scan(totalL, tmp, tt);
        totalL = tt;
    }
    for (int j0 = 0; j0 <= nb; j0 += 256) {
        const int j = j0 + (int)threadIdx.x;
        const int vh = j < nb ? hh[j] : 0, vl = j < nb ? hl[j] : 0;
        // hi and lo in one scan: a block-row holds at most kBlk < 2^16 points
        int tt;
        const int e = block_excl_scan(vh | (vl << 16), tmp, tt);
        const int eh = e & 0xFFFF, el = e >> 16, th = tt & 0xFFFF, tl = tt >> 16;
        if (j <= nb) {
            ch[(int64_t)j * W + i] = carryH + eh;
            cl[(int64_t)j * W + i] = totalL - (carryL + el);
        }
        carryH += th;
        carryL += tl;
    }
}

// Both passes of the count tables in one block per subproblem, for
// subproblems of at most kSmallTabNb blocks per side: the hi and lo counts of
// every (block-row, block-column) cell histogrammed in LDS, then cntHi[j][i]
// = #hi with row-block < i and col-block < j (a 2D prefix) and cntLo[j][i] =
// #lo with row-block >= i and col-block >= j (a 2D suffix), written whole.
// One launch instead of table_rows + table_cols.
__global__ __launch_bounds__(256) void table_small(const uint32_t *__restrict__ U, const SaSeg *__restrict__ segs,
                                                   int32_t *__restrict__ cntHi, int32_t *__restrict__ cntLo) {
    constexpr int T = kSmallTabNb + 1;
    __shared__ int hh[T * T], hl[T * T];   // [i][j]
    const SaSeg sg = segs[blockIdx.x];
    const int nb = sg.nb, W = nb + 1;
    if (nb == 0) return;   // not split at this level: no walk, no table
    for (int k = threadIdx.x; k < T * T; k += 256) hh[k] = hl[k] = 0;
    __syncthreads();
    for (int rr = threadIdx.x; rr < sg.s; rr += 256) {
        const uint32_t u = U[sg.x0 + rr];
        const int i = rr / kBlk, j = (int)(u & kValMask) / kBlk;
        atomicAdd((u & kHiBit) ? &hh[i * T + j] : &hl[i * T + j], 1);
    }
    __syncthreads();
    // hi: prefix along j (exclusive, shifted one column right), then along i;
    // lo: suffix along j, then along i.  Row/column nb of the hi counts and
    // the lo sums past nb are empty.
    if (threadIdx.x < W) {   // one thread per block-row i
        const int i = threadIdx.x;
        int ch = 0, cl = 0;
        for (int j = 0; j < W; ++j) {
            const int h = i < nb && j < nb ? hh[i * T + j] : 0;
            hh[i * T + j] = ch;
            ch += h;
        }
        for (int j = nb; j >= 0; --j) {
            cl += i < nb && j < nb ? hl[i * T + j] : 0;
            hl[i * T + j] = cl;
        }
    }
    __syncthreads();
    if (threadIdx.x < W) {   // one thread per block-column j
        const int j = threadIdx.x;
        int ch = 0, cl = 0;
        for (int i = 0; i < W; ++i) {
            const int h = hh[i * T + j];
            hh[i * T + j] = ch;
            ch += h;
        }
        for (int i = nb; i >= 0; --i) {
            cl += hl[i * T + j];
            hl[i * T + j] = cl;
        }
    }
    __syncthreads();
    int32_t *ch = cntHi + sg.tab_off2, *cl = cntLo + sg.tab_off2;
    for (int k = threadIdx.x; k < W * W; k += 256) {
        const int j = k / W, i = k - j * W;
        ch[k] = hh[i * T + j];
        cl[k] = hl[i * T + j];
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
    // loads issued G chunks of 64 entries at a time, so a column of nb + 1
    // entries costs a few memory round trips, not one per chunk
    constexpr int G = 8;
    int total = 0;   // lo over all block-rows
    for (int i0 = 0; i0 < nb; i0 += 64 * G) {
        int v[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = i0 + 64 * g + lane;
            v[g] = i < nb ? cl[i] : 0;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) total += v[g];
    }
    total = wave_sum(total);
    int carryH = 0, carryL = 0;
    for (int i0 = 0; i0 <= nb; i0 += 64 * G) {
        int vh[G], vl[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = i0 + 64 * g + lane;
            vh[g] = i < nb ? ch[i] : 0;
            vl[g] = i < nb ? cl[i] : 0;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int i = i0 + 64 * g + lane;
            const int ih = wave_incl_scan(vh[g]), il = wave_incl_scan(vl[g]);
            if (i <= nb) {
                ch[i] = carryH + ih - vh[g];
                cl[i] = total - (carryL + il - vl[g]);
            }
            carryH += __shfl(ih, 63);
            carryL += __shfl(il, 63);
        }
    }
}

// delta(r, c) of one subproblem, evaluated by a whole wave (uniform result);
// r, c in [0, s].  The two partial scans read 16 bytes per lane and load:
// block starts are multiples of kBlk from a 1024-aligned subproblem base, and
// every problem's region is a whole number of kTile tiles, so a group of 4
// that starts below r (or c) lies inside the level arrays; groups past it are
// not read, and entries past r (or c) inside a group are masked.  All loads
// are issued before the first use.
__device__ int wave_delta(const uint32_t *__restrict__ U, const uint32_t *__restrict__ Ui,
                          const int32_t *__restrict__ cntHi, const int32_t *__restrict__ cntLo, const SaSeg &sg,
                          int r, int c) {
    const int lane = threadIdx.x & 63;
    const int i = r / kBlk, j = c / kBlk;
    const int r0 = i * kBlk, c0 = j * kBlk;
    const int64_t x0 = sg.x0;
    const int64_t tix = sg.tab_off2 + (int64_t)j * (sg.nb + 1) + i;
    const int base = cntHi[tix] - cntLo[tix];
    const uint4 *U4 = reinterpret_cast<const uint4 *>(U + x0 + r0);
    const uint4 *Ui4 = reinterpret_cast<const uint4 *>(Ui + x0 + c0);
    constexpr int kGroups = kBlk / 256;   // 4-element groups per lane
    uint4 u[kGroups], v[kGroups];
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        const int g = k * 256 + lane * 4;   // first element of the lane's group
        u[k] = r0 + g < r ? U4[k * 64 + lane] : make_uint4(0u, 0u, 0u, 0u);
        v[k] = c0 + g < c ? Ui4[k * 64 + lane] : make_uint4(0u, 0u, 0u, 0u);
    }
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kGroups; ++k) {
        const int g = k * 256 + lane * 4;
        const uint32_t uu[4] = {u[k].x, u[k].y, u[k].z, u[k].w};
        const uint32_t vv[4] = {v[k].x, v[k].y, v[k].z, v[k].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // rows [r0, r): hi with col < c, lo with col >= c
            const int col = (int)(uu[e] & kValMask);
            if (r0 + g + e < r) acc += ((col >= c) ^ (int)(uu[e] >> 31));
            // cols [c0, c): hi with row < r0, lo with row >= r0
            const int row = (int)(vv[e] & kValMask);
            if (c0 + g + e < c) acc += ((row >= r0) ^ (int)(vv[e] >> 31));
        }
    }
    return __builtin_amdgcn_readfirstlane(base + wave_sum(acc));
}

// delta(r, c) of a subproblem of at most kDirectNb * kBlk elements by one scan
// of all its rows (uniform result): aligned 16-byte groups from the
// subproblem's 1024-aligned start, entries at or past s masked (a group that
// starts below s ends at most 3 entries past it, inside the level arrays).
__device__ int wave_delta_direct(const uint32_t *__restrict__ Ux, int s, int r, int c) {
    const int lane = threadIdx.x & 63;
    constexpr int kG = kDirectNb * kBlk / 256;   // 16-byte groups per lane
    const uint4 *U4 = reinterpret_cast<const uint4 *>(Ux);
    uint4 u[kG];
#pragma unroll
    for (int k = 0; k < kG; ++k) u[k] = k * 256 + lane * 4 < s ? U4[k * 64 + lane] : make_uint4(0u, 0u, 0u, 0u);
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kG; ++k) {
        const uint32_t uu[4] = {u[k].x, u[k].y, u[k].z, u[k].w};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rho = k * 256 + lane * 4 + e;
            const int col = (int)(uu[e] & kValMask);
            if (rho < s) acc += (uu[e] >> 31) ? (rho < r && col < c ? 1 : 0) : (rho >= r && col >= c ? -1 : 0);
        }
    }
    return __builtin_amdgcn_readfirstlane(wave_sum(acc));
}

// Along an anti-diagonal, with R(c) = c + off, D(c) = delta(R(c), c) moves by
//     D(c+1) - D(c) = [row R(c)'s point: hi with col <= c, or lo with col >= c]
//                   + [column c's point: hi with row < R(c), or lo with row > R(c)]
// (the cells that enter {row < R, col < c} and leave {row >= R, col >= c}),
// so D over a short range is D(a) plus a prefix sum of terms that read only
// U[R(c)] and Ui[c], both consecutive in c.  A window pass takes PER
// positions per lane, lane L taking [w0 + PER L, w0 + PER (L + 1)) with w0 = a
// rounded down to a multiple of 4, so it settles b - a <= 64 PER - 4.
constexpr int kWinPer = 16;                   // positions per lane after a binary search
constexpr int kWinSpan = 64 * kWinPer - 4;    // b - a bound for that pass

// Lane terms of one window.  Ui is read as aligned 16-byte groups from w0;
// U from R(w0) rounded down to a multiple of 4 (M = R(w0) mod 4, so entry
// e + M is row R(w0 + PER L + e)), one group more when M > 0.  A group is
// read only when it holds a position below b, and for U when it starts at
// row >= 0 (rows below 0 belong to positions below a).  Such groups end at
// most 3 entries past row or column s - 1, inside the arrays' slack.
// Positions outside [a, b) get term 0.
template <int M, int PER>
__device__ __forceinline__ int window_terms(const uint32_t *__restrict__ Ux, const uint32_t *__restrict__ Uix,
                                            int w0, int off, int a, int b, int (&dl)[PER]) {
    constexpr int GV = PER / 4, GU = M > 0 ? GV + 1 : GV;
    const int lane = threadIdx.x & 63;
    const int p0 = w0 + lane * PER;
    const int rb = w0 + off - M + lane * PER;   // first U row read by this lane (a multiple of 4)
    uint4 u[GU], v[GV];
#pragma unroll
    for (int g = 0; g < GU; ++g) {
        const int gr = rb + 4 * g;
        u[g] = (gr >= 0 && gr - off < b) ? *reinterpret_cast<const uint4 *>(Ux + gr) : make_uint4(0u, 0u, 0u, 0u);
    }
#pragma unroll
    for (int g = 0; g < GV; ++g) {
        const int gc = p0 + 4 * g;
        v[g] = gc < b ? *reinterpret_cast<const uint4 *>(Uix + gc) : make_uint4(0u, 0u, 0u, 0u);
    }
    uint32_t ue[4 * GU], ve[PER];
#pragma unroll
    for (int g = 0; g < GU; ++g) {
        ue[4 * g] = u[g].x;
        ue[4 * g + 1] = u[g].y;
        ue[4 * g + 2] = u[g].z;
        ue[4 * g + 3] = u[g].w;
    }
#pragma unroll
    for (int g = 0; g < GV; ++g) {
        ve[4 * g] = v[g].x;
        ve[4 * g + 1] = v[g].y;
        ve[4 * g + 2] = v[g].z;
        ve[4 * g + 3] = v[g].w;
    }
    int sum = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int c = p0 + e, R = c + off;
        const uint32_t uu = ue[e + M], vv = ve[e];
        const int col = (int)(uu & kValMask), row = (int)(vv & kValMask);
        const int tu = (uu >> 31) ? (col <= c) : (col >= c);
        const int tv = (vv >> 31) ? (row < R) : (row > R);
        dl[e] = (c >= a && c < b) ? tu + tv : 0;
        sum += dl[e];
    }
    return sum;
}

// First c in [a, b] with D(c) >= 0 (b when none below b), given D(a) = Da and
// b - a <= 64 PER - 4; returns (c, D(c)).  Uniform across the wave.
template <int PER>
__device__ int2 wave_window_search(const uint32_t *__restrict__ Ux, const uint32_t *__restrict__ Uix, int off,
                                   int a, int b, int Da) {
    const int w0 = a & ~3;
    int dl[PER];
    int sum;
    switch ((w0 + off) & 3) {
    case 0: sum = window_terms<0, PER>(Ux, Uix, w0, off, a, b, dl); break;
    case 1: sum = window_terms<1, PER>(Ux, Uix, w0, off, a, b, dl); break;
    case 2: sum = window_terms<2, PER>(Ux, Uix, w0, off, a, b, dl); break;
    default: sum = window_terms<3, PER>(Ux, Uix, w0, off, a, b, dl); break;
    }
    const int p0 = w0 + (int)(threadIdx.x & 63) * PER;
    int run = Da + wave_incl_scan(sum) - sum;   // D(p0) for p0 >= a
    int neg = 0, dneg = 0;
#pragma unroll
    for (int e = 0; e < PER; ++e) {
        const int c = p0 + e;
        const int n = (c >= a && c < b && run < 0) ? 1 : 0;   // D(c) < 0 exactly on [a, answer)
        neg += n;
        dneg += n * dl[e];
        run += dl[e];
    }
    const int tot = __builtin_amdgcn_readfirstlane(wave_sum((dneg << 16) | neg));   // neg <= 1024, dneg <= 2048
    return make_int2(a + (tot & 0xFFFF), Da + (tot >> 16));
}

// Row r's term of delta: delta(r+1, c) - delta(r, c) = [hi with col < c, or lo
// with col >= c].
__device__ __forceinline__ int row_term(uint32_t u, int c) {
    return ((int)(u & kValMask) >= c) ^ (int)(u >> 31);
}

// The walk's crossing of anti-diagonal t0 of a split subproblem (t0 <= the
// walk's last step): the first c in [lo, hi] with D(c) = delta(c + off, c)
// >= 0, off = s - t0, and D(c).  Small subproblems (nb <= kDirectNb): exact
// halvings by direct scans.  Larger ones: delta is nondecreasing in r and in
// c, so the count tables' corners around (R, c) bound D(c): delta(i kBlk,
// j kBlk) <= D(c) <= delta((i+1) kBlk, (j+1) kBlk).  Each round probes 64
// positions strictly inside [a, b), one per lane, in one memory round trip:
// the last probe whose upper bound is negative and the first whose lower
// bound is not narrow the range; where every probe straddles 0, one exact
// probe halves it.  At most kWinSpan candidates are left to one window pass.
__device__ int2 diag_crossing(const uint32_t *__restrict__ U, const uint32_t *__restrict__ Ui,
                              const int32_t *__restrict__ cntHi, const int32_t *__restrict__ cntLo, const SaSeg &sg,
                              int t0) {
    const int lane = threadIdx.x & 63;
    const int s = sg.s, off = s - t0;
    const int lo = max(0, t0 - s + 1), hi = min(t0, s);
    const uint32_t *Ux = U + sg.x0;
    if (sg.nb <= kDirectNb) {
        int a = lo, b = hi;
        while (b - a > kWinSpan) {
            const int mid = (a + b) >> 1;
            if (wave_delta_direct(Ux, s, mid + off, mid) >= 0) b = mid;
            else a = mid + 1;
        }
        return wave_window_search<kWinPer>(Ux, Ui + sg.x0, off, a, b, wave_delta_direct(Ux, s, a + off, a));
    }
    const int W = sg.nb + 1;
    const int32_t *th = cntHi + sg.tab_off2, *tl = cntLo + sg.tab_off2;
    int a = lo, b = hi;
    while (b - a > kWinSpan) {
        const int p = a + (int)(((int64_t)(b - a) * (lane + 1)) / 65);   // a < p < b
        const int i = (p + off) / kBlk, j = p / kBlk;
        const int i1 = min(i + 1, sg.nb), j1 = min(j + 1, sg.nb);
        const int64_t t0i = (int64_t)j * W + i, t1i = (int64_t)j1 * W + i1;
        const int lower = th[t0i] - tl[t0i], upper = th[t1i] - tl[t1i];
        const uint64_t neg = __ballot(upper < 0), nonneg = __ballot(lower >= 0);
        const int na = neg ? __builtin_amdgcn_readlane(p, 63 - __clzll((long long)neg)) + 1 : a;
        const int nb = nonneg ? __builtin_amdgcn_readlane(p, ctz64(nonneg)) : b;
        if (na == a && nb == b) {
            const int mid = (a + b) >> 1;
            if (wave_delta(U, Ui, cntHi, cntLo, sg, mid + off, mid) >= 0) b = mid;
            else a = mid + 1;
        } else {
            a = na;
            b = nb;
        }
    }
    return wave_window_search<kWinPer>(Ux, Ui + sg.x0, off, a, b, wave_delta(U, Ui, cntHi, cntLo, sg, a + off, a));
}

// Searches and walks of a large level, one wave per segment of kSegSteps
// evaluations: the wave finds where the walk crosses the segment's first
// anti-diagonal (diag_crossing; delta(r, c) one row above the crossing is
// D(c) less row r's term) and walks the segment from there (wave_walk).
// Blocks of 4 waves; the hardware deals blocks to the 8 XCDs round-robin,
// so block b is mapped to segments in contiguous ranges per XCD and each
// XCD's L2 holds the rows and columns of its part of the level.
__global__ __launch_bounds__(256) void seg_search_walk(const uint32_t *__restrict__ U,
                                                       const uint32_t *__restrict__ Ui,
                                                       const int32_t *__restrict__ cntHi,
                                                       const int32_t *__restrict__ cntLo,
                                                       const SaSeg *__restrict__ segs,
                                                       const int2 *__restrict__ walks, int nwalks,
                                                       const int32_t *__restrict__ c_end, int32_t *__restrict__ R) {
    const int per_xcd = (int)gridDim.x / 8;   // the grid is a multiple of 8 blocks
    const int blk = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    const int w = __builtin_amdgcn_readfirstlane(blk * 4 + (int)(threadIdx.x >> 6));
    if (w >= nwalks) return;
    const int2 job = walks[w];
    const SaSeg sg = segs[job.x];
    const int s = __builtin_amdgcn_readfirstlane(sg.s);
    const int t_end = __builtin_amdgcn_readfirstlane(s - 1 + c_end[job.x]);
    const int t0 = job.y * kSegSteps;
    if (t0 > t_end) return;
    const uint32_t *Us = U + sg.x0, *Uis = Ui + sg.x0;
    const int2 cd = diag_crossing(U, Ui, cntHi, cntLo, sg, t0);
    const int c = cd.x, r = c + (s - t0) - 1;
    int32_t *Rs = R + sg.x0;
    (void)wave_walk(__builtin_amdgcn_readfirstlane(r), __builtin_amdgcn_readfirstlane(c),
                    __builtin_amdgcn_readfirstlane(cd.y - row_term(Us[r], c)), s, min(kSegSteps, t_end - t0 + 1),
                    [&](int x) { return Us[x]; }, [&](int x) { return Uis[x]; }, [&](int rr, int cc) { Rs[rr] = cc; });
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
