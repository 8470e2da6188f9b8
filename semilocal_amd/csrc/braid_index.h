// braid_index.h -- wavelet-matrix index over a resident braid (row -> col),
// for O(log(m+n)) H(i, j) / string-substring / substring-string queries.
#pragma once
#include "common.h"

namespace slhip {

// Device view of one index: `levels` levels of `nw` 64-bit words (bit
// levels-1-l of every value of level l's arrangement), prefix counts of ones
// per word (nw + 1 per level) and the number of zeros per level.
struct BraidIndexView {
    const uint64_t *words = nullptr;
    const uint32_t *ranks = nullptr;
    const uint32_t *zeros = nullptr;
    int32_t m = 0, n = 0;
    int32_t levels = 0;
    uint32_t nw = 0;
};

size_t braid_index_bytes(int64_t size);           // index memory for a kernel of size m+n
size_t braid_index_scratch_bytes(int64_t size);   // temporary memory of one build
hipError_t braid_index_build(const int32_t *rc, int32_t m, int32_t n, void *mem, void *scratch,
                             hipStream_t stream, BraidIndexView *view);
hipError_t braid_index_answer(const BraidIndexView &ix, int nq, const int32_t *kind, const int32_t *qi,
                              const int32_t *qj, int64_t *out, hipStream_t stream);
// ones among the first i bits of a level
__device__ __forceinline__ uint32_t wm_rank1(const uint64_t *words, const uint32_t *ranks, uint32_t i) {
    const uint32_t w = i >> 6, b = i & 63u;
    uint32_t r = ranks[w];
    if (b) r += (uint32_t)__builtin_popcountll(words[w] & ((1ull << b) - 1ull));
    return r;
}

// #{r in [lo, hi) : P[r] < J}: one rank step per level (rank-less-than walk)
__device__ __forceinline__ uint32_t wm_count_less(const BraidIndexView &ix, uint32_t lo, uint32_t hi, int64_t J) {
    if (J <= 0 || lo >= hi) return 0;
    if (J >= ((int64_t)1 << ix.levels)) return hi - lo;   // every value is below J
    uint32_t cnt = 0;
    for (int l = 0; l < ix.levels; ++l) {
        const int bit = ix.levels - 1 - l;
        const uint64_t *words = ix.words + (size_t)l * ix.nw;
        const uint32_t *ranks = ix.ranks + (size_t)l * (ix.nw + 1);
        const uint32_t r1lo = wm_rank1(words, ranks, lo), r1hi = wm_rank1(words, ranks, hi);
        if ((J >> bit) & 1) {
            cnt += (hi - r1hi) - (lo - r1lo);   // a 0 here below equal higher bits: less than J
            lo = ix.zeros[l] + r1lo;
            hi = ix.zeros[l] + r1hi;
        } else {
            lo -= r1lo;
            hi -= r1hi;
        }
    }
    return cnt;
}

// Query kinds of include/semilocal_hip.h: 0 raw H(i, j); 1 LCS(a, b[i:j)) =
// H(i, j); 2 LCS(a[i:j), b) = H(-i, n+m-j) - i - (m-j), with
// H(i, j) = j - i - #{r >= i + m : P[r] < j} (approximate_matching/utils.h:268-359).
__device__ __forceinline__ int64_t wm_query(const BraidIndexView &ix, int kind, int32_t qi, int32_t qj) {
    int64_t hi_ = qi, hj = qj;
    if (kind == 2) {
        hi_ = -(int64_t)qi;
        hj = (int64_t)ix.n + ix.m - qj;
    }
    const int64_t size = (int64_t)ix.m + ix.n;
    int64_t r0 = hi_ + ix.m;
    r0 = r0 < 0 ? 0 : (r0 > size ? size : r0);
    const int64_t h = hj - hi_ - (int64_t)wm_count_less(ix, (uint32_t)r0, (uint32_t)size, hj);
    return kind == 2 ? h - qi - (ix.m - qj) : h;
}

}  // namespace slhip
