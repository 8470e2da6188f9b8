// steady_ant.h -- layout of the level-by-level braid multiplication engine
// (kernel (c)) shared by its kernels and its host planner.
#pragma once
#include "common.h"

namespace slhip {
namespace sa {

constexpr int kTile = 1024;      // elements per block of the split / small-solver kernels
constexpr int kBlk = 1024;       // block size of the per-subproblem count tables
constexpr int kSegSteps = 1024;  // walk steps per segment of the large levels (one wave each)
constexpr int kWalkSlack = 64;   // entries allocated past the end of U and U^-1 (the searches' 16-byte groups)
constexpr int kSplitPer = 4;     // consecutive elements per thread of the large-level split kernels
constexpr int kSplitThreads = kTile / kSplitPer;   // their block size (one tile per block)
constexpr int kMaxNb = 4096;     // count-table blocks per side: subproblems up to kMaxNb * kBlk
constexpr int kSmallTabNb = 8;   // count tables of at most this many blocks per side: built by one block
constexpr int kDirectNb = 2;     // subproblems of at most kDirectNb * kBlk elements: no count tables, the
                                 // segment searches scan the subproblem

// One product of the batch: elements [base, base + n) of every level array.
// base is a multiple of kTile, so every tile belongs to exactly one problem.
struct SaProb {
    int64_t base;
    int32_t n;
    int32_t pad;
};

// One subproblem of a level (large levels only).
struct SaSeg {
    int64_t x0;        // first element
    int64_t tab_off2;  // offset of its (nb+1) x (nb+1) count tables
    int32_t s;         // size
    int32_t tile0;     // first tile
    int32_t ntiles;    // tiles
    int32_t nb;        // count-table blocks per side; 0 when the subproblem is not split
};

// One staggered multiplication (glue of two adjacent block braids).
struct StagJob {
    int64_t p_off, q_off;   // offsets of p and q (row -> col) in the block pool
    int64_t qcr_off;        // offset of q's inverse in the scratch array
    int64_t out_off;        // offset of the product in the output pool
    int32_t pr, qr, k;      // sizes of p, q and of the shared core
    int32_t row_glue;       // 1: q above p (vertical neighbours), 0: p left of q
    int32_t prob;           // problem index of the core in the multiply plan
    int32_t tileP0, ntilesP, tileQ0, ntilesQ;
    int32_t pad;
};

}  // namespace sa
}  // namespace slhip
