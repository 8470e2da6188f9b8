// ref_amatch.cpp -- TEST INFRASTRUCTURE ONLY: the reference's own duplicate
// searches (approximate_matching/ApproximateMatching.h), compiled in place
// from /root/reference by oracle/Makefile into oracle/_ref/libref_amatch.so,
// behind a C entry point.  Built without fused multiply-add contraction, as
// the reference's default x86-64 build computes its thresholds.
#include <chrono>
#include <cstdint>
#include <iostream>
#include <vector>

#include "approximate_matching/utils.h"
#include "approximate_matching/ApproximateMatching.h"

extern "C" int ref_dup_search(int variant, const int32_t *p, int m, const int32_t *t, int n, double k,
                              int32_t *start, int32_t *end, double *score, int cap) {
    std::vector<int> pp(p, p + m), tt(t, t + n);
    std::vector<approximate_matching::utils::Interval> res;
    if (variant == 0) {
        InteractiveDuplicateSearchSemiDense s;
        s.find(pp.data(), m, tt.data(), n, k, res);
    } else {
        InteractiveDuplicateSearchSemiSparse s;
        s.find(pp.data(), m, tt.data(), n, k, res);
    }
    for (size_t i = 0; i < res.size() && (int)i < cap; ++i) {
        start[i] = res[i].start;
        end[i] = res[i].end;
        score[i] = res[i].score;
    }
    return (int)res.size();
}

// The reference's own semi-local query on a braid: H(i, j) = j - i -
// dominanceSumRandom(i, j), the O(m+n) scan of
// SemiLocalStringSubstringWrapper (approximate_matching/utils.h:268-293,
// 331-334) with v = 1 (no blow-up), for nq (i, j) pairs on one permutation
// row_to_col[size] of a pair with m rows.  Used to time the reference's
// follow-up query beside the request loop's.  Returns the seconds spent in
// the queries alone (the braid is held as a Permutation beforehand, as a
// caller of the reference would hold it).
extern "C" double ref_semi_local_H(const int32_t *row_to_col, int size, int m, int nq, const int32_t *qi,
                                   const int32_t *qj, int64_t *out) {
    Permutation perm(size, size);
    for (int r = 0; r < size; ++r) perm.set_point(r, row_to_col[r]);
    approximate_matching::utils::SemiLocalStringSubstringWrapper w(&perm, nullptr, m, 1);
    const auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < nq; ++k) {
        const int range = w.dominanceSumRandom(qi[k], qj[k]);
        out[k] = (int64_t)qj[k] - qi[k] - range;
    }
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
