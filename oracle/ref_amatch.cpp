// ref_amatch.cpp -- TEST INFRASTRUCTURE ONLY: the reference's own duplicate
// searches (approximate_matching/ApproximateMatching.h), compiled in place
// from /root/reference by oracle/Makefile into oracle/_ref/libref_amatch.so,
// behind a C entry point.  Built without fused multiply-add contraction, as
// the reference's default x86-64 build computes its thresholds.
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
