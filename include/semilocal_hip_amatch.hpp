// semilocal_hip_amatch.hpp -- drop-in for the reference's semi-local duplicate
// searches (approximate_matching/ApproximateMatching.h).
//
// A call site switches by qualifying the class with semilocal_hip::
//
//   InteractiveDuplicateSearchSemiDense s;              // reference
//   semilocal_hip::InteractiveDuplicateSearchSemiDense s;   // engine
//   s.find(p, p_size, text, text_size, k, result);
//
// Replaces:
//   InteractiveDuplicateSearchSemiDense::find<T>(T *p, int p_size, T *text,
//     int text_size, double k, std::vector<Interval> &result)
//     -- ApproximateMatching.h:339-432
//   InteractiveDuplicateSearchSemiSparse::find<T>(...)  -- :252-336
// Same arguments and the same intervals, appended to `result` in the same
// order.  `result` may hold the reference's approximate_matching::utils::
// Interval or any type with public start, end and score members.  Symbols of
// 1, 2 or 4 bytes; 0 < k <= 1 and int(p_size / k) <= text_size.  Errors
// (including a device failure) throw std::runtime_error.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "semilocal_hip.h"

namespace semilocal_hip {

namespace detail {

template <class T, class IntervalT>
void duplicate_search(int variant, const T *p, int p_size, const T *text, int text_size, double k,
                      std::vector<IntervalT> &result) {
    static_assert(sizeof(T) == 1 || sizeof(T) == 2 || sizeof(T) == 4, "symbols must be 1, 2 or 4 bytes");
    sl_amatch *am = nullptr;
    int count = 0;
    int rc = sl_amatch_create(&am, p, p_size, text, text_size, (int)sizeof(T), int(p_size / k), 0);
    if (rc == SL_OK) rc = sl_duplicate_search(am, k, variant, &count);
    std::vector<int32_t> start((size_t)count), end((size_t)count);
    std::vector<double> score((size_t)count);
    if (rc == SL_OK && count > 0) rc = sl_duplicate_results(am, start.data(), end.data(), score.data());
    if (rc != SL_OK) {
        const std::string msg = sl_last_error();
        sl_amatch_destroy(am);
        throw std::runtime_error("semilocal_hip duplicate search: " + msg);
    }
    sl_amatch_destroy(am);
    for (int i = 0; i < count; ++i) {
        IntervalT iv;
        iv.start = start[(size_t)i];
        iv.end = end[(size_t)i];
        iv.score = score[(size_t)i];
        result.push_back(iv);
    }
}

}  // namespace detail

class InteractiveDuplicateSearchSemiDense {
public:
    template <class T, class IntervalT>
    void find(T *p, int p_size, T *text, int text_size, double k, std::vector<IntervalT> &result) {
        detail::duplicate_search(0, p, p_size, text, text_size, k, result);
    }
};

class InteractiveDuplicateSearchSemiSparse {
public:
    template <class T, class IntervalT>
    void find(T *p, int p_size, T *text, int text_size, double k, std::vector<IntervalT> &result) {
        detail::duplicate_search(1, p, p_size, text, text_size, k, result);
    }
};

}  // namespace semilocal_hip
