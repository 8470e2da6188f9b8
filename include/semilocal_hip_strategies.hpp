// semilocal_hip_strategies.hpp -- the engine behind the reference's
// namespaced strategy interfaces.
//
// Include AFTER the reference's own headers (this file derives from their
// classes and does not include them itself):
//     include/SemiLocal/SemiLocalLCS.h   semi_local::SemiLocalStrategy<Input>,
//                                        semi_local::AntidiagonalIterativeCombing<Input, WithIf>
//     common/include/Permutations/MongeMultuplicationStrategy.h
//                                        common::BraidMultiplicationStrategy
//     include/LCS/LCS.h                  lcs::LCSStrategy<Input>
//     include/BitwiseLCS/BitParalellLCS.h (optional; bit_parallel::lcs::BitParallel)
// Results are written the way the reference's strategies write them: the
// callee reassigns `kernel = common::Permutation(m + n, m + n)` and fills it
// (SemiLocalLCS.h:117, 259).  compute() is noexcept like the reference's, so a
// device failure terminates (std::terminate) instead of returning a wrong
// kernel.
#pragma once

#include <cmath>
#include <exception>
#include <iostream>

#include "semilocal_hip.hpp"

#ifndef SEMILOCAL_HIP_NO_BASE_DEFINITIONS
// The reference declares these base members virtual but never defines them
// (SemiLocalLCS.h:7-11, LCS.h:9-13).  A build that emits the base vtables (any
// unoptimised build constructing a derived strategy, the reference's own
// included) then fails to link.  They act as pure virtual members: defining
// them here, as a call that terminates, lets every build link.  Define
// SEMILOCAL_HIP_NO_BASE_DEFINITIONS if another header already defines them.
template <class Input>
void semi_local::SemiLocalStrategy<Input>::compute(const Input *, int, const Input *, int,
                                                   common::Permutation &) noexcept {
    std::terminate();
}
template <class Input>
int lcs::LCSStrategy<Input>::compute(const Input *, int, const Input *, int) {
    std::terminate();
}
#endif

namespace semilocal_hip {

namespace detail {
inline void store_common(common::Permutation &kernel, const std::vector<int32_t> &rc) {
    kernel = common::Permutation((int)rc.size(), (int)rc.size());
    for (size_t i = 0; i < rc.size(); ++i) kernel.set((int)i, rc[i]);
}

[[noreturn]] inline void fail(const std::exception &e) {
    std::cerr << e.what() << std::endl;
    std::terminate();
}
}  // namespace detail

// Replaces semi_local::OpenMPIterativeCombing<Input, WithIf> (SemiLocalLCS.h:
// 231-306) and is accepted wherever the reference takes an
// AntidiagonalIterativeCombing<Input, WithIf> (its hybrids, :482, :597-600).
template <class Input, bool WithIf = false>
class HipIterativeCombing : public semi_local::AntidiagonalIterativeCombing<Input, WithIf> {
  public:
    explicit HipIterativeCombing(int numThreads = 1) : threads(numThreads) {}

    void compute(const Input *a, int aSize, const Input *b, int bSize, common::Permutation &kernel) noexcept override {
        try {
            detail::store_common(kernel, detail::comb(a, aSize, b, bSize));
        } catch (const std::exception &e) {
            detail::fail(e);
        }
        kernel.m = aSize;
        kernel.n = bSize;
    }

  private:
    int threads;   // accepted for signature parity; the device decides its own parallelism
};

// Replaces semi_local::OpenMPIterativeCombingCombined<Input, WithIf>
// (SemiLocalLCS.h:308-476): same constructor; same unique kernel.
template <class Input, bool WithIf = false>
class HipIterativeCombingCombined : public HipIterativeCombing<Input, WithIf> {
  public:
    HipIterativeCombingCombined(int numThreads, common::StaggeredStickyBraidMultiplication *staggered = nullptr)
        : HipIterativeCombing<Input, WithIf>(numThreads) {
        (void)staggered;
    }
};

// Replaces semi_local::OpenMPHybridRecursive<Input, WithIf, UseSumBound,
// UseDepth> (SemiLocalLCS.h:478-585): same template parameters and
// constructor.  Its thresholds only decide where the CPU recursion stops; the
// device combs the pair whole (same unique kernel).
template <class Input, bool WithIf = false, bool UseSumBound = true, bool UseDepth = true>
class HipHybridRecursive : public semi_local::SemiLocalStrategy<Input> {
  public:
    HipHybridRecursive(int parallelThreshold, common::StaggeredStickyBraidMultiplication *staggered,
                       int depthThreshold, int sumThreshold,
                       semi_local::AntidiagonalIterativeCombing<Input, WithIf> *combing = nullptr) {
        (void)parallelThreshold;
        (void)staggered;
        (void)depthThreshold;
        (void)sumThreshold;
        (void)combing;
    }

    void compute(const Input *a, int aSize, const Input *b, int bSize, common::Permutation &kernel) noexcept override {
        inner.compute(a, aSize, b, bSize, kernel);
    }

  private:
    HipIterativeCombing<Input, WithIf> inner;
};

// Replaces semi_local::OpenMPHybridIterative<Input, WithIf> (SemiLocalLCS.h:
// 594-794): same constructor, same block heuristic (:606-646); blocks are
// combed in one batched launch and merged by batched steady-ant levels.
template <class Input, bool WithIf = false>
class HipHybridIterative : public semi_local::SemiLocalStrategy<Input> {
  public:
    HipHybridIterative(int numThreads, int colsPerBlock, int rowsPerBlock,
                       common::StaggeredStickyBraidMultiplication *staggered = nullptr,
                       semi_local::AntidiagonalIterativeCombing<Input, WithIf> *combing = nullptr)
        : threads(numThreads), maxCols(colsPerBlock), maxRows(rowsPerBlock) {
        (void)staggered;
        (void)combing;
    }

    void compute(const Input *a, int aSize, const Input *b, int bSize, common::Permutation &kernel) noexcept override {
        int colsPerBlock = maxCols, rowsInBlock = maxRows;
        const int totalInBlock = colsPerBlock + rowsInBlock;
        int mSmall, nSmall;
        if (rowsInBlock >= aSize) {
            rowsInBlock = aSize;
            mSmall = 1;
        } else {
            const int nums = (int)std::ceil((1.0 * aSize) / rowsInBlock);
            rowsInBlock = (int)std::ceil(1.0 * aSize / nums);
            mSmall = (int)std::ceil((aSize * 1.0) / rowsInBlock);
        }
        if (colsPerBlock >= bSize) {
            nSmall = threads;
        } else {
            const int colsPerThread = (int)std::ceil((1.0 * bSize) / threads);
            const int rest = totalInBlock - rowsInBlock;
            nSmall = colsPerThread < rest ? threads : threads * (int)std::ceil(1.0 * colsPerThread / rest);
        }
        mSmall = std::max(1, std::min(mSmall, std::max(aSize, 1)));
        nSmall = std::max(1, std::min(nSmall, std::max(bSize, 1)));
        try {
            Symbols<Input> sa(a, aSize), sb(b, bSize);
            std::vector<int32_t> rc((size_t)(aSize + bSize));
            check(sl_hybrid_iterative(sa.ptr, aSize, sb.ptr, bSize, sa.bytes, mSmall, nSmall, rc.data()));
            detail::store_common(kernel, rc);
        } catch (const std::exception &e) {
            detail::fail(e);
        }
        kernel.m = aSize;
        kernel.n = bSize;
    }

  private:
    template <class T>
    using Symbols = detail::Symbols<T>;
    int threads, maxCols, maxRows;
};

// Replaces the steady-ant strategies behind common::BraidMultiplicationStrategy
// (MongeMultuplicationStrategy.h:18-21; SequentialMemoryOptimizedStick-
// BraidMultiplication :548-594, OpenMPStickyBraid :597-673).
class HipStickyBraid : public common::BraidMultiplicationStrategy {
  public:
    void multiply(const common::Permutation &p, const common::Permutation &q, common::Permutation &res) override {
        const int n = p.rows;
        std::vector<int32_t> pr((size_t)n), qr((size_t)n), r((size_t)n);
        for (int i = 0; i < n; ++i) {
            pr[(size_t)i] = p.getColByRow(i);
            qr[(size_t)i] = q.getColByRow(i);
        }
        check(sl_braid_multiply(pr.data(), qr.data(), n, r.data()));
        res = common::Permutation(n, n);
        for (int i = 0; i < n; ++i) res.set(i, r[(size_t)i]);
    }
};

// LLCS on the device: an lcs::LCSStrategy<Input> (LCS.h:9-13), usable where
// the reference takes lcs::LCSBySemiLocal over a combing strategy or a
// bit_parallel::lcs::BitSemiLocalSticky (BitParalellLCS.h:868-1401); any
// alphabet.  Scored by Hyyro's bit-vector recurrence (sl_llcs_hyyro) when the
// longer string has at most 256 distinct symbols and 2^20 symbols, otherwise
// read off the semi-local kernel (m - #{s < m : P[s] >= n}); the same value.
template <class Input>
class HipLLCS : public lcs::LCSStrategy<Input> {
  public:
    int compute(const Input *a, int sizeA, const Input *b, int sizeB) override {
        detail::Symbols<Input> sa(a, sizeA), sb(b, sizeB);
        int llcs = 0;
        const int st = sl_llcs_hyyro(sa.ptr, sizeA, sb.ptr, sizeB, sa.bytes, &llcs);
        if (st == SL_OK) return llcs;
        if (st != SL_ERR_UNSUPPORTED) check(st);
        const auto rc = detail::comb(a, sizeA, b, sizeB);
        int acc = 0;
        for (int s = 0; s < sizeA; ++s) acc += rc[(size_t)s] >= sizeB;
        return sizeA - acc;
    }
};

// Hyyro's bit-vector LLCS on the device (sl_llcs_hyyro): replaces
// bit_parallel::lcs::Hyyro<Word, Input> (BitParalellLCS.h:337-433), same
// score.  Word is accepted for the signature (the device uses 64-bit words);
// Input must be 1, 2 or 4 bytes; at most 256 distinct symbols in the longer
// string, strings up to 2^20 symbols.
template <class Word, class Input>
class HipHyyro : public lcs::LCSStrategy<Input> {
  public:
    int compute(const Input *a, int sizeA, const Input *b, int sizeB) override {
        static_assert(sizeof(Input) == 1 || sizeof(Input) == 2 || sizeof(Input) == 4, "1, 2 or 4-byte symbols");
        int llcs = 0;
        check(sl_llcs_hyyro(a, sizeA, b, sizeB, (int)sizeof(Input), &llcs));
        return llcs;
    }
};

}  // namespace semilocal_hip
