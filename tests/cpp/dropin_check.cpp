// dropin_check.cpp -- the drop-in claim, checked at the reference's own call
// sites.  TEST INFRASTRUCTURE ONLY: compiled by oracle/Makefile against the
// reference's headers where they lie (nothing of the reference is copied)
// and against include/semilocal_hip*.hpp; linked with libsemilocal_hip.so.
//
// Every check calls a reference entry point and the same call site switched
// to the engine, on the same seeded input, and compares the results:
//   flat API:   semi_local::sticky_braid_mpi / sticky_braid_sequential /
//               hybrid_iterative / hybrid_iterative_wrapper,
//               steady_ant::steady_ant_parallel_wrapper,
//               steady_ant::staggered_sticky_multiplication<RowGlue, false>,
//               binary::llcs_2symbol_smart_combing<uint64_t / uint32_t>
//   namespaced: semi_local::OpenMPIterativeCombing, OpenMPHybridIterative,
//               common::SequentialMemoryOptimizedStickBraidMultiplication,
//               lcs::LCSBySemiLocal over either combing strategy,
//               InteractiveDuplicateSearchSemiDense / SemiSparse::find,
//               bit_parallel::lcs::Hyyro, lllcs_hyyro::hyyro_magic(_mpi_with_precalc).
// Prints one line per check and "dropin_check: N checks, M failures"; exits
// non-zero on any failure.  Needs a GPU to run.
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <optional>
#include <random>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "semi_local.h"
#include "prefix_lcs/bitwise/encoders_and_decoders.h"
#include "prefix_lcs/bitwise/transposition_network_binary_alphabet.h"
#include "include/LCS/LCS.h"
#include "include/SemiLocal/SemiLocalLCS.h"
#include "approximate_matching/utils.h"
#include "approximate_matching/ApproximateMatching.h"
#include "include/BitwiseLCS/BitParalellLCS.h"

#include "semilocal_hip_strategies.hpp"
#include "semilocal_hip_amatch.hpp"

static int checks = 0, failures = 0;

static void report(const char *what, bool ok) {
    ++checks;
    if (!ok) ++failures;
    std::printf("%-64s %s\n", what, ok ? "ok" : "MISMATCH");
}

static bool same(const AbstractPermutation &x, const AbstractPermutation &y) {
    if (x.row_size != y.row_size) return false;
    for (int i = 0; i < x.row_size; ++i)
        if (x.get_col_by_row(i) != y.get_col_by_row(i)) return false;
    return true;
}

static bool same(const common::Permutation &x, const common::Permutation &y) {
    if (x.rows != y.rows) return false;
    for (int i = 0; i < x.rows; ++i)
        if (x.getColByRow(i) != y.getColByRow(i)) return false;
    return true;
}

static std::vector<int> symbols(std::mt19937_64 &g, int n, int sigma) {
    std::vector<int> v((size_t)n);
    for (auto &x : v) x = (int)(g() % (uint64_t)sigma);
    return v;
}

static std::vector<int> shuffled(std::mt19937_64 &g, int n) {
    std::vector<int> v((size_t)n);
    for (int i = 0; i < n; ++i) v[(size_t)i] = i;
    std::shuffle(v.begin(), v.end(), g);
    return v;
}

int main() {
    std::mt19937_64 g(20261203);
    PrecalcMap map;
    precalc(map, 5);
    common::PrecalcMap cmap;   // the namespaced API's own base-case table
    {
        common::NaiveBraidMultiplication naive;
        common::StickyBraidMultiplication::buildPrecalcMap(&naive, cmap, 5);
    }
    char name[160];

    // ---- flat API: combing ------------------------------------------------
    for (auto mn : {std::pair<int, int>{300, 500}, {777, 123}, {1, 900}, {2000, 2000}}) {
        const int m = mn.first, n = mn.second;
        auto a = symbols(g, m, 4), b = symbols(g, n, 4);
        Permutation ref(m + n, m + n), hip(m + n, m + n);
        semi_local::sticky_braid_mpi<int, false, true>(ref, a.data(), m, b.data(), n, 4);
        semilocal_hip::sticky_braid_mpi<int, false, true>(hip, a.data(), m, b.data(), n, 4);
        std::snprintf(name, sizeof name, "sticky_braid_mpi<int,false,true> %dx%d", m, n);
        report(name, same(ref, hip));
        Permutation ref_r(m + n, m + n), hip_r(m + n, m + n);
        semi_local::sticky_braid_mpi<int, false, false>(ref_r, a.data(), m, b.data(), n, 1, true);
        semilocal_hip::sticky_braid_mpi<int, false, false>(hip_r, a.data(), m, b.data(), n, 1, true);
        std::snprintf(name, sizeof name, "sticky_braid_mpi is_reverse=true %dx%d", m, n);
        report(name, same(ref_r, hip_r));
        Permutation ref_s(m + n, m + n), hip_s(m + n, m + n);
        semi_local::sticky_braid_sequential<int, false>(ref_s, a.data(), m, b.data(), n);
        semilocal_hip::sticky_braid_sequential<int, false>(hip_s, a.data(), m, b.data(), n);
        std::snprintf(name, sizeof name, "sticky_braid_sequential<int,false> %dx%d", m, n);
        report(name, same(ref_s, hip_s));
    }
    {   // unsigned short symbols, as some reference call sites use
        const int m = 900, n = 1100;
        std::vector<unsigned short> a((size_t)m), b((size_t)n);
        for (auto &x : a) x = (unsigned short)(g() % 300);
        for (auto &x : b) x = (unsigned short)(g() % 300);
        Permutation ref(m + n, m + n), hip(m + n, m + n);
        semi_local::sticky_braid_mpi<unsigned short, false, true>(ref, a.data(), m, b.data(), n, 2);
        semilocal_hip::sticky_braid_mpi<unsigned short, false, true>(hip, a.data(), m, b.data(), n, 2);
        report("sticky_braid_mpi<unsigned short,...> 900x1100 sigma 300", same(ref, hip));
    }

    {   // first_and_third_phase_combined and the recursive hybrid
        const int m = 1200, n = 1700;
        auto a = symbols(g, m, 4), b = symbols(g, n, 4);
        Permutation ref(m + n, m + n), hip(m + n, m + n);
        semi_local::first_and_third_phase_combined<int, false, true>(ref, a.data(), m, b.data(), n, map, 0, 4);
        semilocal_hip::first_and_third_phase_combined<int, false, true>(hip, a.data(), m, b.data(), n, map, 0, 4);
        report("first_and_third_phase_combined<int,false,true> 1200x1700", same(ref, hip));
        Permutation ref_h(m + n, m + n), hip_h(m + n, m + n);
        semi_local::hybrid<int, false, true, true>(ref_h, a.data(), m, b.data(), n, map, 1, 0, 3, 400, 1);
        semilocal_hip::hybrid<int, false, true, true>(hip_h, a.data(), m, b.data(), n, map, 1, 0, 3, 400, 1);
        report("hybrid<int,false,true,true> (recursive) 1200x1700", same(ref_h, hip_h));
    }

    // ---- flat API: hybrid --------------------------------------------------
    for (auto shape : {std::pair<int, int>{2, 2}, {3, 5}, {8, 8}, {1, 16}}) {
        const int m = 3000, n = 2500;
        auto a = symbols(g, m, 4), b = symbols(g, n, 4);
        Permutation ref(m + n, m + n), hip(m + n, m + n);
        semi_local::hybrid_iterative<int, false>(ref, a.data(), m, b.data(), n, map, shape.first, shape.second, 4);
        semilocal_hip::hybrid_iterative<int, false>(hip, a.data(), m, b.data(), n, map, shape.first, shape.second, 4);
        std::snprintf(name, sizeof name, "hybrid_iterative<int,false> 3000x2500 %dx%d blocks", shape.first,
                      shape.second);
        report(name, same(ref, hip));
    }

    // ---- flat API: hybrid_iterative_wrapper (the block heuristic) ---------
    // shapes reach every branch: one row block and threads_num column blocks;
    // two row blocks with columns per thread below, and above, 64000 - rows
    for (auto sh : {std::array<int, 3>{3000, 2500, 4}, {33000, 40000, 2}, {33000, 70000, 1}}) {
        const int m = sh[0], n = sh[1], thr = sh[2];
        auto a32 = symbols(g, m, 4), b32 = symbols(g, n, 4);
        std::vector<unsigned short> a(a32.begin(), a32.end()), b(b32.begin(), b32.end());
        Permutation ref(m + n, m + n), hip(m + n, m + n);
        semi_local::hybrid_iterative_wrapper<false>(ref, a.data(), m, b.data(), n, map, thr);
        semilocal_hip::hybrid_iterative_wrapper<false>(hip, a.data(), m, b.data(), n, map, thr);
        std::snprintf(name, sizeof name, "hybrid_iterative_wrapper<false> %dx%d, %d threads", m, n, thr);
        report(name, same(ref, hip));
    }

    // ---- flat API: steady ant and staggered --------------------------------
    for (int n : {1, 7, 64, 1000, 5000, 100000}) {
        auto pv = shuffled(g, n), qv = shuffled(g, n);
        Permutation p(n, n), q(n, n), ref(n, n), hip(n, n);
        for (int i = 0; i < n; ++i) {
            p.set_point(i, pv[(size_t)i]);
            q.set_point(i, qv[(size_t)i]);
        }
        Permutation p2(n, n), q2(n, n);   // the reference's routine may reuse its inputs
        for (int i = 0; i < n; ++i) {
            p2.set_point(i, pv[(size_t)i]);
            q2.set_point(i, qv[(size_t)i]);
        }
        distance_unit_monge_product::steady_ant::steady_ant_parallel_wrapper(p, q, ref, map, 2);
        semilocal_hip::steady_ant_parallel_wrapper(p2, q2, hip, map, 2);
        std::snprintf(name, sizeof name, "steady_ant_parallel_wrapper n=%d", n);
        report(name, same(ref, hip));
    }
    for (int rg = 0; rg < 2; ++rg) {
        // block braids of two neighbouring blocks of one grid, as the hybrid glues them
        const int m1 = 400, n1 = 700, m2 = 400, n2 = 350;
        auto a = symbols(g, rg ? m1 + m2 : m1, 4), b = symbols(g, rg ? n1 : n1 + n2, 4);
        const int pm = m1, pn = n1, qm = rg ? m2 : m1, qn = rg ? n1 : n2;
        Permutation P(pm + pn, pm + pn), Q(qm + qn, qm + qn);
        semi_local::sticky_braid_mpi<int, false, false>(P, a.data(), pm, b.data(), pn, 1);
        semi_local::sticky_braid_mpi<int, false, false>(Q, rg ? a.data() + m1 : a.data(), qm,
                                                        rg ? b.data() : b.data() + n1, qn, 1);
        P.m = pm; P.n = pn; Q.m = qm; Q.n = qn;
        const int k = rg ? pn : pm;
        const int N = P.row_size + Q.row_size - k;
        Permutation ref(N, N), hip(N, N);
        if (rg) {
            distance_unit_monge_product::steady_ant::staggered_sticky_multiplication<true, false>(&P, &Q, k, map, &ref);
            semilocal_hip::staggered_sticky_multiplication<true, false>(&P, &Q, k, map, &hip);
        } else {
            distance_unit_monge_product::steady_ant::staggered_sticky_multiplication<false, false>(&P, &Q, k, map, &ref);
            semilocal_hip::staggered_sticky_multiplication<false, false>(&P, &Q, k, map, &hip);
        }
        std::snprintf(name, sizeof name, "staggered_sticky_multiplication<%s,false>", rg ? "true" : "false");
        report(name, same(ref, hip));
    }

    // ---- flat API: bit-parallel --------------------------------------------
    {
        const int bits = 1 << 14;
        std::vector<int> a((size_t)bits), b((size_t)bits);
        for (auto &x : a) x = (int)(g() & 1u);
        for (auto &x : b) x = (int)(g() & 1u);
        std::unordered_map<int, uint64_t> fwd{{0, 0}, {1, 1}};
        std::unordered_map<uint64_t, int> rev{{0, 0}, {1, 1}};
        auto ar = encode_reverse<int, uint64_t>(a, &fwd, &rev, 2);
        auto bw = encode<int, uint64_t>(b, &fwd, &rev, 2);
        const int want = prefix_lcs_via_semi_local::binary::llcs_2symbol_smart_combing<uint64_t>(
            ar.first.first, ar.first.second, bw.first.first, bw.first.second, bits, 1);
        const int got = semilocal_hip::llcs_2symbol_smart_combing<uint64_t>(ar.first.first, ar.first.second,
                                                                             bw.first.first, bw.first.second, bits, 1);
        report("llcs_2symbol_smart_combing<uint64_t> 2^14 bits", want == got);
        std::unordered_map<int, uint32_t> fwd32{{0, 0}, {1, 1}};
        std::unordered_map<uint32_t, int> rev32{{0, 0}, {1, 1}};
        auto ar32 = encode_reverse<int, uint32_t>(a, &fwd32, &rev32, 2);
        auto bw32 = encode<int, uint32_t>(b, &fwd32, &rev32, 2);
        const int got32 = semilocal_hip::llcs_2symbol_smart_combing<uint32_t>(
            ar32.first.first, ar32.first.second, bw32.first.first, bw32.first.second, bits, 1);
        report("llcs_2symbol_smart_combing<uint32_t> 2^14 bits (repacked)", want == got32);
        std::free(ar.first.first);
        std::free(bw.first.first);
        std::free(ar32.first.first);
        std::free(bw32.first.first);
    }

    // ---- namespaced strategies ---------------------------------------------
    {
        const int m = 1500, n = 2600;
        auto a = symbols(g, m, 4), b = symbols(g, n, 4);
        common::Permutation ref, hip;
        semi_local::OpenMPIterativeCombing<int, false> cpu(4);
        semilocal_hip::HipIterativeCombing<int, false> dev(4);
        cpu.compute(a.data(), m, b.data(), n, ref);
        dev.compute(a.data(), m, b.data(), n, hip);
        report("OpenMPIterativeCombing vs HipIterativeCombing 1500x2600", same(ref, hip));

        common::SequentialMemoryOptimizedStickBraidMultiplication seqmul(cmap);
        common::StaggeredStickyBraidMultiplication stag(&seqmul);
        semi_local::OpenMPHybridIterative<int, false> cpu_h(4, 700, 500, &stag, &cpu);
        semilocal_hip::HipHybridIterative<int, false> dev_h(4, 700, 500, nullptr, &dev);
        common::Permutation ref_h, hip_h;
        cpu_h.compute(a.data(), m, b.data(), n, ref_h);
        dev_h.compute(a.data(), m, b.data(), n, hip_h);
        report("OpenMPHybridIterative vs HipHybridIterative 1500x2600", same(ref_h, hip_h) && same(ref, hip_h));

        semi_local::OpenMPIterativeCombingCombined<int, false> cpu_c(4, &stag);
        semilocal_hip::HipIterativeCombingCombined<int, false> dev_c(4, &stag);
        common::Permutation ref_c, hip_c;
        cpu_c.compute(a.data(), m, b.data(), n, ref_c);
        dev_c.compute(a.data(), m, b.data(), n, hip_c);
        report("OpenMPIterativeCombingCombined vs HipIterativeCombingCombined", same(ref_c, hip_c));
        semi_local::OpenMPHybridRecursive<int, false, true, true> cpu_r(1, &stag, 3, 800, &cpu);
        semilocal_hip::HipHybridRecursive<int, false, true, true> dev_r(1, &stag, 3, 800, &dev);
        common::Permutation ref_r, hip_r;
        cpu_r.compute(a.data(), m, b.data(), n, ref_r);
        dev_r.compute(a.data(), m, b.data(), n, hip_r);
        report("OpenMPHybridRecursive vs HipHybridRecursive", same(ref_r, hip_r));

        lcs::LCSBySemiLocal<int> lcs_cpu(&cpu), lcs_dev(&dev);
        semilocal_hip::HipLLCS<int> llcs_dev;
        const int want = lcs_cpu.compute(a.data(), m, b.data(), n);
        report("LCSBySemiLocal over the device combing strategy", lcs_dev.compute(a.data(), m, b.data(), n) == want);
        report("HipLLCS", llcs_dev.compute(a.data(), m, b.data(), n) == want);
    }
    for (int n : {5, 300, 4096, 30000}) {
        auto pv = shuffled(g, n), qv = shuffled(g, n);
        common::Permutation p(n, n), q(n, n), ref, hip;
        for (int i = 0; i < n; ++i) {
            p.set(i, pv[(size_t)i]);
            q.set(i, qv[(size_t)i]);
        }
        common::SequentialMemoryOptimizedStickBraidMultiplication cpu(cmap);
        semilocal_hip::HipStickyBraid dev;
        cpu.multiply(p, q, ref);
        dev.multiply(p, q, hip);
        std::snprintf(name, sizeof name, "SequentialMemoryOptimizedStickBraidMultiplication vs HipStickyBraid n=%d", n);
        report(name, same(ref, hip));
    }

    // duplicate searches: pattern vs a text holding mutated copies of it
    for (auto [m, n, k, sigma] : std::vector<std::tuple<int, int, double, int>>{
             {12, 400, 0.8, 4}, {30, 2000, 0.9, 4}, {25, 1500, 0.6, 2}, {60, 5000, 0.95, 8}}) {
        auto p = symbols(g, m, sigma), t = symbols(g, n, sigma);
        for (int c = 0; c < 6; ++c) {
            const int pos = (int)(g() % (uint64_t)(n - m));
            for (int i = 0; i < m; ++i) t[(size_t)(pos + i)] = (g() % 8 == 0) ? (int)(g() % (uint64_t)sigma) : p[(size_t)i];
        }
        using Iv = approximate_matching::utils::Interval;
        std::vector<Iv> ref_d, hip_d, ref_s, hip_s;
        InteractiveDuplicateSearchSemiDense().find(p.data(), m, t.data(), n, k, ref_d);
        semilocal_hip::InteractiveDuplicateSearchSemiDense().find(p.data(), m, t.data(), n, k, hip_d);
        InteractiveDuplicateSearchSemiSparse().find(p.data(), m, t.data(), n, k, ref_s);
        semilocal_hip::InteractiveDuplicateSearchSemiSparse().find(p.data(), m, t.data(), n, k, hip_s);
        std::snprintf(name, sizeof name, "InteractiveDuplicateSearchSemiDense m=%d n=%d k=%.2f (%zu found)", m, n, k,
                      ref_d.size());
        report(name, ref_d == hip_d);
        std::snprintf(name, sizeof name, "InteractiveDuplicateSearchSemiSparse m=%d n=%d k=%.2f (%zu found)", m, n, k,
                      ref_s.size());
        report(name, ref_s == hip_s);
    }

    // Hyyro's bit-vector LLCS: namespaced strategy and the flat entry points
    for (auto [m, n, sigma] : std::vector<std::tuple<int, int, int>>{
             {1, 1, 2}, {63, 65, 4}, {700, 3000, 4}, {5000, 70000, 20}, {2000, 130000, 4}}) {
        auto a = symbols(g, m, sigma), b = symbols(g, n, sigma);
        bit_parallel::lcs::Hyyro<uint64_t, int> cpu;
        semilocal_hip::HipHyyro<uint64_t, int> dev;
        const int want = cpu.compute(a.data(), m, b.data(), n);
        std::snprintf(name, sizeof name, "bit_parallel::lcs::Hyyro vs HipHyyro m=%d n=%d sigma=%d", m, n, sigma);
        report(name, dev.compute(a.data(), m, b.data(), n) == want);
        std::unordered_set<int> alphabet(a.begin(), a.end());
        std::unordered_map<int, uint32_t *> lookup;
        lllcs_hyyro::preprocess<uint32_t>(b.data(), n, alphabet, lookup);
        const int words = (n + 31) / 32;
        const int ref_flat = lllcs_hyyro::hyyro_magic<uint32_t>(a.data(), m, words, lookup);
        const int dev_flat = semilocal_hip::lllcs_hyyro::hyyro_magic<uint32_t>(a.data(), m, words, lookup);
        const int dev_mpi =
            semilocal_hip::lllcs_hyyro::hyyro_magic_mpi_with_precalc<uint32_t>(a.data(), m, words, lookup, 4);
        std::snprintf(name, sizeof name, "lllcs_hyyro::hyyro_magic<uint32_t> m=%d n=%d", m, n);
        report(name, ref_flat == want && dev_flat == want && dev_mpi == want);
        for (auto &kv : lookup) delete[] kv.second;
    }

    std::printf("dropin_check: %d checks, %d failures\n", checks, failures);
    return failures ? 1 : 0;
}
