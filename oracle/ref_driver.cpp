// ref_driver.cpp -- extern "C" wrappers around the reference's own flat API.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference headers where they lie (REFSRC=<reference>/source/cpu); nothing
// of the reference is copied.  The resulting oracle/_ref/libref.so is used to
// (1) write golden vectors under tests/golden and (2) time the reference's
// OpenMP path as the CPU baseline.  Functions called (reference paths):
//   semi_local::sticky_braid_mpi              semi_local.h:212-276
//   semi_local::sticky_braid_sequential       semi_local.h:135-196
//   semi_local::hybrid_iterative              semi_local.h:577-738
//   steady_ant::steady_ant_parallel_wrapper   unit_monge_mult/steady_ant.h:811-844
//   ::precalc                                 unit_monge_mult/steady_ant.h:860-894
//   binary::llcs_2symbol_smart_combing        prefix_lcs/bitwise/transposition_network_binary_alphabet.h:598-663
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <omp.h>
#include "semi_local.h"
#include "prefix_lcs/bitwise/transposition_network_binary_alphabet.h"

namespace {
PrecalcMap &precalc_map() {
    static PrecalcMap map;
    static std::once_flag once;
    std::call_once(once, [] { precalc(map, 5); });
    return map;
}
void to_rc(const AbstractPermutation &p, int32_t *out, int n) {
    for (int i = 0; i < n; ++i) out[i] = p.get_col_by_row(i);
}
}  // namespace

extern "C" {

void ref_sticky_braid_mpi(const int32_t *a, int m, const int32_t *b, int n, int32_t *perm, int threads) {
    Permutation p(m + n, m + n);
    semi_local::sticky_braid_mpi<int, false, true>(p, a, m, b, n, threads);
    to_rc(p, perm, m + n);
}

void ref_sticky_braid_sequential(const int32_t *a, int m, const int32_t *b, int n, int32_t *perm) {
    Permutation p(m + n, m + n);
    semi_local::sticky_braid_sequential<int, false>(p, a, m, b, n);
    to_rc(p, perm, m + n);
}

// OpenMP parallel-for over independent pairs, each combed single-threaded
// (the reference has no batch API; this is how its OpenMP path is used on a
// ragged batch).
void ref_sticky_braid_batch(const int32_t *a_cat, const int64_t *a_off, const int32_t *a_len,
                            const int32_t *b_cat, const int64_t *b_off, const int32_t *b_len,
                            const int64_t *p_off, int32_t *perm_cat, int npairs, int threads) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int i = 0; i < npairs; ++i) {
        int m = a_len[i], n = b_len[i];
        Permutation p(m + n, m + n);
        semi_local::sticky_braid_mpi<int, false, false>(p, a_cat + a_off[i], m, b_cat + b_off[i], n, 1);
        to_rc(p, perm_cat + p_off[i], m + n);
    }
}

void ref_hybrid_iterative(const int32_t *a, int m, const int32_t *b, int n, int small_m, int small_n,
                          int32_t *perm, int threads) {
    auto &map = precalc_map();
    Permutation p(m + n, m + n);
    semi_local::hybrid_iterative<int, false>(p, a, m, b, n, map, small_m, small_n, threads);
    to_rc(p, perm, m + n);
}

void ref_steady_ant(const int32_t *p, const int32_t *q, int n, int32_t *r, int nested_lvls) {
    auto &map = precalc_map();
    Permutation P(n, n), Q(n, n), R(n, n);
    for (int i = 0; i < n; ++i) {
        P.set_point(i, p[i]);
        Q.set_point(i, q[i]);
    }
    distance_unit_monge_product::steady_ant::steady_ant_parallel_wrapper(P, Q, R, map, nested_lvls);
    to_rc(R, r, n);
}

void ref_precalc_warmup(void) { (void)precalc_map(); }

int ref_llcs_2symbol_smart(const uint64_t *a_rev, int a_words, const uint64_t *b, int b_words, int a_total,
                           int threads, int formula_one) {
    // the reference routine takes non-const pointers; it does not write them
    return prefix_lcs_via_semi_local::binary::llcs_2symbol_smart_combing<uint64_t>(
        const_cast<uint64_t *>(a_rev), a_words, const_cast<uint64_t *>(b), b_words, a_total, threads,
        formula_one != 0);
}

// The reference's Hyyro LLCS (lllcs_hyyro::preprocess + hyyro_magic over
// 64-bit words, transposition_network_binary_alphabet.h:743-820): the
// alphabet of a, its bit tables over b, then the row recurrence.
int ref_hyyro(const int32_t *a, int m, const int32_t *b, int n) {
    if (m == 0 || n == 0) return 0;
    std::unordered_set<int> alphabet(a, a + m);
    std::unordered_map<int, uint64_t *> lookup;
    lllcs_hyyro::preprocess<uint64_t>(b, n, alphabet, lookup);
    const int words = (n + 63) / 64;
    const int score = lllcs_hyyro::hyyro_magic<uint64_t>(a, m, words, lookup);
    for (auto &kv : lookup) delete[] kv.second;
    return score;
}

}  // extern "C"
