// semilocal_hip.hpp -- C++ drop-in layer of the MI355X semi-local LCS engine.
//
// Header-only, over the C ABI in semilocal_hip.h.  It needs none of the
// reference's headers: the flat-API templates below take the result by any
// permutation type with the reference's AbstractPermutation interface
// (unit_monge_mult/matrices.h:53-100: row_size, set_point(row, col),
// get_col_by_row(row)), so a call site written against
//     semi_local::sticky_braid_mpi<int, false, true>(perm, a, m, b, n, threads);
// switches by namespace only:
//     semilocal_hip::sticky_braid_mpi<int, false, true>(perm, a, m, b, n, threads);
// Same template parameter lists, same argument meaning, same results (the
// semi-local kernel is unique).  Arguments that only steer the reference's
// CPU parallelism (threads_num, nested_lvls, WithIf, WithWait, UseParallel,
// the precalc map) are accepted and do not change the result.
//
// Errors: the reference's flat API has no error channel.  These templates
// throw semilocal_hip::Error (status code + message from sl_last_error()).
// The strategy classes in semilocal_hip_strategies.hpp are noexcept like the
// reference's and terminate instead.
#pragma once

#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <cmath>
#include <unordered_map>
#include <vector>

#include "semilocal_hip.h"

namespace semilocal_hip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error("semilocal_hip: " + msg), code(c) {}
};

inline void check(int status) {
    if (status != SL_OK) throw Error(status, sl_last_error());
}

namespace detail {

// Symbols of any integral Input as the widths the C ABI takes: 1-byte and
// 2-byte unsigned symbols pass through, everything else goes as int32 (the
// engine compares symbols for equality only, so the values just need to be
// kept distinct; wider values are checked to fit).
template <class Input>
struct Symbols {
    static_assert(std::is_integral<Input>::value, "Input must be an integral symbol type");
    const void *ptr = nullptr;
    int bytes = 4;
    std::vector<int32_t> wide;
    Symbols(const Input *x, int n) {
        if (std::is_same<Input, uint8_t>::value || std::is_same<Input, unsigned char>::value) {
            ptr = x;
            bytes = 1;
        } else if (std::is_same<Input, uint16_t>::value) {
            ptr = x;
            bytes = 2;
        } else if (sizeof(Input) == 4) {
            ptr = x;
            bytes = 4;
        } else {
            wide.resize((size_t)n);
            for (int i = 0; i < n; ++i) {
                const long long v = (long long)x[i];
                if (v < INT32_MIN || v > INT32_MAX) throw Error(SL_ERR_UNSUPPORTED, "symbol value does not fit int32");
                wide[(size_t)i] = (int32_t)v;
            }
            ptr = wide.data();
            bytes = 4;
        }
    }
};

template <class Perm>
void store(Perm &out, const std::vector<int32_t> &rc) {
    for (size_t i = 0; i < rc.size(); ++i) out.set_point((int)i, rc[i]);
}

template <class Perm>
std::vector<int32_t> load(const Perm &p) {
    std::vector<int32_t> rc((size_t)p.row_size);
    for (int i = 0; i < p.row_size; ++i) rc[(size_t)i] = p.get_col_by_row(i);
    return rc;
}

// The kernel of (b, a) from the kernel of (a, b): P_ba[x] = size-1 - P_ab[size-1-x]
// (semi_local.h:98-108, steady_ant.h:83-89).
inline std::vector<int32_t> reversed(const std::vector<int32_t> &rc) {
    const size_t s = rc.size();
    std::vector<int32_t> out(s);
    for (size_t x = 0; x < s; ++x) out[x] = (int32_t)(s - 1) - rc[s - 1 - x];
    return out;
}

template <class Input>
std::vector<int32_t> comb(const Input *a, int m, const Input *b, int n) {
    Symbols<Input> sa(a, m), sb(b, n);
    std::vector<int32_t> rc((size_t)(m + n));
    check(sl_semi_local_comb(sa.ptr, m, sb.ptr, n, sa.bytes, rc.data(), nullptr));
    return rc;
}

// Bit order of the reference's binary encoders (encoders_and_decoders.h:
// 50-139): b LSB-first, a reversed and MSB-first, for any word type W.
template <class W>
std::vector<uint8_t> unpack_reversed(const W *words, int nwords) {
    constexpr int wb = (int)sizeof(W) * 8;
    std::vector<uint8_t> bits((size_t)nwords * wb);
    for (int k = 0; k < nwords; ++k) {
        const int i = nwords - 1 - k;
        for (int s = 0; s < wb; ++s) bits[(size_t)i * wb + s] = (uint8_t)((words[k] >> (wb - 1 - s)) & 1u);
    }
    return bits;
}

template <class W>
std::vector<uint8_t> unpack(const W *words, int nwords) {
    constexpr int wb = (int)sizeof(W) * 8;
    std::vector<uint8_t> bits((size_t)nwords * wb);
    for (int i = 0; i < nwords; ++i)
        for (int s = 0; s < wb; ++s) bits[(size_t)i * wb + s] = (uint8_t)((words[i] >> s) & 1u);
    return bits;
}

inline std::vector<uint64_t> pack64_reversed(const std::vector<uint8_t> &bits) {
    const int n = (int)((bits.size() + 63) / 64);
    std::vector<uint64_t> w((size_t)n, 0);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < 64 && (size_t)i * 64 + s < bits.size(); ++s)
            w[(size_t)(n - 1 - i)] |= (uint64_t)bits[(size_t)i * 64 + s] << (63 - s);
    return w;
}

inline std::vector<uint64_t> pack64(const std::vector<uint8_t> &bits) {
    const int n = (int)((bits.size() + 63) / 64);
    std::vector<uint64_t> w((size_t)n, 0);
    for (int i = 0; i < n; ++i)
        for (int s = 0; s < 64 && (size_t)i * 64 + s < bits.size(); ++s)
            w[(size_t)i] |= (uint64_t)bits[(size_t)i * 64 + s] << s;
    return w;
}

}  // namespace detail

// ---------------------------------------------------------------------------
// Flat API (the reference's semi_local / steady_ant / binary namespaces).
// ---------------------------------------------------------------------------

// semi_local::sticky_braid_mpi<Input, WithIf, WithWait> -- semi_local.h:212-214.
// is_reverse = true yields the kernel of (b, a), as in the reference.
template <class Input, bool WithIf, bool WithWait, class Perm>
void sticky_braid_mpi(Perm &matrix, const Input *a, int a_size, const Input *b, int b_size, int threads_num = 1,
                      bool is_reverse = false) {
    (void)threads_num;
    auto rc = detail::comb(a, a_size, b, b_size);
    detail::store(matrix, is_reverse ? detail::reversed(rc) : rc);
}

// semi_local::sticky_braid_sequential<Input, WithIf> -- semi_local.h:135-137.
template <class Input, bool WithIf, class Perm>
void sticky_braid_sequential(Perm &permutation, const Input *a, int a_size, const Input *b, int b_size) {
    detail::store(permutation, detail::comb(a, a_size, b, b_size));
}

// semi_local::first_and_third_phase_combined<Input, WithIf, WithWait> --
// semi_local.h:292-295.  The phase split is a CPU barrier optimisation; the
// kernel it produces is the same unique permutation.
template <class Input, bool WithIf, bool WithWait, class Perm, class Map>
void first_and_third_phase_combined(Perm &matrix, const Input *a, int a_size, const Input *b, int b_size, Map &map,
                                    int nested_parall_regions = 0, int threads_num = 1) {
    (void)map;
    (void)nested_parall_regions;
    (void)threads_num;
    detail::store(matrix, detail::comb(a, a_size, b, b_size));
}

// semi_local::hybrid<Input, WithIf, WithWait, UseSumBound> (recursive) --
// semi_local.h:452-456.  Its depth / sum bounds only choose where the CPU
// recursion stops; the kernel is unique, so the device combs the pair whole.
template <class Input, bool WithIf, bool WithWait, bool UseSumBound, class Perm, class Map>
void hybrid(Perm &perm, const Input *a, int m, const Input *b, int n, Map &map, int thds_per_combing_algo,
            int braid_mul_parall_depth, int depth, int sum_bound, int parallel_depth) {
    (void)map;
    (void)thds_per_combing_algo;
    (void)braid_mul_parall_depth;
    (void)depth;
    (void)sum_bound;
    (void)parallel_depth;
    detail::store(perm, detail::comb(a, m, b, n));
}

// semi_local::hybrid_iterative<Input, WithIf> -- semi_local.h:577-580.
template <class Input, bool WithIf, class Perm, class Map>
void hybrid_iterative(Perm &perm, const Input *a, int m, const Input *b, int n, Map &map, int small_m, int small_n,
                      int threads_num = 1) {
    (void)map;
    (void)threads_num;
    detail::Symbols<Input> sa(a, m), sb(b, n);
    std::vector<int32_t> rc((size_t)(m + n));
    check(sl_hybrid_iterative(sa.ptr, m, sb.ptr, n, sa.bytes, small_m, small_n, rc.data()));
    detail::store(perm, rc);
}

// Block counts of semi_local::hybrid_iterative_wrapper (semi_local.h:
// 742-788), so the same call combs the same blocks: rows in equal blocks of
// at most 32000, columns split over threads_num and further while a block
// would exceed 64000 - rows.
inline void hybrid_wrapper_blocks(int m, int n, int threads_num, int *small_m, int *small_n) {
    const int threads = threads_num > 0 ? threads_num : 1;
    int sm = 1, rows = m;
    if (m > 32000) {
        const int parts = (int)std::ceil((double)m / 32000.0);
        rows = (int)std::ceil((double)m / parts);
        sm = (int)std::ceil((double)m / rows);
    }
    int sn = threads;
    if (n > 32000) {
        const int per_thread = (int)std::ceil((double)n / threads);
        const int room = 64000 - rows;
        if (per_thread >= room) sn = threads * (int)std::ceil((double)per_thread / room);
    }
    *small_m = sm;
    *small_n = sn;
}

// semi_local::hybrid_iterative_wrapper<WithIf> -- semi_local.h:742-788.
template <bool WithIf, class Perm, class Map>
void hybrid_iterative_wrapper(Perm &perm, const unsigned short *a, int m, const unsigned short *b, int n, Map &map,
                              int threads_num = 1) {
    int small_m = 1, small_n = 1;
    hybrid_wrapper_blocks(m, n, threads_num, &small_m, &small_n);
    hybrid_iterative<unsigned short, WithIf>(perm, a, m, b, n, map, small_m, small_n, threads_num);
}

// distance_unit_monge_product::steady_ant::steady_ant_parallel_wrapper --
// unit_monge_mult/steady_ant.h:811-812.  p and q must be full permutations.
template <class Perm, class Map>
void steady_ant_parallel_wrapper(Perm &p, Perm &q, Perm &product, Map &map, int nested_lvls = 3) {
    (void)map;
    (void)nested_lvls;
    auto pr = detail::load(p), qr = detail::load(q);
    std::vector<int32_t> r(pr.size());
    check(sl_braid_multiply(pr.data(), qr.data(), (int)pr.size(), r.data()));
    detail::store(product, r);
}

// distance_unit_monge_product::steady_ant::staggered_sticky_multiplication
// <RowGlue, UseParallel> -- unit_monge_mult/steady_ant.h:609-612.
template <bool RowGlue, bool UseParallel, class Perm, class Map>
void staggered_sticky_multiplication(Perm *p, Perm *q, int k, Map &map, Perm *product,
                                     int nested_parall_regions = 0) {
    (void)map;
    (void)nested_parall_regions;
    if (k == p->row_size && k == q->col_size) return;   // the reference leaves the product unset (:616-619)
    auto pr = detail::load(*p), qr = detail::load(*q);
    std::vector<int32_t> out(pr.size() + qr.size() - (size_t)k);
    check(sl_staggered_multiply(pr.data(), (int)pr.size(), qr.data(), (int)qr.size(), k, RowGlue ? 1 : 0,
                                out.data()));
    detail::store(*product, out);
}

// prefix_lcs_via_semi_local::binary::llcs_2symbol_smart_combing<W> --
// prefix_lcs/bitwise/transposition_network_binary_alphabet.h:598-599.
// Same packing as the reference's encoders for any unsigned word W.
template <class W>
int llcs_2symbol_smart_combing(W *a_reverse, int a_size, W *b, int b_size, int a_total_symbols, int threads_num = 1,
                               bool formula_one = false) {
    static_assert(std::is_unsigned<W>::value, "W must be an unsigned machine word");
    int llcs = 0;
    if (std::is_same<W, uint64_t>::value) {
        check(sl_llcs_bitparallel(reinterpret_cast<const uint64_t *>(a_reverse), a_size,
                                  reinterpret_cast<const uint64_t *>(b), b_size, a_total_symbols, threads_num,
                                  formula_one ? 1 : 0, &llcs));
        return llcs;
    }
    const auto ar = detail::pack64_reversed(detail::unpack_reversed(a_reverse, a_size));
    const auto bw = detail::pack64(detail::unpack(b, b_size));
    check(sl_llcs_bitparallel(ar.data(), (int)ar.size(), bw.data(), (int)bw.size(), a_total_symbols, threads_num,
                              formula_one ? 1 : 0, &llcs));
    return llcs;
}

// LLCS read from a kernel (include/LCS/LCS.h:223-233).
template <class Perm>
int llcs_from_kernel(const Perm &kernel, int a_size, int b_size) {
    int acc = 0;
    for (int s = 0; s < a_size; ++s) acc += kernel.get_col_by_row(s) >= b_size;
    return a_size - acc;
}

// Flat Hyyro entry points (prefix_lcs/bitwise/transposition_network_binary_
// alphabet.h:743-916): the caller has built `lookup` with the reference's
// lllcs_hyyro::preprocess (one bit table per symbol of a over the packed
// string b).  The packed string is rebuilt from the tables (positions no
// table marks take a symbol that matches nothing) and scored on the device
// by sl_llcs_hyyro; same score.
namespace lllcs_hyyro {

template <class Input>
int hyyro_magic(const int *str_a, int size_a_bits, int size_b_word, std::unordered_map<int, Input *> &lookup) {
    const int W = (int)sizeof(Input) * 8;
    const int64_t nb = (int64_t)size_b_word * W;
    int sentinel = 0;
    bool first = true;
    for (auto &kv : lookup)
        if (first || kv.first < sentinel) {
            sentinel = kv.first;
            first = false;
        }
    sentinel = first ? 0 : sentinel - 1;   // below every key: matches no symbol of a
    std::vector<int32_t> b((size_t)nb, sentinel);
    for (auto &kv : lookup) {
        const Input *t = kv.second;
        if (!t) continue;
        for (int w = 0; w < size_b_word; ++w)
            for (int j = 0; j < W; ++j)
                if ((t[w] >> j) & Input(1)) b[(size_t)w * W + j] = kv.first;
    }
    int llcs = 0;
    check(sl_llcs_hyyro(str_a, size_a_bits, b.data(), (int)nb, 4, &llcs));
    return llcs;
}

template <class Input>
int hyyro_magic_mpi_with_precalc(int *str_a, int size_a_bits, int size_b_word,
                                 std::unordered_map<int, Input *> &lookup, int num_threads = 1) {
    (void)num_threads;
    return hyyro_magic<Input>(str_a, size_a_bits, size_b_word, lookup);
}

}  // namespace lllcs_hyyro

}  // namespace semilocal_hip
