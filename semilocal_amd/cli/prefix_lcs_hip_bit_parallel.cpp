// prefix_lcs_hip_bit_parallel -- the bit-parallel LCS driver of the
// reference (cpu_impl/prefix_lcs_semi_bit_parallel_2form.cpp: binary strings,
// same arguments, same stdout), on the GPU.  The strings are packed as the
// reference packs them (a reversed into MSB-first words, b LSB-first;
// encoders_and_decoders.h:50-139), here in 64-bit words; lengths that are not
// multiples of 64 take the combing path, which gives the same score.
//   prefix_lcs_hip_bit_parallel <threads> <a_file> <b_file>
#include "cli_common.h"

static std::vector<uint64_t> pack_reverse(const std::vector<int32_t> &bits) {
    const size_t words = bits.size() / 64;
    std::vector<uint64_t> out(words, 0);
    for (size_t w = 0; w < words; ++w) {
        uint64_t v = 0;
        for (int k = 0; k < 64; ++k) v = (v << 1) | (uint64_t)(bits[64 * w + (size_t)k] & 1);
        out[words - 1 - w] = v;   // word w: bit 63 = bits[64w]; stored in reverse word order
    }
    return out;
}

static std::vector<uint64_t> pack(const std::vector<int32_t> &bits) {
    const size_t words = bits.size() / 64;
    std::vector<uint64_t> out(words, 0);
    for (size_t i = 0; i < words * 64; ++i) out[i / 64] |= (uint64_t)(bits[i] & 1) << (i % 64);
    return out;
}

int main(int argc, char **argv) {
    using namespace slcli;
    if (argc < 4) fail("usage: prefix_lcs_hip_bit_parallel <threads> <a_file> <b_file>");
    const int threads = (int)std::strtol(argv[1], nullptr, 10);
    const Sequence a = read_sequence(argv[2]), b = read_sequence(argv[3]);
    const int m = (int)a.symbols.size(), n = (int)b.symbols.size();
    for (int32_t v : a.symbols) if (v != 0 && v != 1) fail("a: only the symbols 0 and 1 are allowed");
    for (int32_t v : b.symbols) if (v != 0 && v != 1) fail("b: only the symbols 0 and 1 are allowed");
    auto t0 = std::chrono::high_resolution_clock::now();
    check(sl_init(0), "sl_init");
    const long pre = elapsed_ms(t0);
    int llcs = 0;
    if (m % 64 == 0 && n % 64 == 0 && m > 0 && n > 0) {
        const auto ar = pack_reverse(a.symbols), bw = pack(b.symbols);
        t0 = std::chrono::high_resolution_clock::now();
        check(sl_llcs_bitparallel(ar.data(), (int)ar.size(), bw.data(), (int)bw.size(), m, threads, 0, &llcs),
              "sl_llcs_bitparallel");
    } else {
        std::vector<int32_t> rc((size_t)(m + n));
        t0 = std::chrono::high_resolution_clock::now();
        check(sl_semi_local_comb(a.symbols.data(), m, b.symbols.data(), n, 4, rc.data(), nullptr),
              "sl_semi_local_comb");
        llcs = sl_llcs_from_kernel(rc.data(), m, n);
    }
    const long algo = elapsed_ms(t0);
    print_result(pre, algo, llcs, a, b);
    return 0;
}
