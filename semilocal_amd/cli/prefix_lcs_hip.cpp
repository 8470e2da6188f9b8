// prefix_lcs_hip -- an LCS driver with the reference's LCS stdout (the
// score in place of the hash; utils/runners.py:59-74), any alphabet: Hyyro's
// bit-vector LLCS on the device (sl_llcs_hyyro) when the longer string has at
// most 256 distinct symbols and 2^20 symbols, otherwise the semi-local kernel
// by combing and LLCS = m - #{s < m : P[s] >= n} (include/LCS/LCS.h:223-233).
//   prefix_lcs_hip <threads> <a_file> <b_file>
#include "cli_common.h"

int main(int argc, char **argv) {
    using namespace slcli;
    if (argc < 4) fail("usage: prefix_lcs_hip <threads> <a_file> <b_file>");
    const Sequence a = read_sequence(argv[2]), b = read_sequence(argv[3]);
    const int m = (int)a.symbols.size(), n = (int)b.symbols.size();
    auto t0 = std::chrono::high_resolution_clock::now();
    check(sl_init(0), "sl_init");
    const long pre = elapsed_ms(t0);
    int llcs = 0;
    t0 = std::chrono::high_resolution_clock::now();
    if (sl_llcs_hyyro(a.symbols.data(), m, b.symbols.data(), n, 4, &llcs) != SL_OK) {
        // more than 256 distinct symbols or a string past 2^20: the kernel
        std::vector<int32_t> rc((size_t)(m + n));
        check(sl_semi_local_comb(a.symbols.data(), m, b.symbols.data(), n, 4, rc.data(), nullptr),
              "sl_semi_local_comb");
        llcs = sl_llcs_from_kernel(rc.data(), m, n);
    }
    const long algo = elapsed_ms(t0);
    print_result(pre, algo, llcs, a, b);
    return 0;
}
