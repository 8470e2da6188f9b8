// prefix_lcs_hip -- an LCS driver with the reference's LCS stdout (the
// score in place of the hash; utils/runners.py:59-74), any alphabet: the
// semi-local kernel by combing, then LLCS = m - #{s < m : P[s] >= n}
// (include/LCS/LCS.h:223-233).
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
    std::vector<int32_t> rc((size_t)(m + n));
    int llcs = 0;
    t0 = std::chrono::high_resolution_clock::now();
    check(sl_semi_local_comb(a.symbols.data(), m, b.symbols.data(), n, 4, rc.data(), nullptr), "sl_semi_local_comb");
    llcs = sl_llcs_from_kernel(rc.data(), m, n);
    const long algo = elapsed_ms(t0);
    print_result(pre, algo, llcs, a, b);
    return 0;
}
