// semi_local_hip_hybrid_iterative -- the hybrid driver of the reference
// (cpu_impl/semi_local_hybrid_iterative.cpp: same arguments, same stdout), on
// the GPU: the blocks of semi_local::hybrid_iterative_wrapper's heuristic
// (semi_local.h:742-788) combed in one launch and merged by batched
// steady-ant levels.
//   semi_local_hip_hybrid_iterative <threads> <a_file> <b_file>
// Preprocess = device initialisation (the reference's is its precalc table,
// which the GPU path does not need).
#include "cli_common.h"
#include "semilocal_hip.hpp"

int main(int argc, char **argv) {
    using namespace slcli;
    if (argc < 4) fail("usage: semi_local_hip_hybrid_iterative <threads> <a_file> <b_file>");
    const int threads = (int)std::strtol(argv[1], nullptr, 10);
    const Sequence a = read_sequence(argv[2]), b = read_sequence(argv[3]);
    const int m = (int)a.symbols.size(), n = (int)b.symbols.size();
    int small_m = 1, small_n = 1;
    semilocal_hip::hybrid_wrapper_blocks(m, n, threads, &small_m, &small_n);
    auto t0 = std::chrono::high_resolution_clock::now();
    check(sl_init(0), "sl_init");
    const long pre = elapsed_ms(t0);
    std::vector<int32_t> rc((size_t)(m + n));
    t0 = std::chrono::high_resolution_clock::now();
    check(sl_hybrid_iterative(a.symbols.data(), m, b.symbols.data(), n, 4, small_m, small_n, rc.data()),
          "sl_hybrid_iterative");
    const long algo = elapsed_ms(t0);
    print_result(pre, algo, permutation_hash(rc), a, b);
    return 0;
}
