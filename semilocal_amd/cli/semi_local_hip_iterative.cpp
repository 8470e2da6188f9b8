// semi_local_hip_iterative -- the combing driver of the reference
// (cpu_impl/semi_local_parallel_iterative.cpp and its withif / withoutif /
// 1and3_combined / naive variants: same arguments, same stdout), on the GPU.
//   semi_local_hip_iterative <threads> <a_file> <b_file>
// threads is accepted and ignored.  Preprocess = device initialisation; the
// algorithm time is sl_semi_local_comb, host buffers in and out.
#include "cli_common.h"

int main(int argc, char **argv) {
    using namespace slcli;
    if (argc < 4) fail("usage: semi_local_hip_iterative <threads> <a_file> <b_file>");
    const Sequence a = read_sequence(argv[2]), b = read_sequence(argv[3]);
    const int m = (int)a.symbols.size(), n = (int)b.symbols.size();
    auto t0 = std::chrono::high_resolution_clock::now();
    check(sl_init(0), "sl_init");
    const long pre = elapsed_ms(t0);
    std::vector<int32_t> rc((size_t)(m + n));
    t0 = std::chrono::high_resolution_clock::now();
    check(sl_semi_local_comb(a.symbols.data(), m, b.symbols.data(), n, 4, rc.data(), nullptr), "sl_semi_local_comb");
    const long algo = elapsed_ms(t0);
    print_result(pre, algo, permutation_hash(rc), a, b);
    return 0;
}
