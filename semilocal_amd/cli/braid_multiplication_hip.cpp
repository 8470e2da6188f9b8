// braid_multiplication_hip -- the braid-multiplication driver's interface
// (cpu_impl/braid_multiplication_parallel.cpp: <depth> <n> <seed>; stdout
// preprocess ms, algorithm ms, hash of the product, n, seed; utils/runners.py
// :16-31), on the GPU.  The reference's drivers do not compile (their
// fillPermutationMatrix is missing), so the inputs come from this driver's
// own generator: P and Q are std::shuffle permutations of [0, n) under
// std::mt19937_64(seed) and std::mt19937_64(~seed).  depth is accepted and
// ignored.  Preprocess = device initialisation.  An optional fourth argument
// names a file that receives P, Q and the product R, one line each,
// space-separated (for checking the product outside the driver).
#include <algorithm>
#include <numeric>
#include <random>

#include "cli_common.h"

int main(int argc, char **argv) {
    using namespace slcli;
    if (argc < 4) fail("usage: braid_multiplication_hip <depth> <n> <seed> [dump_file]");
    const long n = std::strtol(argv[2], nullptr, 10);
    const long long seed = std::strtoll(argv[3], nullptr, 10);
    if (n < 1 || n > (1L << 22)) fail("n must be in [1, 4194304]");
    std::vector<int32_t> p((size_t)n), q((size_t)n), r((size_t)n);
    std::iota(p.begin(), p.end(), 0);
    std::iota(q.begin(), q.end(), 0);
    std::mt19937_64 gp((unsigned long long)seed), gq(~(unsigned long long)seed);
    std::shuffle(p.begin(), p.end(), gp);
    std::shuffle(q.begin(), q.end(), gq);
    auto t0 = std::chrono::high_resolution_clock::now();
    check(sl_init(0), "sl_init");
    const long pre = elapsed_ms(t0);
    t0 = std::chrono::high_resolution_clock::now();
    check(sl_braid_multiply(p.data(), q.data(), (int)n, r.data()), "sl_braid_multiply");
    const long algo = elapsed_ms(t0);
    if (argc > 4) {
        std::ofstream dump(argv[4]);
        if (!dump) fail(std::string("cannot write ") + argv[4]);
        for (const auto *v : {&p, &q, &r}) {
            for (size_t i = 0; i < v->size(); ++i) dump << (i ? " " : "") << (*v)[i];
            dump << "\n";
        }
    }
    std::cout << pre << "\n" << algo << "\n" << permutation_hash(r) << "\n" << n << "\n" << seed << std::flush;
    return 0;
}
