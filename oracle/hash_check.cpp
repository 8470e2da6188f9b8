// hash_check -- pins the command-line drivers' permutation hash
// (semilocal_amd/cli/cli_common.h) to the reference's own `hash`
// (test_utils.h:12-19), compiled in place from the reference headers with
// the same compiler: both over the same random permutations, sizes 1 .. 2^20.
// Prints "hash ok <cases>" or the first mismatch and exits non-zero.
// `hash_check <file>`: reads row_to_col (whitespace-separated) from the file
// and prints the reference's hash of it.
#include <algorithm>
#include <cstdio>
#include <iostream>
#include <numeric>
#include <random>
#include <vector>

#include "../semilocal_amd/cli/cli_common.h"
#include "unit_monge_mult/matrices.h"   // needs <iostream> first
#include "test_utils.h"

int main(int argc, char **argv) {
    if (argc > 1) {
        std::ifstream in(argv[1]);
        std::vector<int32_t> row_to_col;
        for (int v; in >> v;) row_to_col.push_back(v);
        const int n = (int)row_to_col.size();
        Permutation perm(n, n);
        for (int r = 0; r < n; ++r) perm.set_point(r, row_to_col[(size_t)r]);
        std::printf("%lld\n", hash(perm, perm.row_size));
        return 0;
    }
    std::mt19937_64 g(7);
    int cases = 0;
    for (int n : {1, 2, 3, 17, 1000, 4096, 99991, 1 << 20}) {
        for (int rep = 0; rep < 3; ++rep) {
            std::vector<int32_t> row_to_col((size_t)n);
            std::iota(row_to_col.begin(), row_to_col.end(), 0);
            if (rep > 0) std::shuffle(row_to_col.begin(), row_to_col.end(), g);
            Permutation perm(n, n);
            for (int r = 0; r < n; ++r) perm.set_point(r, row_to_col[(size_t)r]);
            const long long want = hash(perm, perm.row_size);
            const long long got = slcli::permutation_hash(row_to_col);
            if (want != got) {
                std::printf("mismatch n=%d rep=%d reference=%lld driver=%lld\n", n, rep, want, got);
                return 1;
            }
            ++cases;
        }
    }
    std::printf("hash ok %d\n", cases);
    return 0;
}
