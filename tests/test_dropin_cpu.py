"""The C++ drop-in headers compile where the reference's call sites do:
include/semilocal_hip.h as C, include/semilocal_hip.hpp on its own, and
include/semilocal_hip_strategies.hpp together with the reference's headers
at -O0 and -O2 (compiled where the reference tree lies; no GPU needed)."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "semilocal_amd")
REFSRC = "/root/reference/source/cpu"


def _compile(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


def test_c_header_compiles_as_c99():
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.c")
        with open(src, "w") as f:
            f.write('#include "semilocal_hip.h"\nint main(void) { sl_server_stats_t s; (void)s; '
                    'return sl_version() > 0 ? 0 : 1; }\n')
        _compile(["gcc", "-std=c99", "-Wall", "-Werror", "-I", INC, "-c", src, "-o", os.path.join(d, "t.o")])


def test_flat_templates_need_no_reference_header():
    # any type with the AbstractPermutation members works as the result
    code = r'''
#include "semilocal_hip.hpp"
#include <map>
struct Perm {
    int row_size, col_size; std::vector<int> rc;
    Perm(int r, int c) : row_size(r), col_size(c), rc(r, -1) {}
    void set_point(int r, int c) { rc[r] = c; }
    int get_col_by_row(int r) const { return rc[r]; }
};
int main() {
    std::map<int, int> precalc_stand_in;
    int a[3] = {1, 2, 3}, b[2] = {2, 3};
    Perm p(5, 5), q(5, 5), r(5, 5);
    semilocal_hip::sticky_braid_mpi<int, false, true>(p, a, 3, b, 2, 8);
    semilocal_hip::sticky_braid_sequential<int, false>(q, a, 3, b, 2);
    semilocal_hip::hybrid_iterative<int, false>(r, a, 3, b, 2, precalc_stand_in, 1, 1);
    semilocal_hip::steady_ant_parallel_wrapper(p, q, r, precalc_stand_in, 3);
    semilocal_hip::staggered_sticky_multiplication<true, false>(&p, &q, 2, precalc_stand_in, &r);
    unsigned long long w[1] = {0};
    return semilocal_hip::llcs_2symbol_smart_combing<unsigned long long>(w, 1, w, 1, 64) + semilocal_hip::llcs_from_kernel(p, 3, 2);
}
'''
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "t.cpp")
        with open(src, "w") as f:
            f.write(code)
        _compile(["g++", "-std=c++17", "-Wall", "-Werror", "-I", INC, src, "-L", LIBDIR, "-lsemilocal_hip",
                  "-o", os.path.join(d, "t")])


@pytest.mark.skipif(not os.path.isdir(REFSRC), reason="reference tree not present")
@pytest.mark.parametrize("opt", ["-O0", "-O2"])
def test_strategies_compile_and_link_with_the_reference_headers(opt):
    with tempfile.TemporaryDirectory() as d:
        _compile(["g++", opt, "-std=c++17", "-fopenmp", "-w", "-I", REFSRC, "-I", INC,
                  os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp"), "-L", LIBDIR, "-lsemilocal_hip",
                  "-o", os.path.join(d, "dropin_check")])
