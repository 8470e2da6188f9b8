// cli_common.h -- shared parts of the command-line drivers: the reference
// drivers' input files and stdout contract, so the reference's own Python
// harness (utils/runners.py:36-53, 59-74, 16-31) runs these drivers unchanged.
//
//   input file (parsers.h:38-47): line 1 the name, line 2 the length, line 3
//   the symbols as comma-separated integers;
//   stdout: preprocess ms, algorithm ms, then the permutation hash (combing
//   drivers) or the score (LCS drivers), then the two sizes and the two names,
//   one per line, the last without a newline.
#pragma once
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "semilocal_hip.h"

namespace slcli {

struct Sequence {
    std::string name;
    std::vector<int32_t> symbols;
};

inline void fail(const std::string &msg) {
    std::cerr << msg << std::endl;
    std::exit(1);
}

inline void check(int status, const char *what) {
    if (status != SL_OK) fail(std::string(what) + ": " + sl_last_error());
}

// name / length / comma-separated symbols (parsers.h:38-47)
inline Sequence read_sequence(const std::string &path) {
    std::ifstream in(path);
    if (!in) fail("cannot open " + path);
    Sequence s;
    std::string length, line;
    std::getline(in, s.name);
    std::getline(in, length);
    std::getline(in, line);
    const long n = std::strtol(length.c_str(), nullptr, 10);
    if (n < 0) fail("bad length in " + path);
    s.symbols.reserve((size_t)n);
    const char *p = line.c_str();
    while (*p && (long)s.symbols.size() < n) {
        char *end = nullptr;
        const long v = std::strtol(p, &end, 10);
        if (end == p) break;
        s.symbols.push_back((int32_t)v);
        p = end;
        while (*p == ',' || *p == ' ') ++p;
    }
    if ((long)s.symbols.size() != n) fail("length and symbols disagree in " + path);
    return s;
}

// The reference's permutation hash (test_utils.h:12-19): h = (R h + P^-1[i])
// mod M over i = 0 .. size-1, R = 4294967279, M = 4294967291.  R h exceeds
// 64 bits there and the compiled reference wraps it; the wrap is explicit here
// so the printed values agree.
inline long long permutation_hash(const std::vector<int32_t> &row_to_col) {
    const size_t n = row_to_col.size();
    std::vector<int32_t> col_to_row(n);
    for (size_t r = 0; r < n; ++r) col_to_row[(size_t)row_to_col[r]] = (int32_t)r;
    long long h = 0;
    for (size_t i = 0; i < n; ++i) {
        const unsigned long long t = 4294967279ull * (unsigned long long)h + (unsigned long long)(long long)col_to_row[i];
        h = (long long)t % 4294967291LL;
    }
    return h;
}

inline long elapsed_ms(std::chrono::high_resolution_clock::time_point t0) {
    return long(std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0).count());
}

// the combing / LCS drivers' seven lines
inline void print_result(long preprocess_ms, long algo_ms, long long value, const Sequence &a, const Sequence &b) {
    std::cout << preprocess_ms << "\n" << algo_ms << "\n" << value << "\n" << a.symbols.size() << "\n"
              << b.symbols.size() << "\n" << a.name << "\n" << b.name << std::flush;
}

}  // namespace slcli
