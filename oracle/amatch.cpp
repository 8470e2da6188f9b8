// amatch.cpp -- TEST INFRASTRUCTURE ONLY (the checker, never the product):
// a CPU restatement of the reference's approximate-matching duplicate
// searches (approximate_matching/ApproximateMatching.h), built into
// oracle/libamatch.so.
//
// The window values come from direct LCS dynamic programming, not from a
// braid: for every start s one DP row set gives LCS(p, text[s, e)) for all
// e <= s + w, so score(s, e) = 2 LCS - m - (e - s) (the fixed scheme 0/-2/-1
// of InteractiveDuplicateSearchSemi*, utils.h:194-235) and each window's
// maximum is taken by brute force over its triangle.  The sequential phases
// restate ApproximateMatching.h:212-228 (phase3), :252-336 (SemiSparse) and
// :339-432 (SemiDense).  Pinned against the reference itself compiled in place
// (oracle/ref_amatch.cpp) by tests/test_amatch_cpu.py.
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <unordered_set>
#include <vector>

namespace {

struct Iv {
    int start;
    int end;
    double score;
    int len() const { return end - start; }
};

// (score, length, end) lexicographic: the window maximum's order
bool better(int sc, int len, int e, int bsc, int blen, int be) {
    if (sc != bsc) return sc > bsc;
    if (len != blen) return len > blen;
    return e > be;
}

void windows(const int32_t *p, int m, const int32_t *t, int n, int w, int L, int32_t *ws, int32_t *bs, int32_t *be,
             int32_t *bsc) {
    const int nw = n - w + 1;
    // lcs[s][d] = LCS(p, t[s, s+d)), d = 0..w
    std::vector<std::vector<int>> lcs((size_t)n + 1);
    std::vector<int> prev((size_t)w + 1), cur((size_t)w + 1);
    for (int s = 0; s <= n; ++s) {
        const int span = std::min(w, n - s);
        std::fill(prev.begin(), prev.end(), 0);
        for (int i = 1; i <= m; ++i) {
            cur[0] = 0;
            for (int d = 1; d <= span; ++d)
                cur[(size_t)d] = p[i - 1] == t[s + d - 1] ? prev[(size_t)d - 1] + 1
                                                          : std::max(prev[(size_t)d], cur[(size_t)d - 1]);
            std::swap(prev, cur);
        }
        lcs[(size_t)s].assign(prev.begin(), prev.begin() + span + 1);
    }
    auto score = [&](int s, int e) { return 2 * lcs[(size_t)s][(size_t)(e - s)] - m - (e - s); };
    for (int i = 0; i < nw; ++i) {
        ws[i] = score(i, i + w);
        int bsc_ = 0, blen = -1, be_ = -1;
        for (int e = i + L; e <= i + w; ++e)
            for (int s = i; s <= e - L; ++s) {
                const int sc = score(s, e);
                if (blen < 0 || better(sc, e - s, e, bsc_, blen, be_)) {
                    bsc_ = sc;
                    blen = e - s;
                    be_ = e;
                }
            }
        bs[i] = be_ - blen;
        be[i] = be_;
        bsc[i] = bsc_;
    }
}

long long key(int l, int r) {
    long long k = l;
    return (k << 32) | r;
}

void phase3(std::vector<Iv> &w3, std::vector<Iv> &out) {
    std::sort(w3.begin(), w3.end(), [](const Iv &a1, const Iv &a2) {
        return (a1.start < a2.start) || (a1.start == a2.start && a1.score > a2.score);
    });
    int lb = 0, rb = 0;
    for (auto &c : w3)
        if (c.start > lb && c.start >= rb) {
            out.push_back(c);
            lb = c.start;
            rb = std::max(rb, c.end);
        }
}

}  // namespace

extern "C" {

// Per-window arrays (n - w + 1 entries each); 0 on success.
int orc_am_windows(const int32_t *p, int m, const int32_t *t, int n, int w, int L, int32_t *ws, int32_t *bs,
                   int32_t *be, int32_t *bsc) {
    if (m < 1 || w < 1 || w > n || L < 0 || L > w) return 1;
    windows(p, m, t, n, w, L, ws, bs, be, bsc);
    return 0;
}

// variant 0: InteractiveDuplicateSearchSemiDense, 1: ...SemiSparse.  Writes
// up to cap intervals and returns their total number (-1 on bad arguments).
int orc_dup_search(int variant, const int32_t *p, int m, const int32_t *t, int n, double k, int32_t *start,
                   int32_t *end, double *score, int cap) {
    const int w = int(m / k);
    const int left_w = int(m * k) - 1;
    const double kdi = -(m * (1 / k + 1) * (1 - k * k));
    const int L = left_w + 1;
    if (m < 1 || w < 1 || w > n || L < 0 || L > w) return -1;
    const int nw = n - w + 1;
    std::vector<int32_t> ws((size_t)nw), bs((size_t)nw), be((size_t)nw), bsc((size_t)nw);
    windows(p, m, t, n, w, L, ws.data(), bs.data(), be.data(), bsc.data());
    std::unordered_set<long long> seen;
    std::vector<Iv> w2, out;
    if (variant == 0) {
        for (int i = nw - 1; i >= 0; --i) {
            const double dist = ws[(size_t)i];
            if (dist >= kdi && !seen.count(key(bs[(size_t)i], be[(size_t)i]))) {
                w2.push_back(Iv{bs[(size_t)i], be[(size_t)i], (double)bsc[(size_t)i]});
                seen.insert(key(bs[(size_t)i], be[(size_t)i]));
            }
        }
        std::reverse(w2.begin(), w2.end());
    } else {
        std::vector<int> w1;
        for (int e = w; e < n + 1;) {
            const double dist = ws[(size_t)(e - w)];
            if (dist >= kdi) w1.push_back(e - w);
            // int abs on both operands, as the reference's unqualified abs binds
            if (dist < kdi - 1) e += std::max(1, (std::abs((int)dist) - std::abs((int)kdi)) / 2);
            else ++e;
        }
        for (int i : w1)
            if (!seen.count(key(bs[(size_t)i], be[(size_t)i]))) {
                seen.insert(key(bs[(size_t)i], be[(size_t)i]));
                w2.push_back(Iv{bs[(size_t)i], be[(size_t)i], (double)bsc[(size_t)i]});
            }
    }
    phase3(w2, out);
    for (size_t i = 0; i < out.size() && (int)i < cap; ++i) {
        start[i] = out[i].start;
        end[i] = out[i].end;
        score[i] = out[i].score;
    }
    return (int)out.size();
}

}  // extern "C"
