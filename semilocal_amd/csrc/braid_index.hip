// braid_index.hip -- O(log(m+n)) follow-up queries on a resident braid.
//
// Every query on a semi-local kernel P (row -> col, size N = m+n) reduces to
// one dominance count (reference approximate_matching/utils.h:268-359):
//
//     H(i, j) = j - i - #{r >= i + m : P[r] < j}
//     LCS(a, b[l:r))  = H(l, r)
//     LCS(a[s:e), b)  = H(-s, n + m - e) - s - (m - e)
//
// A wavelet matrix over the sequence P[0..N) answers #{r in [lo, hi) : P[r]
// < J} with one rank step per bit of the values (L = bits(N-1) levels): level
// l holds bit L-1-l of every value in the level's arrangement as 64-bit words
// with a prefix count of ones per word; the next level's arrangement is the
// stable partition of this one by that bit (zeros first).  The reference
// answers the same counts with a scan per query or a 2-D prefix structure
// (dominance_sum_queries.h:16-120); this is the GPU-side index for braids that
// stay resident.
//
// Build, per level (all on one stream): one wave per 64 values turns the bits
// into a word with a ballot and counts it; one block scans the counts; then
// every value moves to its slot in the next arrangement.
#include <algorithm>
#include <vector>

#include "../../include/semilocal_hip.h"
#include "braid_index.h"

namespace slhip {

namespace {

// bits of level `lvl` (bit `bit` of every value of `cur`) -> words, counts
__global__ __launch_bounds__(256) void wm_bits(const int32_t *__restrict__ cur, uint32_t n, int bit,
                                               uint64_t *__restrict__ words, uint32_t *__restrict__ cnt) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    const uint32_t w = i >> 6;
    const bool one = i < n && ((((uint32_t)cur[i]) >> bit) & 1u);
    const uint64_t word = __builtin_amdgcn_ballot_w64(one);
    if ((threadIdx.x & 63) == 0 && (uint64_t)w * 64u < n) {
        words[w] = word;
        cnt[w] = (uint32_t)__builtin_popcountll(word);
    }
}

// ranks[w] = ones in words [0, w) for w in [0, nw]; zeros = n - ones.  One block.
__global__ __launch_bounds__(1024) void wm_scan(const uint32_t *__restrict__ cnt, uint32_t nw, uint32_t n,
                                                uint32_t *__restrict__ ranks, uint32_t *__restrict__ zeros) {
    __shared__ uint32_t part[16];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nw; base += 1024) {
        const uint32_t k = base + threadIdx.x;
        const uint32_t v = k < nw ? cnt[k] : 0u;
        uint32_t inc = v;   // inclusive scan within the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t o = (uint32_t)__shfl_up((int)inc, off);
            if (lane >= off) inc += o;
        }
        if (lane == 63) part[wid] = inc;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wid; ++w) before += part[w];
        if (k < nw) ranks[k] = before + inc - v;
        uint32_t total = 0;
        for (int w = 0; w < 16; ++w) total += part[w];
        __syncthreads();
        carry += total;
    }
    if (threadIdx.x == 0) {
        ranks[nw] = carry;
        *zeros = n - carry;
    }
}

// stable partition of `cur` by the level's bit into `nxt` (zeros first)
__global__ __launch_bounds__(256) void wm_scatter(const int32_t *__restrict__ cur, uint32_t n,
                                                  const uint64_t *__restrict__ words,
                                                  const uint32_t *__restrict__ ranks,
                                                  const uint32_t *__restrict__ zeros, int bit,
                                                  int32_t *__restrict__ nxt) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const int32_t v = cur[i];
    const uint32_t r1 = wm_rank1(words, ranks, i);
    const uint32_t dst = ((((uint32_t)v) >> bit) & 1u) ? *zeros + r1 : i - r1;
    nxt[dst] = v;
}

__global__ __launch_bounds__(256) void wm_answer(BraidIndexView ix, int nq, const int32_t *__restrict__ kind,
                                                 const int32_t *__restrict__ qi, const int32_t *__restrict__ qj,
                                                 int64_t *__restrict__ out) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < nq) out[k] = wm_query(ix, kind[k], qi[k], qj[k]);
}

}  // namespace

size_t braid_index_bytes(int64_t size) {
    if (size <= 0) return 0;
    int levels = 1;
    while (levels < 31 && ((uint64_t)1 << levels) < (uint64_t)size) ++levels;
    const size_t nw = (size_t)((size + 63) / 64);
    const size_t b = sizeof(uint64_t) * nw * levels + sizeof(uint32_t) * (nw + 1) * levels + sizeof(uint32_t) * 64;
    return (b + 255) & ~(size_t)255;
}

size_t braid_index_scratch_bytes(int64_t size) {
    const size_t nw = (size_t)((std::max<int64_t>(size, 1) + 63) / 64);
    return ((2 * sizeof(int32_t) * (size_t)std::max<int64_t>(size, 1) + sizeof(uint32_t) * nw) + 255) & ~(size_t)255;
}

// Builds the index of `rc` (device, size m+n) into `mem` (braid_index_bytes)
// using `scratch` (braid_index_scratch_bytes), all on `stream`.
hipError_t braid_index_build(const int32_t *rc, int32_t m, int32_t n, void *mem, void *scratch,
                             hipStream_t stream, BraidIndexView *view) {
    const int64_t size = (int64_t)m + n;
    BraidIndexView v{};
    v.m = m;
    v.n = n;
    if (size <= 0) {
        *view = v;
        return hipSuccess;
    }
    int levels = 1;
    while (levels < 31 && ((uint64_t)1 << levels) < (uint64_t)size) ++levels;
    const uint32_t N = (uint32_t)size, nw = (N + 63u) / 64u;
    uint64_t *all_words = reinterpret_cast<uint64_t *>(mem);
    uint32_t *all_ranks = reinterpret_cast<uint32_t *>(all_words + (size_t)nw * levels);
    uint32_t *zeros = all_ranks + (size_t)(nw + 1) * levels;
    v.levels = levels;
    v.nw = nw;
    v.words = all_words;
    v.ranks = all_ranks;
    v.zeros = zeros;
    int32_t *a = reinterpret_cast<int32_t *>(scratch), *b = a + N;
    uint32_t *cnt = reinterpret_cast<uint32_t *>(b + N);
    hipError_t e = hipMemcpyAsync(a, rc, sizeof(int32_t) * N, hipMemcpyDeviceToDevice, stream);
    const unsigned blocks = (N + 255u) / 256u;
    for (int l = 0; l < levels && e == hipSuccess; ++l) {
        const int bit = levels - 1 - l;
        uint64_t *words = all_words + (size_t)l * nw;
        uint32_t *ranks = all_ranks + (size_t)l * (nw + 1);
        hipLaunchKernelGGL(wm_bits, dim3(blocks), dim3(256), 0, stream, a, N, bit, words, cnt);
        hipLaunchKernelGGL(wm_scan, dim3(1), dim3(1024), 0, stream, cnt, nw, N, ranks, zeros + l);
        if (l + 1 < levels) {
            hipLaunchKernelGGL(wm_scatter, dim3(blocks), dim3(256), 0, stream, a, N, words, ranks, zeros + l, bit, b);
            std::swap(a, b);
        }
        e = hipGetLastError();
    }
    *view = v;
    return e;
}

hipError_t braid_index_answer(const BraidIndexView &ix, int nq, const int32_t *kind, const int32_t *qi,
                              const int32_t *qj, int64_t *out, hipStream_t stream) {
    if (nq <= 0) return hipSuccess;
    hipLaunchKernelGGL(wm_answer, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, stream, ix, nq, kind, qi, qj,
                       out);
    return hipGetLastError();
}

}  // namespace slhip

// ---------------------------------------------------------------------------
// C ABI: a standalone index over a host permutation.
// ---------------------------------------------------------------------------
struct sl_braid_index {
    slhip::BraidIndexView view{};
    void *mem = nullptr;
    hipStream_t stream = nullptr;
};

extern "C" {

int sl_braid_index_create(sl_braid_index **out, const int32_t *row_to_col, int m, int n) {
    using namespace slhip;
    if (!out || m < 0 || n < 0 || ((int64_t)m + n > 0 && !row_to_col) || (int64_t)m + n >= (1LL << 31)) {
        set_error("sl_braid_index_create: bad argument");
        return kErrArg;
    }
    *out = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    const int64_t size = (int64_t)m + n;
    sl_braid_index *h = new sl_braid_index();
    void *scratch = nullptr;
    int32_t *d_rc = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess && size > 0) e = hipMalloc(&h->mem, braid_index_bytes(size));
    if (e == hipSuccess && size > 0) e = hipMalloc(&scratch, braid_index_scratch_bytes(size) + sizeof(int32_t) * size);
    if (e == hipSuccess && size > 0) {
        d_rc = reinterpret_cast<int32_t *>(reinterpret_cast<char *>(scratch) + braid_index_scratch_bytes(size));
        e = hipMemcpyAsync(d_rc, row_to_col, sizeof(int32_t) * size, hipMemcpyHostToDevice, h->stream);
    }
    if (e == hipSuccess) e = braid_index_build(d_rc, m, n, h->mem, scratch, h->stream, &h->view);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (scratch) (void)hipFree(scratch);
    if (e != hipSuccess) {
        set_error(std::string("sl_braid_index_create: ") + hipGetErrorString(e));
        if (h->mem) (void)hipFree(h->mem);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;
    }
    *out = h;
    return kOk;
}

int sl_braid_index_query(sl_braid_index *h, int nq, const int32_t *kind, const int32_t *qi, const int32_t *qj,
                         int64_t *out) {
    using namespace slhip;
    if (!h || nq < 0 || (nq > 0 && (!kind || !qi || !qj || !out))) {
        set_error("sl_braid_index_query: bad argument");
        return kErrArg;
    }
    if (nq == 0) return kOk;
    for (int k = 0; k < nq; ++k)
        if (kind[k] < SL_QUERY_H || kind[k] > SL_QUERY_SUBSTRING_STRING) {
            set_error("sl_braid_index_query: unknown query kind");
            return kErrArg;
        }
    if (h->view.m + (int64_t)h->view.n == 0) {   // empty kernel: H(i, j) = j - i
        for (int k = 0; k < nq; ++k) out[k] = kind[k] == SL_QUERY_SUBSTRING_STRING ? 0 : (int64_t)qj[k] - qi[k];
        return kOk;
    }
    const size_t in_bytes = sizeof(int32_t) * (size_t)nq;
    char *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, 3 * in_bytes + sizeof(int64_t) * (size_t)nq);
    int32_t *dk = reinterpret_cast<int32_t *>(d), *di = dk + nq, *dj = di + nq;
    int64_t *dout = reinterpret_cast<int64_t *>(d + 3 * in_bytes);
    if (e == hipSuccess) e = hipMemcpyAsync(dk, kind, in_bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(di, qi, in_bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dj, qj, in_bytes, hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = braid_index_answer(h->view, nq, dk, di, dj, dout, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, sizeof(int64_t) * (size_t)nq, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (d) (void)hipFree(d);
    if (e != hipSuccess) {
        set_error(std::string("sl_braid_index_query: ") + hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

int sl_braid_index_stats(const sl_braid_index *h, int *levels, int64_t *bytes) {
    using namespace slhip;
    if (!h) {
        set_error("sl_braid_index_stats: null handle");
        return kErrArg;
    }
    if (levels) *levels = h->view.levels;
    if (bytes) *bytes = (int64_t)braid_index_bytes((int64_t)h->view.m + h->view.n);
    return kOk;
}

void sl_braid_index_destroy(sl_braid_index *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->mem) (void)hipFree(h->mem);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

}  // extern "C"
