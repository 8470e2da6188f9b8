// amatch.hip -- approximate matching over resident braids.
//
// A pattern p (m symbols) is matched against every window [i, i+w) of a text
// (n symbols).  For the scoring scheme of the reference's duplicate searches
// (match 0, mismatch -2, gap -1; approximate_matching/ApproximateMatching.h
// :252-432) the normalised mismatch is 0, so the blown-up strings equal the
// originals and every score is read from the semi-local kernel of (p, text):
//
//   score(s, e) = 2 LCS(p, text[s, e)) - m - (e - s)      (= -indel distance)
//   LCS(p, text[s, e)) = (e - s) - D(s, e),  D(s, e) = #{r >= s + m : P[r] < e}
//
// (approximate_matching/utils.h:268-359: canonicalDecompositionWithKnown and
// FixedScoringScheme::getOriginalScoreFunc.)
//
// Device side.  The text is cut into overlapping chunks of B + W - 1 columns
// (W = the longest window the context serves), so every substring of a window
// lies inside one chunk; all (p, chunk) pairs are combed in ONE launch of the
// ragged batch kernel and stay in HBM.  The window kernel then gives one wave
// a run of consecutive windows of one chunk and walks them right to left
// keeping, per column e of the current window, D(i, e) and the best (score,
// length) over the starts seen so far -- the reference's per-column "stacks"
// (ApproximateMatching.h:373-413), held in registers, S = w - L + 1 columns
// spread over the 64 lanes.  Each step a 64-lane max gives the window's best
// substring; the window's own score is the column e = i + w.
//
// Host side: the duplicate searches' sequential phases (the skip heuristic of
// phase 1, deduplication, phase 3's sort and greedy pass) run on the per-window
// arrays exactly as the reference runs them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "common.h"
#include "semilocal_hip.h"

namespace slhip {

struct Batch;
int batch_create_flags_internal(Batch **out, int npairs, const void *a_cat, const int64_t *a_off,
                                const int32_t *a_len, const void *b_cat, const int64_t *b_off, const int32_t *b_len,
                                int sym_bytes, int flags);
int batch_launch_internal(Batch *bt);
int batch_wait_internal(Batch *bt);
void batch_destroy_internal(Batch *bt);
hipStream_t batch_stream_internal(const Batch *bt);
const int32_t *batch_rc_internal(const Batch *bt);
const int32_t *batch_cr_internal(const Batch *bt);
int64_t batch_perm_off_internal(const Batch *bt, int pair);
int64_t batch_cells_internal(const Batch *bt);

namespace am {

constexpr int kMaxColumns = 4096;   // S = w - L + 1 columns per window: at most 64 per lane

// One wave's work: chunk-local windows [lo, hi) of one chunk.
struct AmItem {
    int64_t perm_off;   // the chunk's braid in the batch output (m + nc entries, row->col and col->row)
    int32_t nc;         // chunk text length
    int32_t lo, hi;
    int32_t gbase;      // text offset of the chunk (global index of its window 0)
};

// One step of a 64-bit max over DPP-moved lanes; rows outside ROW_MASK (and
// lanes without a source) keep their own value.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint64_t dpp_max_step(uint64_t v) {
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    const uint32_t tlo = (uint32_t)__builtin_amdgcn_update_dpp(lo, lo, CTRL, ROW_MASK, 0xf, false);
    const uint32_t thi = (uint32_t)__builtin_amdgcn_update_dpp(hi, hi, CTRL, ROW_MASK, 0xf, false);
    const uint64_t t = ((uint64_t)thi << 32) | tlo;
    return t > v ? t : v;
}

// Max over the 64 lanes, returned to every lane: quad swaps, half-row and
// row mirrors leave each 16-lane row's max in all its lanes; row_bcast:15 and
// row_bcast:31 carry it into rows 1 and 3, then 2 and 3, so lane 63 holds the
// wave's max.  VALU moves only (no LDS round trips).
__device__ inline uint64_t wave_max_u64(uint64_t v) {
    v = dpp_max_step<0xB1, 0xf>(v);    // quad_perm [1,0,3,2]
    v = dpp_max_step<0x4E, 0xf>(v);    // quad_perm [2,3,0,1]
    v = dpp_max_step<0x141, 0xf>(v);   // row_half_mirror
    v = dpp_max_step<0x140, 0xf>(v);   // row_mirror
    v = dpp_max_step<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
    v = dpp_max_step<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
    return ((uint64_t)hi << 32) | lo;
}

__device__ inline int wave_sum_i32(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Column slot q (0 <= q < S) holds the window column e with e = q (mod S):
// the S columns of a window are consecutive, so the slots never collide and a
// step (i -> i - 1) retires column i + w exactly where column i - 1 + L
// enters.  Lane l holds slots l, l + 64, ..., in registers: per slot its
// column E, D = D(i, E) and best = ((score + bias) << len_bits) | (len - L),
// the lexicographic max of (score, length) over the starts s in [i, E - L].
// The reference keeps a start only if its score is higher, or equal with a
// longer substring (DuplicateSearch::compare, ApproximateMatching.h:144-148);
// across columns the window's best is the max of (score, length, end): its
// loop visits ends from the right and keeps the first of equal (score,
// length) (ApproximateMatching.h:392-405).
//
// Two running diagonals live on the scalar unit:
//   D(i, i+L) = D(i+1, i+1+L) + [P[i+m] < i+1+L] - [P^-1[i+L] >= i+m]   (entering column)
//   D(i, i+w) = D(i+1, i+1+w) + [P[i+m] < i+1+w] - [P^-1[i+w] >= i+m]   (the window itself)
constexpr int kNoColumn = 0x7fffffff;

__device__ __forceinline__ int count_below(const int32_t *rc, int r0, int r1, int bound, int lane) {
    int cnt = 0;
#pragma unroll 8
    for (int r = r0 + lane; r < r1; r += 64) cnt += rc[r] < bound ? 1 : 0;
    return wave_sum_i32(cnt);
}

template <int SPT>
__global__ __launch_bounds__(64) void am_windows(const int32_t *__restrict__ rc_all,
                                                 const int32_t *__restrict__ cr_all,
                                                 const AmItem *__restrict__ items, int m, int w, int L, int S,
                                                 int bias, int len_bits, int32_t *__restrict__ win_score,
                                                 int32_t *__restrict__ best_start, int32_t *__restrict__ best_end,
                                                 int32_t *__restrict__ best_score) {
    const AmItem it = items[blockIdx.x];
    const int lane = threadIdx.x;
    const int32_t *rc = rc_all + it.perm_off;
    const int32_t *cr = cr_all + it.perm_off;
    const int lo = it.lo, hi = it.hi, nc = it.nc;
    // warm-up from i_start: the window hi - 1 needs its columns' starts from
    // e - L down to hi - 1, and its rightmost column (hi - 1 + w) enters at
    // i = hi - 1 + w - L.  i_start + L <= nc because hi - 1 <= nc - w.
    const int i_start = hi + S - 2;
    int ddiag = count_below(rc, i_start + m, m + nc, i_start + L, lane);
    int dwin = count_below(rc, hi - 1 + m, m + nc, hi - 1 + w, lane);

    int E[SPT], D[SPT];
    uint32_t B[SPT];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
        E[k] = kNoColumn;
        D[k] = 0;
        B[k] = 0;
    }
    int qe = (i_start + L) % S;   // slot of the column entering at step i
    int pbuf = 0, cbuf = 0, wbuf = 0;

    // warm-up: columns enter, no window is complete yet
    for (int i = i_start; i >= hi; --i) {
        const int j = (i_start - i) & 63;
        if (j == 0) {
            // P[i_start + m] is never used (it feeds step i_start - 1 onwards)
            // and lies past the braid when L = 0 and i_start = nc
            const int src = i - lane;
            pbuf = (src >= hi && src < nc) ? rc[src + m] : 0;
            cbuf = src >= hi ? cr[src + L] : 0;
        }
        const int pv = __builtin_amdgcn_readlane(pbuf, j);
        const int cv = __builtin_amdgcn_readlane(cbuf, j);
        if (i < i_start) ddiag += (pv < i + 1 + L ? 1 : 0) - (cv >= i + m ? 1 : 0);
        const int A = bias - i - m, Bc = -i - L;
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int q = lane + 64 * k;
            if (k < SPT / 2 || q < S) {   // S > 32 SPT: groups below SPT / 2 are full
                const bool enter = q == qe;
                const bool live = E[k] != kNoColumn;
                const int e = enter ? i + L : E[k];
                const int d = enter ? ddiag : D[k] + (pv < e ? 1 : 0);
                const uint32_t cand = ((uint32_t)(e - 2 * d + A) << len_bits) | (uint32_t)(e + Bc);
                if (enter || live) {
                    E[k] = e;
                    D[k] = d;
                    B[k] = enter ? cand : (cand > B[k] ? cand : B[k]);
                }
            }
        }
        qe = qe == 0 ? S - 1 : qe - 1;
    }

    // windows hi - 1 .. lo: every slot holds a column of the window
    const uint32_t len_mask = (1u << len_bits) - 1u;
    int r_score = 0, r_start = 0, r_end = 0, r_win = 0;
    for (int i = hi - 1; i >= lo; --i) {
        const int idx = hi - 1 - i;
        const int j = idx & 63;
        if (j == 0) {
            const int src = i - lane;
            pbuf = src >= lo ? rc[src + m] : 0;
            cbuf = src >= lo ? cr[src + L] : 0;
            wbuf = src >= lo ? cr[src + w] : 0;
        }
        const int pv = __builtin_amdgcn_readlane(pbuf, j);
        const int cv = __builtin_amdgcn_readlane(cbuf, j);
        const int cw = __builtin_amdgcn_readlane(wbuf, j);
        if (i < i_start) ddiag += (pv < i + 1 + L ? 1 : 0) - (cv >= i + m ? 1 : 0);
        if (i < hi - 1) dwin += (pv < i + 1 + w ? 1 : 0) - (cw >= i + m ? 1 : 0);
        const int A = bias - i - m, Bc = -i - L;
        uint64_t kmax = 0;
#pragma unroll
        for (int k = 0; k < SPT; ++k) {
            const int q = lane + 64 * k;
            if (k < SPT / 2 || q < S) {   // S > 32 SPT: groups below SPT / 2 are full
                const bool enter = q == qe;
                const int e = enter ? i + L : E[k];
                const int d = enter ? ddiag : D[k] + (pv < e ? 1 : 0);
                const uint32_t cand = ((uint32_t)(e - 2 * d + A) << len_bits) | (uint32_t)(e + Bc);
                const uint32_t b = enter ? cand : (cand > B[k] ? cand : B[k]);
                E[k] = e;
                D[k] = d;
                B[k] = b;
                const uint64_t key = ((uint64_t)b << 32) | (uint32_t)(e - i);
                kmax = key > kmax ? key : kmax;
            }
        }
        kmax = wave_max_u64(kmax);
        if (lane == j) {
            const uint32_t b = (uint32_t)(kmax >> 32);
            const int e = i + (int)(uint32_t)kmax;
            r_score = (int)(b >> len_bits) - bias;
            r_end = e;
            r_start = e - ((int)(b & len_mask) + L);
            r_win = w - 2 * dwin - m;
        }
        if (j == 63 || i == lo) {
            if (lane <= j) {
                const int64_t g = (int64_t)it.gbase + (hi - 1 - (idx - j + lane));
                win_score[g] = r_win;
                best_score[g] = r_score;
                best_start[g] = it.gbase + r_start;
                best_end[g] = it.gbase + r_end;
            }
        }
        qe = qe == 0 ? S - 1 : qe - 1;
    }
}

static int bits_for(int64_t v) {   // bits to hold 0..v
    int b = 0;
    while (b < 62 && (int64_t(1) << b) <= v) ++b;
    return b;
}

// Selection of the windows whose score reaches a threshold, in window
// order: the duplicate searches only ever keep those (a few percent of the
// windows), so only they cross to the host.  Blocks of kSelThreads threads
// take kSelPer consecutive windows per thread; a count pass, one scan of the
// block counts, and a write pass that places each selected window at its
// block's offset plus its rank in the block.
constexpr int kSelThreads = 256, kSelPer = 4, kSelBlock = kSelThreads * kSelPer;

// exclusive prefix of v over the block's threads; *total receives the sum
__device__ __forceinline__ int sel_block_scan(int v, int *tmp /* kSelThreads / 64 */, int *total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tmp[wave] = x;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int k = 0; k < kSelThreads / 64; ++k) {
        before += k < wave ? tmp[k] : 0;
        all += tmp[k];
    }
    *total = all;
    return before + x - v;
}

__global__ __launch_bounds__(kSelThreads) void am_sel_count(const int32_t *__restrict__ ws, int nw, double kdi,
                                                             int32_t *__restrict__ cnt) {
    __shared__ int tmp[kSelThreads / 64];
    const int64_t i0 = (int64_t)blockIdx.x * kSelBlock + (int64_t)threadIdx.x * kSelPer;
    int c = 0;
#pragma unroll
    for (int e = 0; e < kSelPer; ++e) c += (i0 + e < nw && (double)ws[i0 + e] >= kdi) ? 1 : 0;
    int total = 0;
    (void)sel_block_scan(c, tmp, &total);
    if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// one block: cnt[0 .. nb) -> exclusive offsets, cnt[nb] = the total
__global__ __launch_bounds__(kSelThreads) void am_sel_scan(int32_t *__restrict__ cnt, int nb) {
    __shared__ int tmp[kSelThreads / 64];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kSelThreads) {   // uniform
        const int b = b0 + (int)threadIdx.x;
        const int v = b < nb ? cnt[b] : 0;
        int total = 0;
        const int ex = sel_block_scan(v, tmp, &total);
        if (b < nb) cnt[b] = carry + ex;
        carry += total;
        __syncthreads();   // tmp is reused by the next chunk
    }
    if (threadIdx.x == 0) cnt[nb] = carry;
}

// selected window i -> (i, best start, best end, best score) at its rank
__global__ __launch_bounds__(kSelThreads) void am_sel_write(const int32_t *__restrict__ ws,
                                                             const int32_t *__restrict__ bs,
                                                             const int32_t *__restrict__ be,
                                                             const int32_t *__restrict__ bsc, int nw, double kdi,
                                                             const int32_t *__restrict__ off, int4 *__restrict__ out) {
    __shared__ int tmp[kSelThreads / 64];
    const int64_t i0 = (int64_t)blockIdx.x * kSelBlock + (int64_t)threadIdx.x * kSelPer;
    bool sel[kSelPer];
    int c = 0;
#pragma unroll
    for (int e = 0; e < kSelPer; ++e) {
        sel[e] = i0 + e < nw && (double)ws[i0 + e] >= kdi;
        c += sel[e] ? 1 : 0;
    }
    int total = 0;
    int pos = off[blockIdx.x] + sel_block_scan(c, tmp, &total);
#pragma unroll
    for (int e = 0; e < kSelPer; ++e)
        if (sel[e]) {
            const int64_t i = i0 + e;
            out[pos++] = make_int4((int)i, bs[i], be[i], bsc[i]);
        }
}

}  // namespace am

// ---------------------------------------------------------------------------
// host object

struct AmInterval {   // the reference's Interval (approximate_matching/utils.h:242-262)
    int start;
    int end;
    double score;
    int len() const { return end - start; }
};

struct AMatch {
    int m = 0, n = 0, max_window = 0, chunk = 0, nchunks = 0;
    Batch *batch = nullptr;
    hipStream_t stream = nullptr;
    std::vector<int32_t> chunk_len;
    std::vector<int64_t> perm_off;
    am::AmItem *d_items = nullptr;
    size_t items_cap = 0;
    int32_t *d_out = nullptr;   // window score, best start, best end, best score (4 x cap)
    int32_t *h_out = nullptr;   // pinned host copy of d_out
    int64_t out_cap = 0;
    // the windows a duplicate search keeps: block counts / offsets, the
    // selected (window, best start, best end, best score), pinned copies
    int32_t *d_selcnt = nullptr, *h_seltotal = nullptr;
    int4 *d_sel = nullptr, *h_sel = nullptr;
    int64_t sel_cap = 0;
    std::vector<am::AmItem> h_items;   // work list of the last (w, L) launched
    int items_w = -1, items_L = -1;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<AmInterval> results;
    int64_t cells = 0;
    double last_device_ms = 0, last_host_ms = 0;   // last search: kernel + copy, then the host phases
};

static void amatch_free(AMatch *a) {
    if (!a) return;
    (void)hipFree(a->d_items);
    (void)hipFree(a->d_out);
    if (a->h_out) (void)hipHostFree(a->h_out);
    (void)hipFree(a->d_selcnt);
    (void)hipFree(a->d_sel);
    if (a->h_seltotal) (void)hipHostFree(a->h_seltotal);
    if (a->h_sel) (void)hipHostFree(a->h_sel);
    for (auto &e : a->ev)
        if (e) (void)hipEventDestroy(e);
    if (a->batch) batch_destroy_internal(a->batch);
    delete a;
}

// Launches the window kernel for (w, L) on the context's stream; outputs in
// a->d_out (4 arrays of n - w + 1).
static int amatch_launch(AMatch *a, int w, int L) {
    const int S = w - L + 1;
    const int nw = a->n - w + 1;
    const int bias = a->m + w + 1;
    const int len_bits = std::max(1, am::bits_for(S - 1));
    // one wave per run of windows: enough waves to fill the device (8 per
    // CU), runs at least 2S long so the warm-up stays below half the work
    const int cus = device_cus();
    int64_t run = std::max<int64_t>(2 * (int64_t)S, ((int64_t)nw + 8 * cus - 1) / (8 * cus));
    run = std::max<int64_t>(64, (run + 63) / 64 * 64);
    if (w != a->items_w || L != a->items_L) {
        // the work list depends on (w, L) only; rebuilt (after the stream has
        // drained, so no earlier upload still reads it) when they change
        SL_HIP_TRY(hipStreamSynchronize(a->stream));
        a->items_w = a->items_L = -1;
        a->h_items.clear();
        for (int c = 0; c < a->nchunks; ++c) {
            const int64_t base = (int64_t)c * a->chunk;
            const int64_t nwc = std::min<int64_t>(a->chunk, (int64_t)nw - base);
            for (int64_t lo = 0; lo < nwc; lo += run)
                a->h_items.push_back(am::AmItem{a->perm_off[(size_t)c], a->chunk_len[(size_t)c], (int32_t)lo,
                                                (int32_t)std::min<int64_t>(nwc, lo + run), (int32_t)base});
        }
        const size_t nitems = a->h_items.size();
        if (nitems > a->items_cap) {
            (void)hipFree(a->d_items);
            a->d_items = nullptr;
            a->items_cap = 0;
            SL_HIP_TRY(hipMalloc((void **)&a->d_items, sizeof(am::AmItem) * nitems));
            a->items_cap = nitems;
        }
        if ((int64_t)nw > a->out_cap) {
            (void)hipFree(a->d_out);
            if (a->h_out) (void)hipHostFree(a->h_out);
            a->d_out = a->h_out = nullptr;
            a->out_cap = 0;
            SL_HIP_TRY(hipMalloc((void **)&a->d_out, sizeof(int32_t) * 4 * (size_t)nw));
            // cached pinned memory: the host reads it only after the stream has synced
            SL_HIP_TRY(hipHostMalloc((void **)&a->h_out, sizeof(int32_t) * 4 * (size_t)nw, hipHostMallocNonCoherent));
            a->out_cap = nw;
        }
        SL_HIP_TRY(hipMemcpyAsync(a->d_items, a->h_items.data(), sizeof(am::AmItem) * nitems,
                                  hipMemcpyHostToDevice, a->stream));
        a->items_w = w;
        a->items_L = L;
    }
    const size_t nitems = a->h_items.size();
    const int32_t *rc = batch_rc_internal(a->batch);
    const int32_t *cr = batch_cr_internal(a->batch);
    int32_t *o = a->d_out;
    const int64_t cap = a->out_cap;
    int32_t *ws = o, *bs = o + cap, *be = o + 2 * cap, *bsc = o + 3 * cap;
    const dim3 grid((unsigned)nitems), block(64);
#define SL_AM_LAUNCH(SPT)                                                                                   \
    hipLaunchKernelGGL(am::am_windows<SPT>, grid, block, 0, a->stream, rc, cr, a->d_items, a->m, w, L, S, \
                       bias, len_bits, ws, bs, be, bsc)
    if (S <= 64) SL_AM_LAUNCH(1);
    else if (S <= 128) SL_AM_LAUNCH(2);
    else if (S <= 256) SL_AM_LAUNCH(4);
    else if (S <= 512) SL_AM_LAUNCH(8);
    else if (S <= 1024) SL_AM_LAUNCH(16);
    else if (S <= 2048) SL_AM_LAUNCH(32);
    else SL_AM_LAUNCH(64);
#undef SL_AM_LAUNCH
    SL_HIP_TRY(hipGetLastError());
    return kOk;
}

static int amatch_check_window(const AMatch *a, int w, int L, const char *who) {
    if (w < 1 || w > a->n || w > a->max_window) {
        set_error(std::string(who) + ": window length must be in [1, min(text length, max_window)]");
        return kErrArg;
    }
    if (L < 0 || L > w) {
        set_error(std::string(who) + ": min_len must be in [0, w]");
        return kErrArg;
    }
    const int S = w - L + 1;
    if (S > am::kMaxColumns) {
        set_error(std::string(who) + ": w - min_len + 1 exceeds 4096 window columns");
        return kErrUnsupported;
    }
    if (am::bits_for((int64_t)a->m + w + 1) + std::max(1, am::bits_for(S - 1)) > 32) {
        set_error(std::string(who) + ": pattern + window too long for the 32-bit (score, length) key");
        return kErrUnsupported;
    }
    return kOk;
}

// Runs the window kernel and copies its four arrays (score, best start, best
// end, best score; n - w + 1 entries each, out_cap apart) to the pinned host
// buffer; returns it in *host.
static int amatch_windows_host(AMatch *a, int w, int L, const int32_t **host) {
    int rc = amatch_launch(a, w, L);
    if (rc) return rc;
    SL_HIP_TRY(hipMemcpyAsync(a->h_out, a->d_out, sizeof(int32_t) * 4 * (size_t)a->out_cap, hipMemcpyDeviceToHost,
                              a->stream));
    SL_HIP_TRY(hipStreamSynchronize(a->stream));
    *host = a->h_out;
    return kOk;
}

static long long interval_key(int l, int r) {   // DuplicateSearch::buildKey (ApproximateMatching.h:151-154)
    long long key = l;
    return (key << 32) | r;
}

// Set of interval keys (both halves non-negative, so -1 marks an empty slot):
// open addressing with linear probing, kept at most half full.
class KeySet {
public:
    bool insert(long long key) {   // false if present
        if (2 * (used_ + 1) > slots_.size()) grow();
        return put(key);
    }

private:
    std::vector<long long> slots_;
    size_t used_ = 0;
    size_t mask() const { return slots_.size() - 1; }
    size_t home(long long key) const {
        return (size_t)(((unsigned long long)key * 0x9E3779B97F4A7C15ull) >> 20) & mask();
    }
    bool put(long long key) {
        for (size_t h = home(key);; h = (h + 1) & mask()) {
            if (slots_[h] == key) return false;
            if (slots_[h] == -1) {
                slots_[h] = key;
                ++used_;
                return true;
            }
        }
    }
    void grow() {
        std::vector<long long> old;
        old.swap(slots_);
        slots_.assign(old.empty() ? 1024 : 2 * old.size(), -1);
        used_ = 0;
        for (long long k : old)
            if (k != -1) put(k);
    }
};

// Runs the window kernel, selects on the device the windows whose score
// reaches kdi (in window order), and copies them -- and, with `with_scores`,
// every window's score -- to pinned host memory: *sel, *nsel, *scores.
static int amatch_select_host(AMatch *a, int w, int L, double kdi, bool with_scores, const int4 **sel, int *nsel,
                              const int32_t **scores) {
    int rc = amatch_launch(a, w, L);
    if (rc) return rc;
    const int nw = a->n - w + 1;
    const int nb = (nw + am::kSelBlock - 1) / am::kSelBlock;
    if (a->sel_cap < a->out_cap) {
        (void)hipFree(a->d_selcnt);
        (void)hipFree(a->d_sel);
        if (a->h_seltotal) (void)hipHostFree(a->h_seltotal);
        if (a->h_sel) (void)hipHostFree(a->h_sel);
        a->d_selcnt = a->h_seltotal = nullptr;
        a->d_sel = a->h_sel = nullptr;
        a->sel_cap = 0;
        const int64_t nbcap = (a->out_cap + am::kSelBlock - 1) / am::kSelBlock + 1;
        SL_HIP_TRY(hipMalloc((void **)&a->d_selcnt, sizeof(int32_t) * (size_t)nbcap));
        SL_HIP_TRY(hipMalloc((void **)&a->d_sel, sizeof(int4) * (size_t)a->out_cap));
        // cached pinned memory: the host reads it only after the stream has synced
        SL_HIP_TRY(hipHostMalloc((void **)&a->h_seltotal, sizeof(int32_t), hipHostMallocNonCoherent));
        SL_HIP_TRY(hipHostMalloc((void **)&a->h_sel, sizeof(int4) * (size_t)a->out_cap, hipHostMallocNonCoherent));
        a->sel_cap = a->out_cap;
    }
    const int64_t cap = a->out_cap;
    const int32_t *o = a->d_out;
    hipLaunchKernelGGL(am::am_sel_count, dim3((unsigned)nb), dim3(am::kSelThreads), 0, a->stream, o, nw, kdi,
                       a->d_selcnt);
    hipLaunchKernelGGL(am::am_sel_scan, dim3(1), dim3(am::kSelThreads), 0, a->stream, a->d_selcnt, nb);
    hipLaunchKernelGGL(am::am_sel_write, dim3((unsigned)nb), dim3(am::kSelThreads), 0, a->stream, o, o + cap,
                       o + 2 * cap, o + 3 * cap, nw, kdi, a->d_selcnt, a->d_sel);
    SL_HIP_TRY(hipGetLastError());
    SL_HIP_TRY(hipMemcpyAsync(a->h_seltotal, a->d_selcnt + nb, sizeof(int32_t), hipMemcpyDeviceToHost, a->stream));
    if (with_scores)
        SL_HIP_TRY(hipMemcpyAsync(a->h_out, o, sizeof(int32_t) * (size_t)nw, hipMemcpyDeviceToHost, a->stream));
    SL_HIP_TRY(hipStreamSynchronize(a->stream));
    const int total = *a->h_seltotal;
    if (total < 0 || total > nw) {
        set_error("sl_duplicate_search: window selection count out of range");
        return kErrHip;
    }
    if (total > 0) {
        SL_HIP_TRY(hipMemcpyAsync(a->h_sel, a->d_sel, sizeof(int4) * (size_t)total, hipMemcpyDeviceToHost,
                                  a->stream));
        SL_HIP_TRY(hipStreamSynchronize(a->stream));
    }
    *sel = a->h_sel;
    *nsel = total;
    *scores = with_scores ? a->h_out : nullptr;
    return kOk;
}

// DuplicateSearch::phase3 (ApproximateMatching.h:212-228): sort by start, then
// score descending (std::sort, so the input order decides among equal keys),
// then keep clones that start strictly right of the previous kept start and
// not left of any kept end.
static void phase3(std::vector<AmInterval> &set_w3, std::vector<AmInterval> &result) {
    std::sort(set_w3.begin(), set_w3.end(), [](const AmInterval &a1, const AmInterval &a2) {
        return (a1.start < a2.start) || (a1.start == a2.start && a1.score > a2.score);
    });
    int left_border = 0, right_border = 0;
    for (auto &clone : set_w3) {
        if (clone.start > left_border && clone.start >= right_border) {
            result.push_back(clone);
            left_border = clone.start;
            right_border = std::max(right_border, clone.end);
        }
    }
}

// The searches' parameters (ApproximateMatching.h:343-355, 258-260), in the
// same double arithmetic, with no fused multiply-add (the reference's default
// x86-64 build has none).
static void search_params(int m, double k, int *w, int *left_w, double *kdi) {
#pragma clang fp contract(off)
    *w = int(m / k);
    *left_w = int(m * k) - 1;
    *kdi = -(m * (1 / k + 1) * (1 - k * k));
}

static int duplicate_search(AMatch *a, double k, int variant) {
    a->results.clear();
    if (!(k > 0.0 && k <= 1.0) || (variant != 0 && variant != 1)) {
        set_error("sl_duplicate_search: k must be in (0, 1] and variant 0 (semi-dense) or 1 (semi-sparse)");
        return kErrArg;
    }
    int w = 0, left_w = 0;
    double kdi = 0;
    search_params(a->m, k, &w, &left_w, &kdi);
    const int L = left_w + 1;
    if (L < 0 || L > w) {
        set_error("sl_duplicate_search: int(m k) must lie in [0, int(m / k)]");
        return kErrArg;
    }
    int rc = amatch_check_window(a, w, L, "sl_duplicate_search");
    if (rc) return rc;
    // Only the windows whose score reaches kdi are kept by either search;
    // they come from the device already selected, in window order.  The
    // semi-sparse search also walks the scores of the windows it skips over.
    const int4 *sel = nullptr;
    const int32_t *ws = nullptr;
    int nsel = 0;
    const auto t0 = std::chrono::steady_clock::now();
    rc = amatch_select_host(a, w, L, kdi, variant == 1, &sel, &nsel, &ws);
    if (rc) return rc;
    const auto t1 = std::chrono::steady_clock::now();
    const int n = a->n;
    KeySet seen;
    std::vector<AmInterval> set_w2;
    if (variant == 0) {
        // InteractiveDuplicateSearchSemiDense::find (ApproximateMatching.h:339-432):
        // windows from the right end of the text to the left; a window whose
        // own score reaches kdi contributes its best substring once; then the
        // list is reversed and passed to phase 3.
        long long last = -1;   // the previous window's key: neighbours mostly share one
        for (int k = nsel - 1; k >= 0; --k) {
            const int4 e = sel[k];   // (window, best start, best end, best score)
            const long long key = interval_key(e.y, e.z);
            if (key == last) continue;
            last = key;
            if (seen.insert(key)) set_w2.push_back(AmInterval{e.y, e.z, (double)e.w});
        }
        std::reverse(set_w2.begin(), set_w2.end());
    } else {
        // InteractiveDuplicateSearchSemiSparse::find (ApproximateMatching.h:252-336):
        // phase 1 walks window ends left to right and skips ahead while the
        // window is far below kdi.  Its skip is int arithmetic: the unqualified
        // abs there binds to int abs(int), so both operands are truncated to
        // int before the halving.
        std::vector<int> set_w1;
        for (int end = w; end < n + 1;) {
            const double dist = ws[end - w];
            if (dist >= kdi) set_w1.push_back(end - w);
            if (dist < kdi - 1) {
                end += std::max(1, (std::abs((int)dist) - std::abs((int)kdi)) / 2);
            } else {
                ++end;
            }
        }
        // phase 2: each kept window's best substring, once; the kept windows
        // are selected ones, in increasing order, so one pass over the
        // selection finds them
        int k = 0;
        long long last = -1;
        for (int i : set_w1) {
            while (k < nsel && sel[k].x < i) ++k;
            if (k >= nsel || sel[k].x != i) {
                set_error("sl_duplicate_search: a kept window is missing from the device selection");
                return kErrHip;
            }
            const int4 e = sel[k];
            const long long key = interval_key(e.y, e.z);
            if (key == last) continue;
            last = key;
            if (seen.insert(key)) set_w2.push_back(AmInterval{e.y, e.z, (double)e.w});
        }
    }
    phase3(set_w2, a->results);
    const auto t2 = std::chrono::steady_clock::now();
    a->last_device_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    a->last_host_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
    return kOk;
}

}  // namespace slhip

using namespace slhip;

extern "C" {

int sl_amatch_create(sl_amatch **out, const void *p, int m, const void *text, int n, int sym_bytes, int max_window,
                     int chunk) {
    if (!out) {
        set_error("sl_amatch_create: null handle pointer");
        return kErrArg;
    }
    *out = nullptr;
    if (m < 1 || n < 1 || !p || !text || max_window < 1 || chunk < 0 ||
        (sym_bytes != 1 && sym_bytes != 2 && sym_bytes != 4)) {
        set_error("sl_amatch_create: bad argument (m, n >= 1, max_window >= 1, sym_bytes 1/2/4)");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    AMatch *a = new AMatch();
    a->m = m;
    a->n = n;
    a->max_window = std::min(max_window, n);
    // windows per chunk: default about 4 chunks per CU, at least the window
    // length (so the overlap costs at most as much as the chunk) and 1024
    int64_t B = chunk;
    if (B == 0) {
        B = std::max<int64_t>({(int64_t)a->max_window, 1024, ((int64_t)n + 4 * device_cus() - 1) / (4 * device_cus())});
        B = (B + 255) / 256 * 256;
    }
    a->chunk = (int)std::min<int64_t>(B, n);
    a->nchunks = (int)(((int64_t)n + a->chunk - 1) / a->chunk);
    std::vector<int64_t> a_off((size_t)a->nchunks, 0), b_off((size_t)a->nchunks);
    std::vector<int32_t> a_len((size_t)a->nchunks, m);
    a->chunk_len.resize((size_t)a->nchunks);
    for (int c = 0; c < a->nchunks; ++c) {
        const int64_t off = (int64_t)c * a->chunk;
        b_off[(size_t)c] = off;
        a->chunk_len[(size_t)c] = (int32_t)std::min<int64_t>((int64_t)a->chunk + a->max_window - 1, n - off);
    }
    rc = batch_create_flags_internal(&a->batch, a->nchunks, p, a_off.data(), a_len.data(), text, b_off.data(),
                                     a->chunk_len.data(), sym_bytes, SL_BATCH_WITH_COL_TO_ROW);
    if (rc) {
        amatch_free(a);
        return rc;
    }
    a->stream = batch_stream_internal(a->batch);
    a->cells = batch_cells_internal(a->batch);
    a->perm_off.resize((size_t)a->nchunks);
    for (int c = 0; c < a->nchunks; ++c) a->perm_off[(size_t)c] = batch_perm_off_internal(a->batch, c);
    for (auto &e : a->ev) {
        if (hipEventCreate(&e) != hipSuccess) {
            amatch_free(a);
            set_error("sl_amatch_create: hipEventCreate failed");
            return kErrHip;
        }
    }
    rc = batch_launch_internal(a->batch);
    if (!rc) rc = batch_wait_internal(a->batch);
    if (rc) {
        amatch_free(a);
        return rc;
    }
    *out = reinterpret_cast<sl_amatch *>(a);
    return kOk;
}

int sl_amatch_stats(const sl_amatch *h, int *chunks, int *chunk_windows, int64_t *cells) {
    const AMatch *a = reinterpret_cast<const AMatch *>(h);
    if (!a) {
        set_error("sl_amatch_stats: null handle");
        return kErrArg;
    }
    if (chunks) *chunks = a->nchunks;
    if (chunk_windows) *chunk_windows = a->chunk;
    if (cells) *cells = a->cells;
    return kOk;
}

int sl_amatch_windows(sl_amatch *h, int w, int min_len, int32_t *window_score, int32_t *best_start,
                      int32_t *best_end, int32_t *best_score) {
    AMatch *a = reinterpret_cast<AMatch *>(h);
    if (!a || !window_score || !best_start || !best_end || !best_score) {
        set_error("sl_amatch_windows: null argument");
        return kErrArg;
    }
    int rc = amatch_check_window(a, w, min_len, "sl_amatch_windows");
    if (rc) return rc;
    const int32_t *hbuf = nullptr;
    rc = amatch_windows_host(a, w, min_len, &hbuf);
    if (rc) return rc;
    const size_t bytes = sizeof(int32_t) * (size_t)(a->n - w + 1);
    const int64_t cap = a->out_cap;
    std::memcpy(window_score, hbuf, bytes);
    std::memcpy(best_start, hbuf + cap, bytes);
    std::memcpy(best_end, hbuf + 2 * cap, bytes);
    std::memcpy(best_score, hbuf + 3 * cap, bytes);
    return kOk;
}

int sl_amatch_time(sl_amatch *h, int w, int min_len, int warmup, int iters, double *comb_ms, double *windows_ms) {
    AMatch *a = reinterpret_cast<AMatch *>(h);
    if (!a || iters < 1 || warmup < 0) {
        set_error("sl_amatch_time: bad argument");
        return kErrArg;
    }
    int rc = amatch_check_window(a, w, min_len, "sl_amatch_time");
    if (rc) return rc;
    float ms = 0;
    if (comb_ms) {
        for (int r = 0; r < warmup && !rc; ++r) rc = batch_launch_internal(a->batch);
        SL_HIP_TRY(hipEventRecord(a->ev[0], a->stream));
        for (int r = 0; r < iters && !rc; ++r) rc = batch_launch_internal(a->batch);
        SL_HIP_TRY(hipEventRecord(a->ev[1], a->stream));
        if (!rc) rc = batch_wait_internal(a->batch);
        if (rc) return rc;
        SL_HIP_TRY(hipEventSynchronize(a->ev[1]));
        SL_HIP_TRY(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
        *comb_ms = ms / iters;
    }
    if (windows_ms) {
        for (int r = 0; r < warmup && !rc; ++r) rc = amatch_launch(a, w, min_len);
        if (rc) return rc;
        SL_HIP_TRY(hipEventRecord(a->ev[0], a->stream));
        for (int r = 0; r < iters && !rc; ++r) rc = amatch_launch(a, w, min_len);
        SL_HIP_TRY(hipEventRecord(a->ev[1], a->stream));
        if (rc) return rc;
        SL_HIP_TRY(hipEventSynchronize(a->ev[1]));
        SL_HIP_TRY(hipEventElapsedTime(&ms, a->ev[0], a->ev[1]));
        *windows_ms = ms / iters;
    }
    return kOk;
}

int sl_duplicate_search(sl_amatch *h, double k, int variant, int *count) {
    AMatch *a = reinterpret_cast<AMatch *>(h);
    if (!a || !count) {
        set_error("sl_duplicate_search: null argument");
        return kErrArg;
    }
    *count = 0;
    const int rc = duplicate_search(a, k, variant);
    if (rc) return rc;
    *count = (int)a->results.size();
    return kOk;
}

int sl_duplicate_results(const sl_amatch *h, int32_t *start, int32_t *end, double *score) {
    const AMatch *a = reinterpret_cast<const AMatch *>(h);
    if (!a || (!a->results.empty() && (!start || !end || !score))) {
        set_error("sl_duplicate_results: null argument");
        return kErrArg;
    }
    for (size_t i = 0; i < a->results.size(); ++i) {
        start[i] = a->results[i].start;
        end[i] = a->results[i].end;
        score[i] = a->results[i].score;
    }
    return kOk;
}

int sl_amatch_last_times(const sl_amatch *h, double *device_ms, double *host_ms) {
    const AMatch *a = reinterpret_cast<const AMatch *>(h);
    if (!a) {
        set_error("sl_amatch_last_times: null handle");
        return kErrArg;
    }
    if (device_ms) *device_ms = a->last_device_ms;
    if (host_ms) *host_ms = a->last_host_ms;
    return kOk;
}

void sl_amatch_destroy(sl_amatch *h) { amatch_free(reinterpret_cast<AMatch *>(h)); }

}  // extern "C"
