// hyyro.hip -- score-only LLCS by Hyyro's bit-vector recurrence, batched.
//
// For a pair (a, b) with b packed into K = ceil(n / 64) words and one match
// mask per symbol (M_s bit j = [b_j == s]), every symbol x of a updates the
// bit vector V (initially all ones):
//
//   p = M_x & V;   V = (V ^ p) | (V + p)        (one n-bit addition)
//
// and LLCS = the number of zero bits of V at the end (include/BitwiseLCS/
// BitParalellLCS.h:337-433, prefix_lcs/bitwise/transposition_network_binary_
// alphabet.h:743-916: hyyro_magic and its anti-diagonal form).
//
// The n-bit addition is the only coupling between words; two kernels handle
// it.
//  * hy_rows (pairs up to 64 words: one wave): lane l holds word l and a row
//    is two passes -- with carry-in 0 each lane learns whether it generates
//    a carry (g) and whether it is all ones after the add (q); two ballots
//    G, X = G | Q and one 64-bit add give every lane its carry-in (wave64
//    carry-lookahead: carry into lane l = bit l of (X + G) ^ X ^ G), and the
//    lane redoes its word with it.
//  * hy_sys (wider pairs, up to 2^20 columns): systolic -- lane L holds R
//    words and processes row t - L at step t, taking its carry from lane
//    L - 1 of the previous step by one DPP move; carries cross waves through
//    LDS once per 64 steps (the waves run one 64-step group apart).
// Masks come from a table built on the device (symbols of the packed string:
// at most 256 per pair) and are prefetched D rows ahead; rows whose symbol
// does not occur in the packed string are dropped on the host (their p is 0
// and they leave V unchanged).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "semilocal_hip.h"

namespace slhip {

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t stream);

namespace hy {

constexpr int kMaxSymbols = 256;
constexpr int kMaxColumns = 16 * 64 * 64 * 16;   // 16 waves x 64 lanes x 16 words: 2^20

struct HyPair {
    int64_t rows_off;   // row symbols (int32 codes into the pair's table) in the row pool
    int64_t tab_off;    // the pair's mask table (sigma x K words) in the table pool
    int64_t cols_off;   // column symbols (uint8 codes) in the column pool
    int32_t m, n, K, sigma;
    int32_t out;        // caller index
    int32_t pad_;
};

__global__ void hy_table(const HyPair *__restrict__ pairs, const uint8_t *__restrict__ cols,
                         uint64_t *__restrict__ tab) {
    const HyPair P = pairs[blockIdx.x];
    const uint8_t *b = cols + P.cols_off;
    const int64_t total = (int64_t)P.sigma * P.K;
    for (int64_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int s = (int)(idx / P.K), w = (int)(idx % P.K);
        const int j0 = 64 * w, j1 = min(P.n, j0 + 64);
        uint64_t mask = 0;
        for (int j = j0; j < j1; ++j) mask |= (uint64_t)(b[j] == s ? 1 : 0) << (j - j0);
        tab[P.tab_off + idx] = mask;
    }
}

// 64-bit add with carry in and out
__device__ __forceinline__ uint64_t add_c(uint64_t x, uint64_t y, int &c) {
    const uint64_t t = x + y;
    const int o1 = t < x ? 1 : 0;
    const uint64_t t2 = t + (uint64_t)c;
    const int o2 = t2 < t ? 1 : 0;
    c = o1 | o2;
    return t2;
}

template <int R, int D, int NW>
__global__ __launch_bounds__(64 * NW) void hy_rows(const HyPair *__restrict__ pairs, const int32_t *__restrict__ list,
                                                   const int32_t *__restrict__ rows,
                                                   const uint64_t *__restrict__ tab, const uint64_t *__restrict__ zero,
                                                   int32_t *__restrict__ llcs) {
    static_assert(64 % D == 0, "the row ring must divide the 64-row symbol chunks");
    static_assert(NW == 1, "one wave per pair: the carry-in of lane 0 is 0");
    const HyPair P = pairs[list[blockIdx.x]];
    const int lane = threadIdx.x & 63;
    const int w0 = lane * R;   // first word of this lane
    const uint64_t *T = tab + P.tab_off;
    const int32_t *A = rows + P.rows_off;
    const int K = P.K, m = P.m;
    uint64_t V[R];
#pragma unroll
    for (int r = 0; r < R; ++r) V[r] = ~0ull;
    // row symbols, 64 per lane-parallel load: rows [c, c + 64) and [c + 64, c + 128)
    // loads are unconditional (see hy_sys): symbols of rows past m are read
    // from a clamped index and never used
    int scur = A[lane < m ? lane : 0];
    int snxt = A[64 + lane < m ? 64 + lane : 0];
    // mask ring: masks of rows i .. i + D - 1
    uint64_t Mq[D][R];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int sd = __builtin_amdgcn_readlane(scur, d);
#pragma unroll
        for (int r = 0; r < R; ++r) Mq[d][r] = *(d < m && w0 + r < K ? T + (int64_t)sd * K + w0 + r : zero);
    }
    for (int i0 = 0; i0 < m; i0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = i0 + d;
            if (i < m) {   // uniform
                uint64_t p[R];
                int c = 0, all_ones = 1;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    p[r] = Mq[d][r] & V[r];
                    const uint64_t t = add_c(V[r], p[r], c);
                    all_ones &= t == ~0ull ? 1 : 0;
                }
                const uint64_t G = __ballot(c);
                const uint64_t X = G | __ballot(all_ones);
                const uint64_t C = (X + G) ^ X ^ G;
                c = (int)((C >> lane) & 1ull);
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const uint64_t t = add_c(V[r], p[r], c);
                    V[r] = (V[r] ^ p[r]) | t;
                }
                // refill this ring slot with row i + D (its symbol from the
                // lane buffers: (i + D) - c lies in [D, 64 + D) where c is the
                // 64-aligned chunk holding i)
                const int nx = i + D;
                if (nx < m) {
                    const int off = nx & 63;
                    const int sx = (nx >> 6) == (i >> 6) ? __builtin_amdgcn_readlane(scur, off)
                                                         : __builtin_amdgcn_readlane(snxt, off);
#pragma unroll
                    for (int r = 0; r < R; ++r) Mq[d][r] = *(w0 + r < K ? T + (int64_t)sx * K + w0 + r : zero);
                }
            }
        }
        if (((i0 + D) & 63) == 0) {   // next chunk of 64 rows: shift the symbol buffers
            scur = snxt;
            const int nr = i0 + D + 64 + lane;
            snxt = A[nr < m ? nr : 0];
        }
    }
    // LLCS = zero bits of V (padding bits beyond n stay one)
    int z = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) z += 64 - __popcll(V[r]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
    if (lane == 0) llcs[P.out] = z;
}

// Systolic form for wide pairs.  Lane L (0 .. 64 NW - 1) owns words
// [L R, L R + R) and processes row i = t - L at step t, so its carry-in is
// the carry-out of lane L - 1 at step t - 1 (same row): one DPP wave_shr:1
// per step inside a wave.  Across waves the carries of lane 63 are collected
// per group of 64 steps into a 64-bit word in LDS; wave w runs the steps
// [64 (g - w), 64 (g - w) + 64) in group g (one group behind wave w - 1), so
// everything it reads was written in groups g - 1 and g - 2 and one barrier
// per group suffices.  Masks of the row a lane meets at step t + D are
// gathered at step t into ring slot t mod D (uniform across lanes), their
// symbols D steps earlier still.
constexpr int kDppWaveShr1 = 0x138;

template <int R, int D, int NW>
__global__ __launch_bounds__(64 * NW) void hy_sys(const HyPair *__restrict__ pairs, const int32_t *__restrict__ list,
                                                  const int32_t *__restrict__ rows, const uint64_t *__restrict__ tab,
                                                  const uint64_t *__restrict__ zero, int32_t *__restrict__ llcs) {
    static_assert(64 % D == 0, "the ring must divide a group of 64 steps");
    __shared__ uint64_t hand[3][NW];   // lane 63's carry-outs per group (slot g mod 3)
    __shared__ int zsum[NW];
    const HyPair P = pairs[list[blockIdx.x]];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int L = wave * 64 + lane;
    const int w0 = L * R;
    const uint64_t *T = tab + P.tab_off;
    const int32_t *A = rows + P.rows_off;
    const int K = P.K, m = P.m;
    uint64_t V[R];
#pragma unroll
    for (int r = 0; r < R; ++r) V[r] = ~0ull;
    uint64_t Mq[D][R];
    int Sq[D];
    // step t of this wave is 64 (g - wave) + j; prime the rings for its first
    // 2D steps (t = -64 wave + 0 .. 2D - 1): rows below 0 have no masks
    const int t_first = -64 * wave;
    // Every load below is unconditional and its value is used as loaded:
    // masks of rows outside [0, m) or words past K come from `zero` (a word
    // holding 0 after the table pool), symbols of such rows from a clamped
    // index (they are never used).  Guarded loads would become branches, and
    // masking a loaded value makes the compiler wait for the load at once;
    // either empties the prefetch rings every step.
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int i = t_first + d - L;
        const bool ok = i >= 0 && i < m;
        const int s = A[ok ? i : 0];
#pragma unroll
        for (int r = 0; r < R; ++r) Mq[d][r] = *(ok && w0 + r < K ? T + (int64_t)s * K + w0 + r : zero);
        const int i2 = i + D;
        Sq[d] = A[i2 >= 0 && i2 < m ? i2 : 0];
    }
    const int ngroups = (NW - 1) + (m + 64 * NW - 1 + 63) / 64 + 1;
    int cprev = 0;   // this lane's carry-out at the previous step
    for (int g = 0; g < ngroups; ++g) {
        const int t0 = 64 * (g - wave);
        uint64_t in = 0;
        if (wave > 0) {
            const uint64_t prev1 = g >= 1 ? hand[(g - 1) % 3][wave - 1] : 0ull;
            const uint64_t prev2 = g >= 2 ? hand[(g - 2) % 3][wave - 1] : 0ull;
            in = (prev1 << 1) | (prev2 >> 63);
        }
        uint64_t outbits = 0;
        for (int j0 = 0; j0 < 64; j0 += D) {
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int j = j0 + d;
                const int t = t0 + j;
                const int i = t - L;
                const int from_left = __builtin_amdgcn_update_dpp(0, cprev, kDppWaveShr1, 0xf, 0xf, false);
                int c = lane == 0 ? (int)((in >> j) & 1ull) : from_left;
                if (i >= 0 && i < m) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const uint64_t p = Mq[d][r] & V[r];
                        const uint64_t t1 = V[r] + p;
                        const int o1 = t1 < V[r] ? 1 : 0;
                        const uint64_t t2 = t1 + (uint64_t)c;
                        const int o2 = t2 < t1 ? 1 : 0;
                        V[r] = (V[r] ^ p) | t2;
                        c = o1 | o2;
                    }
                } else {
                    c = 0;
                }
                cprev = c;
                outbits |= (uint64_t)(__builtin_amdgcn_readlane(c, 63) & 1) << j;
                // refill: masks for step t + D (symbol already in Sq[d]),
                // symbol for step t + 2D
                const int in_ = i + D;
                const bool ok = in_ >= 0 && in_ < m;
#pragma unroll
                for (int r = 0; r < R; ++r)
                    Mq[d][r] = *(ok && w0 + r < K ? T + (int64_t)Sq[d] * K + w0 + r : zero);
                const int i2 = i + 2 * D;
                Sq[d] = A[i2 >= 0 && i2 < m ? i2 : 0];
            }
        }
        if (lane == 0) hand[g % 3][wave] = outbits;
        __syncthreads();
    }
    int z = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) z += 64 - __popcll(V[r]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
    if (lane == 0) zsum[wave] = z;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int j = 0; j < NW; ++j) tot += zsum[j];
        llcs[P.out] = tot;
    }
}

// Segmented systolic form for pairs wider than 256 words: the packed string
// is cut into segments of 256 R words, each a 4-wave workgroup running the
// systolic scheme of hy_sys with its global wave index Wg = 4 seg + wave (so
// lanes, rows and the one-group skew continue across segments).  The carry
// into a segment's lane 0 is the previous segment's lane-63 output: its last
// wave publishes, per 64-step group, the 64 carry bits as two tagged 64-bit
// entries (high word: (epoch << 16) | group, low word: 32 carry bits),
// written and read with agent-scope atomics like the combing kernel's strand
// hand-off, so a reader that sees the tag has the bits.  Workgroups are
// persistent and take (pair, segment) tickets in order: a segment waits only
// for the segment before it, whose ticket a running workgroup already holds.
// The wait is bounded; running out sets ctl->error and the run reports a
// timeout.
struct SegItem {
    int32_t pair, seg;
};

constexpr int kSegWaves = 4;

__device__ __forceinline__ uint64_t seg_ld(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void seg_st(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int R, int D>
__global__ __launch_bounds__(64 * kSegWaves) void hy_seg(const HyPair *__restrict__ pairs,
                                                         const SegItem *__restrict__ items, uint32_t nitems,
                                                         const int64_t *__restrict__ xoff,
                                                         const int32_t *__restrict__ rows,
                                                         const uint64_t *__restrict__ tab,
                                                         const uint64_t *__restrict__ zero, uint64_t *__restrict__ xbuf,
                                                         uint32_t epoch, LaunchCtl *ctl, uint32_t spin_limit,
                                                         int32_t *__restrict__ llcs) {
    static_assert(64 % D == 0, "the ring must divide a group of 64 steps");
    constexpr int NW = kSegWaves;
    __shared__ uint64_t hand[3][NW];
    __shared__ uint32_t ticket;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // The ticket is read into a scalar register, so the compiler sees a
    // uniform loop exit, and thread 0's ticket add is the loop's only
    // thread-0 region: with a second one at the tail (a block-wide score
    // add), the compiler merged the two into an outer loop and left the
    // barriers and the segment work in an inner loop that only thread 0
    // exits at the end of an item -- the other threads then re-read the
    // unchanged ticket and redo the same segment without end.
    for (;;) {
        if (threadIdx.x == 0) ticket = atomicAdd(&ctl->ticket, 1u);
        __syncthreads();
        const uint32_t tk = __builtin_amdgcn_readfirstlane(ticket);
        __syncthreads();   // every wave has read the ticket before it is replaced
        if (tk >= nitems) break;
        const SegItem it = items[tk];
        const HyPair P = pairs[it.pair];
        const int seg = it.seg;
        const int K = P.K, m = P.m;
        const int nseg = (K + 64 * NW * R - 1) / (64 * NW * R);
        const int Wg = NW * seg + wave;
        const int L = 64 * Wg + lane;   // global lane of the pair
        const int w0 = L * R;
        const uint64_t *T = tab + P.tab_off;
        const int32_t *A = rows + P.rows_off;
        // groups this workgroup runs: wave Wg meets rows in groups [2 Wg, 2 Wg + c]
        const int c = (m + 63) / 64 + 1;
        const int G0 = 2 * NW * seg, G1 = 2 * NW * seg + 2 * (NW - 1) + c;
        const int ng = G1 - G0 + 1;   // the same for every segment of the pair
        uint64_t *xmine = xbuf + xoff[it.pair] + (int64_t)seg * ng * 2;
        const uint64_t *xprevseg = xbuf + xoff[it.pair] + (int64_t)(seg - 1) * ng * 2;
        uint64_t V[R];
#pragma unroll
        for (int r = 0; r < R; ++r) V[r] = ~0ull;
        uint64_t Mq[D][R];
        int Sq[D];
        const int t_first = 64 * (G0 - Wg);
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int i = t_first + d - L;
            const bool ok = i >= 0 && i < m;
            const int sy = A[ok ? i : 0];
#pragma unroll
            for (int r = 0; r < R; ++r) Mq[d][r] = *(ok && w0 + r < K ? T + (int64_t)sy * K + w0 + r : zero);
            const int i2 = i + D;
            Sq[d] = A[i2 >= 0 && i2 < m ? i2 : 0];
        }
        // the previous segment's carry word of group gp (zero outside the
        // groups it ran: rows there lie outside [0, m))
        auto seg_word = [&](int gp) -> uint64_t {
            if (seg == 0 || gp < G0 - 2 * NW || gp > G1 - 2 * NW) return 0ull;
            const uint64_t *e = xprevseg + (int64_t)(gp - (G0 - 2 * NW)) * 2;
            const uint32_t want = (epoch << 16) | ((uint32_t)gp & 0xFFFFu);
            uint64_t lo = seg_ld(e), hi = seg_ld(e + 1);
            uint32_t spins = 0;
            while ((uint32_t)(lo >> 32) != want || (uint32_t)(hi >> 32) != want) {
                if (++spins >= spin_limit ||
                    __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    return 0ull;   // results of this run are void and reported
                }
                __builtin_amdgcn_s_sleep(2);
                lo = seg_ld(e);
                hi = seg_ld(e + 1);
            }
            return ((hi & 0xFFFFFFFFull) << 32) | (lo & 0xFFFFFFFFull);
        };
        uint64_t xprev = wave == 0 ? seg_word(G0 - 2) : 0ull;
        int cprev = 0;
        for (int g = G0; g <= G1; ++g) {
            const int t0 = 64 * (g - Wg);
            uint64_t in = 0;
            if (wave > 0) {
                const uint64_t p1 = g - 1 >= G0 ? hand[(g - 1) % 3][wave - 1] : 0ull;
                const uint64_t p2 = g - 2 >= G0 ? hand[(g - 2) % 3][wave - 1] : 0ull;
                in = (p1 << 1) | (p2 >> 63);
            } else if (seg > 0) {
                const uint64_t x1 = seg_word(g - 1);
                in = (x1 << 1) | (xprev >> 63);
                xprev = x1;
            }
            uint64_t outbits = 0;
            for (int j0 = 0; j0 < 64; j0 += D) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const int j = j0 + d;
                    const int t = t0 + j;
                    const int i = t - L;
                    const int from_left = __builtin_amdgcn_update_dpp(0, cprev, kDppWaveShr1, 0xf, 0xf, false);
                    int cc = lane == 0 ? (int)((in >> j) & 1ull) : from_left;
                    if (i >= 0 && i < m) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            const uint64_t p = Mq[d][r] & V[r];
                            const uint64_t t1 = V[r] + p;
                            const int o1 = t1 < V[r] ? 1 : 0;
                            const uint64_t t2 = t1 + (uint64_t)cc;
                            const int o2 = t2 < t1 ? 1 : 0;
                            V[r] = (V[r] ^ p) | t2;
                            cc = o1 | o2;
                        }
                    } else {
                        cc = 0;
                    }
                    cprev = cc;
                    outbits |= (uint64_t)(__builtin_amdgcn_readlane(cc, 63) & 1) << j;
                    const int in_ = i + D;
                    const bool ok = in_ >= 0 && in_ < m;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        Mq[d][r] = *(ok && w0 + r < K ? T + (int64_t)Sq[d] * K + w0 + r : zero);
                    const int i2 = i + 2 * D;
                    Sq[d] = A[i2 >= 0 && i2 < m ? i2 : 0];
                }
            }
            if (lane == 0) {
                hand[g % 3][wave] = outbits;
                if (wave == NW - 1 && seg + 1 < nseg) {
                    const uint64_t tag = (uint64_t)((epoch << 16) | ((uint32_t)g & 0xFFFFu)) << 32;
                    uint64_t *e = xmine + (int64_t)(g - G0) * 2;
                    seg_st(e, tag | (outbits & 0xFFFFFFFFull));
                    seg_st(e + 1, tag | (outbits >> 32));
                }
            }
            __syncthreads();
        }
        int z = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) z += w0 + r < K ? 64 - __popcll(V[r]) : 0;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) z += __shfl_xor(z, o, 64);
        if (lane == 0) atomicAdd(&llcs[P.out], z);   // each wave adds its own zero count
    }
}

}  // namespace hy

// ---------------------------------------------------------------------------
// host side

struct Hyyro {
    int npairs = 0;
    hipStream_t stream = nullptr;
    std::vector<hy::HyPair> pairs;
    std::vector<std::vector<int32_t>> cls_list;   // launch class -> pair indices
    std::vector<int> wpl;                         // launch class -> words per lane of its kernel
    hy::HyPair *d_pairs = nullptr;
    int32_t *d_list = nullptr;          // the one-workgroup classes' lists, concatenated
    std::vector<int64_t> list_off;
    // segmented classes: (pair, segment) tickets for each segment width
    // (1 or 2 words per lane), exchange buffer (sized for the narrower
    // segments, which are more), control
    std::vector<std::vector<hy::SegItem>> seg_items;   // [class - first][R - 1]
    hy::SegItem *d_items = nullptr;
    std::vector<int64_t> item_off;
    int64_t *d_xoff = nullptr;          // per pair: its segments' exchange entries in d_xbuf
    uint64_t *d_xbuf = nullptr;
    int64_t xbuf_entries = 0;
    uint64_t cleared_at = 0;            // run serial when d_xbuf was last cleared
    LaunchCtl *d_ctl = nullptr;         // one per segmented class
    // a segment waits a few groups (well under a millisecond) for its
    // neighbour; 2^20 polls is over a second of margin and fails fast
    uint32_t spin_limit = 1u << 20;
    int32_t *d_rows = nullptr;
    uint8_t *d_cols = nullptr;
    uint64_t *d_tab = nullptr;
    int32_t *d_llcs = nullptr;
    int64_t tab_words = 0, word_rows = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // launch classes are independent: each runs on a stream of its own
    // (default priority), forked from and joined back into `stream`
    std::vector<hipStream_t> cls_stream;
    std::vector<hipEvent_t> cls_done;
    hipEvent_t fork = nullptr;
};

namespace {

// Launch classes (length buckets) by packed width K (words): up to 64, 128,
// 256, 2,048 and 16,384 words (2^20 columns).  Each class offers kernels that
// differ in words per lane (R), and the batch runs one of them per class:
// the default below, or the one sl_hyyro_autotune measured fastest on the
// batch's own pairs.
//  * up to 64 words: one wave of the lookahead kernel (hy_rows), R = 1 or 2;
//  * up to 128: the systolic kernel (hy_sys) on 2 waves (R = 1), or the
//    lookahead kernel on one wave (R = 2);
//  * up to 256: hy_sys on 4 waves (R = 1) or 2 waves (R = 2), or the
//    lookahead kernel on one wave (R = 4);
//  * wider pairs are cut into segments of 256 R words on workgroups of their
//    own (hy_seg, R = 1 or 2), so a long pair spreads over as many compute
//    units as it has segments.
int hy_class(int K) {
    static const int limits[] = {64, 128, 256, 2048, 16384};
    int c = 0;
    while (K > limits[c]) ++c;
    return c;
}
constexpr int kClasses = 5;
constexpr int kClassMaxWords[kClasses] = {64, 128, 256, 2048, 16384};
constexpr int kFirstSegClass = 3;
constexpr int kSegClasses = 2;   // classes 3 and 4
bool is_seg_class(int cls) { return cls >= kFirstSegClass && cls < kFirstSegClass + kSegClasses; }
// words per lane each class offers, default first
const std::vector<int> &class_wpl_options(int cls) {
    static const std::vector<int> opts[kClasses] = {{1, 2}, {1, 2}, {1, 2, 4}, {1, 2}, {2, 1}};
    return opts[cls];
}
bool class_offers(int cls, int R) {
    for (int r : class_wpl_options(cls))
        if (r == R) return true;
    return false;
}
// waves per pair (one-workgroup classes) or per segment (segmented ones)
int class_waves(int cls, int R) {
    if (is_seg_class(cls)) return 4;   // hy::kSegWaves
    if (cls == 0) return 1;
    if (cls == 1) return R == 1 ? 2 : 1;
    return R == 1 ? 4 : (R == 2 ? 2 : 1);
}

void hy_free(Hyyro *h) {
    if (!h) return;
    (void)hipFree(h->d_pairs);
    (void)hipFree(h->d_list);
    (void)hipFree(h->d_items);
    (void)hipFree(h->d_xoff);
    (void)hipFree(h->d_xbuf);
    (void)hipFree(h->d_ctl);
    (void)hipFree(h->d_rows);
    (void)hipFree(h->d_cols);
    (void)hipFree(h->d_tab);
    (void)hipFree(h->d_llcs);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->cls_done)
        if (e) (void)hipEventDestroy(e);
    if (h->fork) (void)hipEventDestroy(h->fork);
    for (auto &cs : h->cls_stream)
        if (cs) (void)hipStreamDestroy(cs);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

// b's symbols -> dense codes in order of first appearance; a's symbols ->
// codes, dropping those absent from b.  `table` maps a symbol value to its
// code (-1: none) for 1- and 2-byte symbols and is left all -1 on return.
template <class Sym>
int hy_encode(const Sym *a, int m, const Sym *b, int n, std::vector<int32_t> &rows, std::vector<uint8_t> &cols,
              int *sigma, std::vector<int> &table) {
    int count = 0;
    std::vector<uint32_t> used;
    auto finish = [&](int rc) {
        for (uint32_t v : used) table[v] = -1;
        return rc;
    };
    for (int j = 0; j < n; ++j) {
        const uint32_t v = (uint32_t)b[j];
        int c = table[v];
        if (c < 0) {
            if (count >= hy::kMaxSymbols) return finish(kErrUnsupported);
            c = table[v] = count++;
            used.push_back(v);
        }
        cols.push_back((uint8_t)c);
    }
    for (int i = 0; i < m; ++i) {
        const int c = table[(uint32_t)a[i]];
        if (c >= 0) rows.push_back(c);
    }
    *sigma = count;
    return finish(kOk);
}

int hy_encode_u32(const uint32_t *a, int m, const uint32_t *b, int n, std::vector<int32_t> &rows,
                  std::vector<uint8_t> &cols, int *sigma) {
    std::unordered_map<uint32_t, int> code;
    code.reserve(512);
    for (int j = 0; j < n; ++j) {
        auto it = code.find(b[j]);
        int c;
        if (it == code.end()) {
            c = (int)code.size();
            if (c >= hy::kMaxSymbols) return kErrUnsupported;
            code.emplace(b[j], c);
        } else {
            c = it->second;
        }
        cols.push_back((uint8_t)c);
    }
    for (int i = 0; i < m; ++i) {
        auto it = code.find(a[i]);
        if (it != code.end()) rows.push_back(it->second);
    }
    *sigma = (int)code.size();
    return kOk;
}

// Launches the classes in class_mask (bit c: class c) on their streams,
// forked from and joined back into h->stream.
int hy_launch(Hyyro *h, unsigned class_mask = ~0u) {
    // segmented pairs add their segments' zero counts: clear the scores, the
    // tickets and the error flags first
    SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_llcs, 0u, (uint64_t)std::max(h->npairs, 1), h->stream));
    SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_ctl, 0u, 2u * kSegClasses, h->stream));
    uint32_t epoch = 0;
    if (next_run_tag(h->cleared_at, &epoch) && h->xbuf_entries > 0)
        SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_xbuf, 0u, 2u * (uint64_t)h->xbuf_entries, h->stream));
    SL_HIP_TRY(hipEventRecord(h->fork, h->stream));
    const int cus = device_cus();
    // widest class first: it holds the longest pairs
    for (int c = kClasses - 1; c >= 0; --c) {
        if (!((class_mask >> c) & 1u)) continue;
        hipStream_t cs = h->cls_stream[(size_t)c];
        const int R = h->wpl[(size_t)c];
        if (is_seg_class(c)) {
            const size_t k = (size_t)(c - kFirstSegClass);
            const size_t v = (size_t)(R - 1);
            const uint32_t nit = (uint32_t)h->seg_items[2 * k + v].size();
            if (!nit) continue;
            SL_HIP_TRY(hipStreamWaitEvent(cs, h->fork, 0));
            const hy::SegItem *items = h->d_items + h->item_off[2 * k + v];
            const unsigned blocks = (unsigned)std::min<int64_t>(nit, 2 * cus);
#define SL_HY_SEG(R, D)                                                                                         \
    hipLaunchKernelGGL((hy::hy_seg<R, D>), dim3(blocks), dim3(64 * hy::kSegWaves), 0, cs, h->d_pairs, items, nit, \
                       h->d_xoff, h->d_rows, h->d_tab, h->d_tab + h->tab_words, h->d_xbuf, epoch, h->d_ctl + k,     \
                       h->spin_limit, h->d_llcs)
            if (R == 1) SL_HY_SEG(1, 32);
            else SL_HY_SEG(2, 16);
#undef SL_HY_SEG
        } else {
            const unsigned cnt = (unsigned)h->cls_list[(size_t)c].size();
            if (!cnt) continue;
            SL_HIP_TRY(hipStreamWaitEvent(cs, h->fork, 0));
            const int32_t *lst = h->d_list + h->list_off[(size_t)c];
#define SL_HY_LAUNCH(KERNEL, R, D, NW)                                                                       \
    hipLaunchKernelGGL((hy::KERNEL<R, D, NW>), dim3(cnt), dim3(64 * NW), 0, cs, h->d_pairs, lst, h->d_rows, \
                       h->d_tab, h->d_tab + h->tab_words, h->d_llcs)
            // every form covers the class's widest pair: hy_rows 64 R words,
            // hy_sys 64 NW R words
            if (c == 0) {
                if (R == 1) SL_HY_LAUNCH(hy_rows, 1, 16, 1);
                else SL_HY_LAUNCH(hy_rows, 2, 16, 1);
            } else if (c == 1) {
                if (R == 1) SL_HY_LAUNCH(hy_sys, 1, 32, 2);
                else SL_HY_LAUNCH(hy_rows, 2, 16, 1);
            } else {
                if (R == 1) SL_HY_LAUNCH(hy_sys, 1, 32, 4);
                else if (R == 2) SL_HY_LAUNCH(hy_sys, 2, 16, 2);
                else SL_HY_LAUNCH(hy_rows, 4, 8, 1);
            }
#undef SL_HY_LAUNCH
        }
        SL_HIP_TRY(hipGetLastError());
        SL_HIP_TRY(hipEventRecord(h->cls_done[(size_t)c], cs));
        SL_HIP_TRY(hipStreamWaitEvent(h->stream, h->cls_done[(size_t)c], 0));
    }
    return kOk;
}

// After a run (stream synchronised): a bounded wait of a segmented class ran out?
int hy_check(Hyyro *h) {
    LaunchCtl ctl[kSegClasses];
    SL_HIP_TRY(hipMemcpy(ctl, h->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost));
    for (auto &c : ctl)
        if (c.error) {
            set_error("Hyyro LLCS: a segment's bounded wait for its neighbour timed out; results are void");
            return kErrTimeout;
        }
    return kOk;
}

int hy_create(Hyyro **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
              const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes) {
    *out = nullptr;
    if (npairs < 0 || (npairs > 0 && (!a_cat || !a_off || !a_len || !b_cat || !b_off || !b_len)) ||
        (sym_bytes != 1 && sym_bytes != 2 && sym_bytes != 4)) {
        set_error("sl_hyyro_create: bad argument");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    Hyyro *h = new Hyyro();
    h->npairs = npairs;
    h->cls_list.assign(kClasses, {});
    std::vector<int32_t> rows;
    std::vector<uint8_t> cols;
    std::vector<int> table(sym_bytes == 1 ? 256 : (sym_bytes == 2 ? 65536 : 0), -1);
    for (int i = 0; i < npairs; ++i) {
        int m = a_len[i], n = b_len[i];
        if (m < 0 || n < 0) {
            hy_free(h);
            set_error("sl_hyyro_create: negative length");
            return kErrArg;
        }
        const char *pa = (const char *)a_cat + (size_t)a_off[i] * sym_bytes;
        const char *pb = (const char *)b_cat + (size_t)b_off[i] * sym_bytes;
        if (m > n) {   // pack the longer string into the bit vector
            std::swap(m, n);
            std::swap(pa, pb);
        }
        if (n > hy::kMaxColumns) {
            hy_free(h);
            set_error("sl_hyyro_create: a string longer than 2^20 symbols (use the semi-local kernel)");
            return kErrUnsupported;
        }
        hy::HyPair P{};
        P.rows_off = (int64_t)rows.size();
        P.cols_off = (int64_t)cols.size();
        P.n = n;
        P.K = (n + 63) / 64;
        P.out = i;
        int sigma = 0;
        if (sym_bytes == 1)
            rc = hy_encode((const uint8_t *)pa, m, (const uint8_t *)pb, n, rows, cols, &sigma, table);
        else if (sym_bytes == 2)
            rc = hy_encode((const uint16_t *)pa, m, (const uint16_t *)pb, n, rows, cols, &sigma, table);
        else
            rc = hy_encode_u32((const uint32_t *)pa, m, (const uint32_t *)pb, n, rows, cols, &sigma);
        if (rc) {
            hy_free(h);
            set_error("sl_hyyro_create: more than 256 distinct symbols in one string (use the semi-local kernel)");
            return rc;
        }
        P.m = (int)((int64_t)rows.size() - P.rows_off);
        P.sigma = sigma;
        P.tab_off = h->tab_words;
        h->tab_words += (int64_t)sigma * P.K;
        h->word_rows += (int64_t)P.m * P.K;
        h->pairs.push_back(P);
        h->cls_list[(size_t)hy_class(P.K)].push_back(i);
    }
    std::vector<int32_t> all_lists;
    h->list_off.assign(kClasses, 0);
    for (int c = 0; c < kClasses; ++c) {
        if (is_seg_class(c)) continue;
        h->list_off[(size_t)c] = (int64_t)all_lists.size();
        all_lists.insert(all_lists.end(), h->cls_list[(size_t)c].begin(), h->cls_list[(size_t)c].end());
    }
    // segmented classes: pairs by work (rows x words) descending, each pair's
    // segments in order; exchange entries: per segment, 2 per group it runs
    std::vector<hy::SegItem> all_items;
    std::vector<int64_t> xoff((size_t)std::max(npairs, 1), 0);
    int64_t xtotal = 0;
    h->seg_items.assign(2 * kSegClasses, {});
    h->item_off.assign(2 * kSegClasses, 0);
    for (int c = kFirstSegClass; c < kFirstSegClass + kSegClasses; ++c) {
        auto lst = h->cls_list[(size_t)c];
        std::stable_sort(lst.begin(), lst.end(), [&](int x, int y) {
            return (int64_t)h->pairs[(size_t)x].m * h->pairs[(size_t)x].K >
                   (int64_t)h->pairs[(size_t)y].m * h->pairs[(size_t)y].K;
        });
        const size_t k = (size_t)(c - kFirstSegClass);
        for (int pi : lst) {   // exchange entries for the narrower (R = 1) segments, which are more
            const hy::HyPair &P = h->pairs[(size_t)pi];
            const int nseg = (P.K + 64 * hy::kSegWaves - 1) / (64 * hy::kSegWaves);
            const int64_t ng = (P.m + 63) / 64 + 8;   // groups a segment runs (hy_seg: G1 - G0 + 1)
            xoff[(size_t)pi] = xtotal;
            xtotal += (int64_t)nseg * ng * 2;
        }
        for (int R = 1; R <= 2; ++R) {
            const int segw = 64 * hy::kSegWaves * R;
            auto &items = h->seg_items[2 * k + (size_t)(R - 1)];
            for (int pi : lst) {
                const int nseg = (h->pairs[(size_t)pi].K + segw - 1) / segw;
                for (int sg = 0; sg < nseg; ++sg) items.push_back(hy::SegItem{pi, sg});
            }
            h->item_off[2 * k + (size_t)(R - 1)] = (int64_t)all_items.size();
            all_items.insert(all_items.end(), items.begin(), items.end());
        }
    }
    h->wpl.assign(kClasses, 0);
    for (int c = 0; c < kClasses; ++c) h->wpl[(size_t)c] = class_wpl_options(c)[0];
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (auto &ev : h->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
    h->cls_stream.assign(kClasses, nullptr);
    h->cls_done.assign(kClasses, nullptr);
    for (int c = 0; c < kClasses && e == hipSuccess; ++c) {
        if (h->cls_list[(size_t)c].empty()) continue;
        e = hipStreamCreateWithFlags(&h->cls_stream[(size_t)c], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->cls_done[(size_t)c], hipEventDisableTiming);
    }
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 8));
    };
    alloc((void **)&h->d_pairs, sizeof(hy::HyPair) * h->pairs.size());
    alloc((void **)&h->d_list, sizeof(int32_t) * all_lists.size());
    alloc((void **)&h->d_items, sizeof(hy::SegItem) * all_items.size());
    alloc((void **)&h->d_xoff, sizeof(int64_t) * xoff.size());
    alloc((void **)&h->d_xbuf, sizeof(uint64_t) * (size_t)xtotal);
    alloc((void **)&h->d_ctl, sizeof(LaunchCtl) * kSegClasses);
    alloc((void **)&h->d_rows, sizeof(int32_t) * (rows.size() + 1));   // + a spare: clamped loads read index 0
    alloc((void **)&h->d_cols, cols.size());
    alloc((void **)&h->d_tab, sizeof(uint64_t) * (size_t)(h->tab_words + 1));   // + the zero word
    alloc((void **)&h->d_llcs, sizeof(int32_t) * (size_t)std::max(npairs, 1));
    if (e == hipSuccess && !h->pairs.empty())
        e = hipMemcpy(h->d_pairs, h->pairs.data(), sizeof(hy::HyPair) * h->pairs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !all_lists.empty())
        e = hipMemcpy(h->d_list, all_lists.data(), sizeof(int32_t) * all_lists.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !all_items.empty())
        e = hipMemcpy(h->d_items, all_items.data(), sizeof(hy::SegItem) * all_items.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_xoff, xoff.data(), sizeof(int64_t) * xoff.size(), hipMemcpyHostToDevice);
    // exchange entries start with tag 0, which no run's (nonzero) epoch matches
    if (e == hipSuccess && xtotal > 0) e = launch_fill_u32((uint32_t *)h->d_xbuf, 0u, 2u * (uint64_t)xtotal, h->stream);
    h->xbuf_entries = xtotal;
    h->cleared_at = run_serial_now();
    if (e == hipSuccess && !rows.empty())
        e = hipMemcpy(h->d_rows, rows.data(), sizeof(int32_t) * rows.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !cols.empty()) e = hipMemcpy(h->d_cols, cols.data(), cols.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)   // the zero word after the table pool (masks of rows and words outside a pair)
        e = hipMemsetAsync(h->d_tab + h->tab_words, 0, sizeof(uint64_t), h->stream);
    if (e == hipSuccess && npairs > 0) {
        hipLaunchKernelGGL(hy::hy_table, dim3((unsigned)npairs), dim3(256), 0, h->stream, h->d_pairs, h->d_cols,
                           h->d_tab);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        set_error(std::string("sl_hyyro_create: ") + hipGetErrorString(e));
        hy_free(h);
        return kErrHip;
    }
    *out = h;
    return kOk;
}

}  // namespace
}  // namespace slhip

using namespace slhip;

extern "C" {

int sl_hyyro_create(sl_hyyro **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                    const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes) {
    if (!out) {
        set_error("sl_hyyro_create: null handle pointer");
        return kErrArg;
    }
    Hyyro *h = nullptr;
    const int rc = hy_create(&h, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes);
    *out = reinterpret_cast<sl_hyyro *>(h);
    return rc;
}

int sl_hyyro_run(sl_hyyro *hh) {
    Hyyro *h = reinterpret_cast<Hyyro *>(hh);
    if (!h) {
        set_error("sl_hyyro_run: null handle");
        return kErrArg;
    }
    return hy_launch(h);
}

int sl_hyyro_result(sl_hyyro *hh, int32_t *llcs) {
    Hyyro *h = reinterpret_cast<Hyyro *>(hh);
    if (!h || (h->npairs > 0 && !llcs)) {
        set_error("sl_hyyro_result: null argument");
        return kErrArg;
    }
    if (h->npairs == 0) return kOk;
    SL_HIP_TRY(hipMemcpyAsync(llcs, h->d_llcs, sizeof(int32_t) * (size_t)h->npairs, hipMemcpyDeviceToHost,
                              h->stream));
    SL_HIP_TRY(hipStreamSynchronize(h->stream));
    return hy_check(h);
}

int sl_hyyro_time(sl_hyyro *hh, int warmup, int iters, double *ms_per_run) {
    Hyyro *h = reinterpret_cast<Hyyro *>(hh);
    if (!h || iters < 1 || warmup < 0 || !ms_per_run) {
        set_error("sl_hyyro_time: bad argument");
        return kErrArg;
    }
    int rc = kOk;
    for (int r = 0; r < warmup && !rc; ++r) rc = hy_launch(h);
    if (rc) return rc;
    SL_HIP_TRY(hipEventRecord(h->ev[0], h->stream));
    for (int r = 0; r < iters && !rc; ++r) rc = hy_launch(h);
    SL_HIP_TRY(hipEventRecord(h->ev[1], h->stream));
    if (rc) return rc;
    SL_HIP_TRY(hipEventSynchronize(h->ev[1]));
    float ms = 0;
    SL_HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
    *ms_per_run = ms / iters;
    return hy_check(h);
}

int sl_hyyro_buckets(const sl_hyyro *hh, int32_t *max_words, int32_t *pairs, int32_t *words_per_lane,
                     int32_t *waves) {
    const Hyyro *h = reinterpret_cast<const Hyyro *>(hh);
    if (!h) {
        set_error("sl_hyyro_buckets: null handle");
        return kErrArg;
    }
    for (int c = 0; c < kClasses; ++c) {
        if (max_words) max_words[c] = kClassMaxWords[c];
        if (pairs) pairs[c] = (int32_t)h->cls_list[(size_t)c].size();
        if (words_per_lane) words_per_lane[c] = h->wpl[(size_t)c];
        if (waves) waves[c] = class_waves(c, h->wpl[(size_t)c]);
    }
    return kOk;
}

int sl_hyyro_set_words_per_lane(sl_hyyro *hh, const int32_t *words_per_lane) {
    Hyyro *h = reinterpret_cast<Hyyro *>(hh);
    if (!h || !words_per_lane) {
        set_error("sl_hyyro_set_words_per_lane: null argument");
        return kErrArg;
    }
    for (int c = 0; c < kClasses; ++c)
        if (words_per_lane[c] != 0 && !class_offers(c, words_per_lane[c])) {
            set_error("sl_hyyro_set_words_per_lane: bucket " + std::to_string(c) + " offers no kernel with " +
                      std::to_string(words_per_lane[c]) + " words per lane");
            return kErrArg;
        }
    for (int c = 0; c < kClasses; ++c)
        if (words_per_lane[c] != 0) h->wpl[(size_t)c] = words_per_lane[c];
    return kOk;
}

// Per bucket with pairs: each offered words-per-lane kernel runs that bucket
// alone, timed with HIP events (one warm-up launch, then two timed launches
// per round); the fastest minimum over the rounds is kept.
int sl_hyyro_autotune(sl_hyyro *hh, int rounds, int32_t *words_per_lane) {
    Hyyro *h = reinterpret_cast<Hyyro *>(hh);
    if (!h || rounds < 1) {
        set_error("sl_hyyro_autotune: bad argument");
        return kErrArg;
    }
    for (int c = 0; c < kClasses; ++c) {
        if (h->cls_list[(size_t)c].empty()) continue;
        const std::vector<int> &opts = class_wpl_options(c);
        std::vector<double> best(opts.size(), 1e30);
        for (int round = 0; round < rounds; ++round)
            for (size_t o = 0; o < opts.size(); ++o) {
                h->wpl[(size_t)c] = opts[o];
                int rc = hy_launch(h, 1u << c);
                if (rc) return rc;
                SL_HIP_TRY(hipEventRecord(h->ev[0], h->stream));
                for (int it = 0; it < 2 && !rc; ++it) rc = hy_launch(h, 1u << c);
                if (rc) return rc;
                SL_HIP_TRY(hipEventRecord(h->ev[1], h->stream));
                SL_HIP_TRY(hipEventSynchronize(h->ev[1]));
                rc = hy_check(h);
                if (rc) return rc;
                float ms = 0;
                SL_HIP_TRY(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
                best[o] = std::min(best[o], (double)ms / 2);
            }
        size_t pick = 0;
        for (size_t o = 1; o < opts.size(); ++o)
            if (best[o] < best[pick]) pick = o;
        h->wpl[(size_t)c] = opts[pick];
    }
    if (words_per_lane)
        for (int c = 0; c < kClasses; ++c) words_per_lane[c] = h->wpl[(size_t)c];
    return kOk;
}

int sl_hyyro_stats(const sl_hyyro *hh, int64_t *word_rows, int64_t *table_words) {
    const Hyyro *h = reinterpret_cast<const Hyyro *>(hh);
    if (!h) {
        set_error("sl_hyyro_stats: null handle");
        return kErrArg;
    }
    if (word_rows) *word_rows = h->word_rows;
    if (table_words) *table_words = h->tab_words;
    return kOk;
}

void sl_hyyro_destroy(sl_hyyro *hh) { hy_free(reinterpret_cast<Hyyro *>(hh)); }

int sl_llcs_hyyro_batch(int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                        const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes, int32_t *llcs) {
    sl_hyyro *h = nullptr;
    int rc = sl_hyyro_create(&h, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes);
    if (!rc) rc = sl_hyyro_run(h);
    if (!rc) rc = sl_hyyro_result(h, llcs);
    sl_hyyro_destroy(h);
    return rc;
}

int sl_llcs_hyyro(const void *a, int m, const void *b, int n, int sym_bytes, int *llcs) {
    if (!llcs || m < 0 || n < 0 || (m > 0 && !a) || (n > 0 && !b)) {
        set_error("sl_llcs_hyyro: bad argument");
        return kErrArg;
    }
    static const uint32_t dummy = 0;
    const int64_t off = 0;
    int32_t out = 0;
    const int rc = sl_llcs_hyyro_batch(1, m ? a : &dummy, &off, &m, n ? b : &dummy, &off, &n, sym_bytes, &out);
    if (!rc) *llcs = out;
    return rc;
}

}  // extern "C"
