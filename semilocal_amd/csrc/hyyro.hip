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
// The n-bit addition is the only coupling between words.  One systolic
// kernel (hy_seg) handles every width: lane L holds R words (R = 1 or 2) and
// processes row t - L at step t, so its carry-in is lane L - 1's carry-out
// of the previous step.  The carries of a wave live in one 64-bit lane mask
// in scalar registers: a lane's addition is a chain of v_addc_co_u32 taking
// its carry from that mask, and the next step's mask is the carry-out mask
// shifted up by one lane.  Carries cross waves of a workgroup through LDS
// once per 64-step group, and cross workgroups (segments of wide pairs)
// through tagged global entries.  Masks come from a table built on the
// device (symbols of the packed string: at most 256 per pair) -- staged in
// LDS per segment when it fits -- and are read ahead of their step; rows
// whose symbol does not occur in the packed string are dropped on the host
// (their p is 0 and they leave V unchanged).
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
constexpr int kMaxColumns = 1 << 20;   // the widest class: 16,384 words

struct HyPair {
    int64_t rows_off;   // row symbols (codes into the pair's table) in the host's row list
    int64_t tab_off;    // the pair's mask table ((sigma + 1) rows of ks words) in the table pool
    int64_t cols_off;   // column symbols (uint8 codes) in the column pool
    int64_t roff_off;   // row 0's entry in the row-code pool
    int32_t m, n, K, sigma;
    int32_t out;        // caller index
    int32_t ks;         // table row stride in words: K rounded up to whole segments
};

// Row-code pool of a pair: entry i (i in [-kRowPadLo, m +
// kRowPadHi)) is row i's symbol code, and rows outside [0, m) hold sigma,
// the table's all-zero row.
constexpr int kRowPadLo = 64, kRowPadHi = 256;

// Mask table of a pair: row s < sigma, word w < K holds bit j = [b[64 w + j]
// == s]; row sigma and the words past K are zero.
__global__ void hy_table(const HyPair *__restrict__ pairs, const uint8_t *__restrict__ cols,
                         uint64_t *__restrict__ tab) {
    const HyPair P = pairs[blockIdx.x];
    const uint8_t *b = cols + P.cols_off;
    const int64_t total = (int64_t)(P.sigma + 1) * P.ks;
    for (int64_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
        const int s = (int)(idx / P.ks), w = (int)(idx % P.ks);
        uint64_t mask = 0;
        if (s < P.sigma && w < P.K) {
            const int j0 = 64 * w, j1 = min(P.n, j0 + 64);
            for (int j = j0; j < j1; ++j) mask |= (uint64_t)(b[j] == s ? 1 : 0) << (j - j0);
        }
        tab[P.tab_off + idx] = mask;
    }
}

// The systolic kernel.  A pair's packed string is cut into segments of
// 64 NW R words, each a workgroup of NW waves with the global wave index
// Wg = NW seg + wave (so lanes, rows and the one-group skew continue across
// segments); pairs up to 2,048 words are one segment.  Wave Wg runs step t =
// 64 (g - Wg) + j as step j of its group g, so everything it reads from the
// wave before it was written in groups g - 1 and g - 2, and one barrier per
// group suffices.  The carry into a segment's lane 0 is the previous
// segment's lane-63 output: its last wave publishes, per 64-step group, the
// 64 carry bits as two tagged 64-bit entries (high word: (epoch << 16) |
// group, low word: 32 carry bits), written and read with agent-scope atomics
// like the combing kernel's strand hand-off, so a reader that sees the tag
// has the bits.  Workgroups are persistent and take (pair, segment) tickets
// in order: a segment waits only for the segment before it, whose ticket a
// running workgroup already holds.  The wait is bounded; running out sets
// ctl->error and the run reports a timeout.
//
// The step is branch-free and has no range tests.  The carries between
// lanes live in one 64-bit lane mask: bit L of `cm` is lane L's carry-in, so
// a lane's R-word addition is a chain of v_addc_co_u32 that takes its first
// carry from `cm` and leaves its carry-out mask in a scalar register pair;
// the next step's mask is that one shifted up by a lane, with lane 0's
// carry-in from the neighbour's bits (scalar ops, no DPP).  Rows outside
// [0, m) meet the table's all-zero row through the padded row-code pool, and
// words past K the table's zero padding, so their p is 0 and their carries
// stay 0.  A wave runs only the groups in which it meets rows of [0, m); in
// the others its carry bits are 0.
struct SegItem {
    int32_t pair, seg;
};

constexpr int kSegWaves = 4;   // waves per segment of the segmented (widest) classes

__device__ __forceinline__ uint64_t seg_ld(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void seg_st(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A lane holds H 32-bit words of the packed string (H = 1, 2 or 4: 32, 64
// or 128 bits, word 0 lowest).  s = v + p + carry-in over the lane's words,
// the carry-in of every lane taken from the lane mask `cin`; returns the
// carry-out mask.  One v_addc_co_u32 per word, the carry passed in scalar
// registers.
template <int H>
__device__ __forceinline__ uint64_t add_words(const uint32_t (&v)[H], const uint32_t (&p)[H], uint64_t cin,
                                              uint32_t (&s)[H]);
template <>
__device__ __forceinline__ uint64_t add_words<1>(const uint32_t (&v)[1], const uint32_t (&p)[1], uint64_t cin,
                                                 uint32_t (&s)[1]) {
    uint64_t co;
    asm("v_addc_co_u32_e64 %0, %1, %2, %3, %4" : "=v"(s[0]), "=s"(co) : "v"(v[0]), "v"(p[0]), "s"(cin));
    return co;
}
template <>
__device__ __forceinline__ uint64_t add_words<2>(const uint32_t (&v)[2], const uint32_t (&p)[2], uint64_t cin,
                                                 uint32_t (&s)[2]) {
    uint64_t c1, co;
    asm("v_addc_co_u32_e64 %0, %2, %4, %5, %8\n\t"
        "v_addc_co_u32_e64 %1, %3, %6, %7, %2"
        : "=&v"(s[0]), "=v"(s[1]), "=&s"(c1), "=s"(co)
        : "v"(v[0]), "v"(p[0]), "v"(v[1]), "v"(p[1]), "s"(cin));
    return co;
}
template <>
__device__ __forceinline__ uint64_t add_words<4>(const uint32_t (&v)[4], const uint32_t (&p)[4], uint64_t cin,
                                                 uint32_t (&s)[4]) {
    uint64_t c1, c2, c3, co;
    asm("v_addc_co_u32_e64 %0, %4, %8, %9, %16\n\t"
        "v_addc_co_u32_e64 %1, %5, %10, %11, %4\n\t"
        "v_addc_co_u32_e64 %2, %6, %12, %13, %5\n\t"
        "v_addc_co_u32_e64 %3, %7, %14, %15, %6"
        : "=&v"(s[0]), "=&v"(s[1]), "=&v"(s[2]), "=v"(s[3]), "=&s"(c1), "=&s"(c2), "=&s"(c3), "=s"(co)
        : "v"(v[0]), "v"(p[0]), "v"(v[1]), "v"(p[1]), "v"(v[2]), "v"(p[2]), "v"(v[3]), "v"(p[3]), "s"(cin));
    return co;
}

// A wave-uniform 64-bit value in scalar registers (free where it is there
// already): the compiler cannot tell that a value carried through the
// group loop, or an inline-asm scalar output, is uniform.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(x >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// V = (V ^ p) | s: one v_bitop3_b32 (truth table 0xDE: (a ^ c) | b)
__device__ __forceinline__ uint32_t hy_update(uint32_t v, uint32_t s, uint32_t p) {
    return __builtin_amdgcn_bitop3_b32(v, s, p, 0xDE);
}

// The H mask words at `q` (global or LDS): one load of 4 H aligned bytes.
template <int H, typename Ptr>
__device__ __forceinline__ void load_masks(Ptr q, uint32_t (&mq)[H]) {
    if (H == 1) {
        mq[0] = *(const uint32_t *)q;
    } else if (H == 2) {
        const uint2 x = *(const uint2 *)q;
        mq[0] = x.x;
        mq[1] = x.y;
    } else {
        const uint4 x = *(const uint4 *)q;
        mq[0] = x.x;
        mq[1] = x.y;
        mq[2] = x.z;
        mq[3] = x.w;
    }
}

// H: 32-bit words per lane (1, 2 or 4).  DM: mask ring depth (steps a lane's
// masks are read ahead); DS: code ring depth (steps the row codes are read
// ahead of their masks).  LT: the segment's table slice ((sigma + 1) rows of
// 64 NW H words) is copied into LDS first and the masks are gathered from
// there.  A step's gather reads one
// word of a different row in every lane (the lanes are on different rows):
// from the global table that is 64 cache lines per wave and step, which the
// vector memory pipeline serves at about one line per cycle; from LDS it is
// one conflict-free read.
// s_waitcnt vmcnt(15) lgkmcnt(3) (expcnt not waited): the encoding of the
// immediate on gfx9 -- vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8
constexpr int kWaitVm15Lgkm3 = 0xF | (0x7 << 4) | (0x3 << 8);

template <int H, int DM, int DS, bool LT, int NW>
__global__ __launch_bounds__(64 * NW) void hy_seg(const HyPair *__restrict__ pairs,
                                                         const SegItem *__restrict__ items, uint32_t nitems,
                                                         const int64_t *__restrict__ xoff,
                                                         const uint32_t *__restrict__ roff,
                                                         const uint64_t *__restrict__ tab, uint64_t *__restrict__ xbuf,
                                                         uint32_t epoch, LaunchCtl *ctl, uint32_t spin_limit,
                                                         int32_t *__restrict__ llcs, int raise_prio) {
    static_assert(H == 1 || H == 2 || H == 4, "32, 64 or 128 bits per lane");
    // the widest classes hold the longest row chains: their waves issue first
    // on a SIMD they share with the other classes' waves
    if (raise_prio) __builtin_amdgcn_s_setprio(2);
    static_assert(64 % DM == 0 && 64 % DS == 0, "the rings must divide a group of 64 steps");
    static_assert(!LT || (DM == 4 && DS == 16), "the LDS form's one wait per step assumes rings of 4 and 16");
    constexpr int SW = 64 * NW * H;   // 32-bit words per segment
    __shared__ uint64_t hand[3][NW];
    extern __shared__ uint32_t stab[];   // LT: the segment's table slice, row s at s * SW
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
        const int m = P.m;
        const int K32 = 2 * P.K;   // 32-bit words of the packed string
        const int nseg = (K32 + SW - 1) / SW;
        const int Wg = NW * seg + wave;
        const int L = 64 * Wg + lane;   // global lane of the pair
        const int w0 = L * H;           // the lane's first 32-bit word
        // the table as 32-bit words: row stride 2 ks
        const uint32_t *tab32 = (const uint32_t *)(tab + P.tab_off);
        const uint32_t ks2 = 2u * (uint32_t)P.ks;
        if (LT) {
            // the slice: rows 0 .. sigma (sigma: zeros), the segment's words
            const uint32_t *src = tab32 + (int64_t)seg * SW;
            for (int idx = threadIdx.x; idx < (P.sigma + 1) * SW; idx += 64 * NW)
                stab[idx] = src[(int64_t)(idx / SW) * ks2 + idx % SW];
            __syncthreads();
        }
        // the lane's masks of row code c: global table word c * 2 ks + w0, or
        // LDS slice word c * SW + (64 wave + lane) H
        const uint32_t *tb = tab32 + w0;
        const uint32_t *sl = stab + (64 * wave + lane) * H;
        auto fetch = [&](uint32_t code, uint32_t (&mq)[H]) {
            if (LT) load_masks<H>(sl + code * SW, mq);
            else load_masks<H>(tb + code * ks2, mq);
        };
        // rl[t]: code of the row this lane meets at step t (row t - L)
        const uint32_t *rl = roff + P.roff_off - L;
        // groups this workgroup runs: wave Wg meets rows in groups [2 Wg, 2 Wg + c]
        const int c = (m + 63) / 64 + 1;
        const int G0 = 2 * NW * seg, G1 = 2 * NW * seg + 2 * (NW - 1) + c;
        const int ng = G1 - G0 + 1;   // the same for every segment of the pair
        // the groups in which this wave meets rows of [0, m): lane 0 meets row
        // 0 in group 2 Wg, lane 63 row m - 1 by group 2 Wg + (m + 62) / 64
        const int ga = 2 * Wg, gb = 2 * Wg + (m + 62) / 64;
        uint64_t *xmine = xbuf + xoff[it.pair] + (int64_t)seg * ng * 2;
        const uint64_t *xprevseg = xbuf + xoff[it.pair] + (int64_t)(seg - 1) * ng * 2;
        uint32_t V[H];
#pragma unroll
        for (int r = 0; r < H; ++r) V[r] = ~0u;
        // The previous segment's carry word of group gp sits in two tagged
        // entries (none outside the groups that segment ran: rows there lie
        // outside [0, m) and the word is 0).  Wave 0 issues the loads of
        // group g's word half-way through its own group g and checks them
        // at the start of group g + 1, so the wait is off the step chain
        // and the ring's loads are not drained; a word that is not there
        // yet is re-read with a bounded wait.
        const bool ext = wave == 0 && seg > 0;
        auto seg_entry = [&](int gp) -> const uint64_t * {
            if (!ext || gp < G0 - 2 * NW || gp > G1 - 2 * NW) return nullptr;
            return xprevseg + (int64_t)(gp - (G0 - 2 * NW)) * 2;
        };
        auto seg_settle = [&](const uint64_t *e, int gp, uint64_t lo, uint64_t hi) -> uint64_t {
            if (!e) return 0ull;
            const uint32_t want = (epoch << 16) | ((uint32_t)gp & 0xFFFFu);
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
        // a valid address for the loads of waves and groups without a word
        const uint64_t *const dummy = xbuf;
        uint64_t xprev = 0;
        const uint64_t *pf_e = nullptr;   // the word in flight: entry and values
        uint64_t pf_lo = 0, pf_hi = 0;
        if (ext) {
            const uint64_t *e = seg_entry(ga - 2);
            xprev = uniform64(seg_settle(e, ga - 2, seg_ld(e ? e : dummy), seg_ld(e ? e + 1 : dummy)));
            pf_e = seg_entry(ga - 1);
            pf_lo = seg_ld(pf_e ? pf_e : dummy);
            pf_hi = seg_ld(pf_e ? pf_e + 1 : dummy);
        }
        // At step j: Mq[j mod DM] holds the step's masks and So[j mod DS]
        // the code of step j + DM.  Primed after the hand-off word's loads
        // above, so that a wait for those (at the first group) leaves the
        // rings' loads in flight.
        uint32_t Mq[DM][H];
        uint32_t So[DS];
        {
            const int ta = 64 * (ga - Wg);   // the wave's first step with a row of [0, m)
#pragma unroll
            for (int d = 0; d < DM; ++d) fetch(rl[ta + d], Mq[d]);
#pragma unroll
            for (int e = 0; e < DS; ++e) So[e] = rl[ta + DM + e];
        }
        uint64_t co = 0;   // carry-out mask of the wave's previous step
        for (int g = G0; g <= G1; ++g) {
            uint64_t ob = 0;   // lane 63's carry-outs, bit j at step j
            if (g >= ga && g <= gb) {
                // lane j of vob: the high half of step j's carry-out mask
                // (bit 31: lane 63's carry), one v_writelane_b32 per step
                uint32_t vob = 0;
                // bit j: the carry into lane 0 at step j of this group
                uint64_t in = 0;
                if (wave > 0) {
                    const uint64_t p1 = g - 1 >= G0 ? hand[(g - 1) % 3][wave - 1] : 0ull;
                    const uint64_t p2 = g - 2 >= G0 ? hand[(g - 2) % 3][wave - 1] : 0ull;
                    in = (p1 << 1) | (p2 >> 63);
                } else if (ext) {
                    const uint64_t x1 = uniform64(seg_settle(pf_e, g - 1, pf_lo, pf_hi));
                    in = (x1 << 1) | (xprev >> 63);
                    xprev = x1;
                }
                in = uniform64(in);
                const uint32_t *rg = rl + 64 * (g - Wg) + DM + DS;
                uint64_t cm = uniform64((co << 1) | (in & 1ull));
#pragma unroll
                for (int j = 0; j < 64; ++j) {
                    const int d = j % DM, e = j % DS;
                    // one wait for both rings: the masks read DM steps ago
                    // (LDS) and the code read DS steps ago (global); each step
                    // issues one read of each
                    if (LT) __builtin_amdgcn_s_waitcnt(kWaitVm15Lgkm3);
                    uint32_t p[H], sum[H];
#pragma unroll
                    for (int r = 0; r < H; ++r) p[r] = Mq[d][r] & V[r];
                    co = uniform64(add_words<H>(V, p, cm, sum));
#pragma unroll
                    for (int r = 0; r < H; ++r) V[r] = hy_update(V[r], sum[r], p[r]);
                    asm("v_writelane_b32 %0, %1, %2" : "+v"(vob) : "s"((uint32_t)(co >> 32)), "i"(j));
                    if (j < 63) cm = (co << 1) | ((in >> (j + 1)) & 1ull);
                    // refill: masks for step j + DM, the code for step j + DM + DS
                    fetch(So[e], Mq[d]);
                    So[e] = rg[j];
                    if (j == 32) {   // the previous segment's word of this group
                        pf_e = seg_entry(g);
                        pf_lo = seg_ld(pf_e ? pf_e : dummy);
                        pf_hi = seg_ld(pf_e ? pf_e + 1 : dummy);
                    }
                    // keeps each step's loads in its step, ahead of their use
                    __builtin_amdgcn_sched_barrier(0);
                }
                ob = __builtin_amdgcn_ballot_w64((int32_t)vob < 0);
            } else {
                co = 0;
            }
            if (lane == 0) {
                hand[g % 3][wave] = ob;
                if (wave == NW - 1 && seg + 1 < nseg) {
                    const uint64_t tag = (uint64_t)((epoch << 16) | ((uint32_t)g & 0xFFFFu)) << 32;
                    uint64_t *e = xmine + (int64_t)(g - G0) * 2;
                    seg_st(e, tag | (ob & 0xFFFFFFFFull));
                    seg_st(e + 1, tag | (ob >> 32));
                }
            }
            // the hand-off words are in LDS: a barrier that waits for LDS
            // only, so the ring's global loads stay in flight across it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
            __builtin_amdgcn_s_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        }
        int z = 0;
#pragma unroll
        for (int r = 0; r < H; ++r) z += w0 + r < K32 ? 32 - __popc(V[r]) : 0;
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
    std::vector<int> cls_sigma;                   // segmented classes: the largest alphabet
    hy::HyPair *d_pairs = nullptr;
    // segmented classes: (pair, segment) tickets for each segment width
    // (1 or 2 words per lane), exchange buffer (sized for the narrower
    // segments, which are more), control
    std::vector<std::vector<hy::SegItem>> seg_items;   // [kWidths class + width_index(H)]
    hy::SegItem *d_items = nullptr;
    std::vector<int64_t> item_off;
    int64_t *d_xoff = nullptr;          // per pair: its segments' exchange entries in d_xbuf
    uint64_t *d_xbuf = nullptr;
    int64_t xbuf_entries = 0;
    uint64_t cleared_at = 0;            // run serial when d_xbuf was last cleared
    LaunchCtl *d_ctl = nullptr;         // one per class
    // a segment waits a few groups (well under a millisecond) for its
    // neighbour; 2^20 polls is over a second of margin and fails fast
    uint32_t spin_limit = 1u << 20;
    uint32_t *d_roff = nullptr;         // row-code pool of the segmented pairs
    uint8_t *d_cols = nullptr;
    uint64_t *d_tab = nullptr;
    int32_t *d_llcs = nullptr;
    int64_t tab_words = 0, word_rows = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
    // launch classes are independent: the segmented (long-chain) classes run
    // on one stream, the others one after another on a second, both forked
    // from and joined back into `stream` (default priorities)
    hipStream_t lane_stream[2] = {nullptr, nullptr};   // [0]: segmented classes, [1]: the others
    hipEvent_t lane_done[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr;
};

namespace {

// Launch classes (length buckets) by packed width K (64-bit words): up to
// 64, 128, 256, 2,048 and 16,384 words (2^20 columns).  Every class runs the
// systolic kernel with H 32-bit words per lane (the reference's analogue is
// Hyyro's word type), one H per class: the default (H = 2, 64 bits) or the
// one sl_hyyro_autotune measured fastest on the batch's own pairs.  Waves per
// segment: the class's width over 64 H, so the classes up to 2,048 words are
// one segment per pair (1, 2 and 4 waves at H = 2 up to 64, 128 and 256
// words); wider pairs are cut into segments of 4 waves, so a long pair
// spreads over as many compute units as it has segments.  H = 1 (half the
// instructions per step, twice the lanes) is offered where it needs at most
// 4 waves per pair.
int hy_class(int K) {
    static const int limits[] = {64, 128, 256, 2048, 16384};
    int c = 0;
    while (K > limits[c]) ++c;
    return c;
}
constexpr int kClasses = 5;
constexpr int kWidths = 3;   // H = 1, 2, 4 32-bit words per lane
int width_index(int H) { return H == 1 ? 0 : (H == 2 ? 1 : 2); }
constexpr size_t kSegLdsBytes = 64 * 1024;   // largest table slice the kernel holds in LDS
constexpr int kClassMaxWords[kClasses] = {64, 128, 256, 2048, 16384};
constexpr int kFirstSegClass = 3;
constexpr int kSegClasses = 2;   // classes 3 and 4: several segments per pair
bool is_seg_class(int cls) { return cls >= kFirstSegClass && cls < kFirstSegClass + kSegClasses; }
// 32-bit words per lane each class offers, default first
const std::vector<int> &class_wpl_options(int cls) {
    static const std::vector<int> wide = {2, 1, 4}, no_half = {2, 4};
    return cls == 2 ? no_half : wide;
}
bool class_offers(int cls, int H) {
    for (int r : class_wpl_options(cls))
        if (r == H) return true;
    return false;
}
// waves per segment
int class_waves(int cls, int H) {
    if (is_seg_class(cls)) return hy::kSegWaves;
    return std::max(1, 2 * kClassMaxWords[cls] / (64 * H));
}
// 64-bit words per segment
int class_segment_words(int cls, int H) { return 64 * class_waves(cls, H) * H / 2; }

void hy_free(Hyyro *h) {
    if (!h) return;
    (void)hipFree(h->d_pairs);
    (void)hipFree(h->d_items);
    (void)hipFree(h->d_xoff);
    (void)hipFree(h->d_xbuf);
    (void)hipFree(h->d_ctl);
    (void)hipFree(h->d_roff);
    (void)hipFree(h->d_cols);
    (void)hipFree(h->d_tab);
    (void)hipFree(h->d_llcs);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : h->lane_done)
        if (e) (void)hipEventDestroy(e);
    if (h->fork) (void)hipEventDestroy(h->fork);
    for (auto &cs : h->lane_stream)
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
    // every wave adds its words' zero count: clear the scores, the tickets
    // and the error flags first
    SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_llcs, 0u, (uint64_t)std::max(h->npairs, 1), h->stream));
    SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_ctl, 0u, 2u * kClasses, h->stream));
    uint32_t epoch = 0;
    if (next_run_tag(h->cleared_at, &epoch) && h->xbuf_entries > 0)
        SL_HIP_TRY(launch_fill_u32((uint32_t *)h->d_xbuf, 0u, 2u * (uint64_t)h->xbuf_entries, h->stream));
    SL_HIP_TRY(hipEventRecord(h->fork, h->stream));
    for (hipStream_t ls : h->lane_stream) SL_HIP_TRY(hipStreamWaitEvent(ls, h->fork, 0));
    const int cus = device_cus();
    // widest class first: it holds the longest pairs.  Every launch is
    // enqueued before `stream` waits for any of them: streams can share a
    // hardware queue, and a launch behind such a wait would start only
    // when the waited-for lane ends.
    for (int c = kClasses - 1; c >= 0; --c) {
        if (!((class_mask >> c) & 1u)) continue;
        hipStream_t cs = h->lane_stream[is_seg_class(c) ? 0 : 1];
        const int H = h->wpl[(size_t)c];
        const size_t v = (size_t)(kWidths * c + width_index(H));
        const uint32_t nit = (uint32_t)h->seg_items[v].size();
        if (!nit) continue;
        const hy::SegItem *items = h->d_items + h->item_off[v];
        // one workgroup of NW waves per segment of 64 NW H 32-bit words;
        // enough persistent workgroups for about 8 waves per compute unit
        const int NW = class_waves(c, H);
        const unsigned blocks = (unsigned)std::min<int64_t>(nit, (int64_t)cus * 8 / NW);
        // the table slice in LDS when it fits in 64 KiB (segmented classes:
        // alphabets up to 63, 31 and 15 symbols at 32, 64 and 128 bits per lane)
        const size_t slice = (size_t)(h->cls_sigma[(size_t)c] + 1) * 64 * NW * H * 4;
        const bool lds = slice <= kSegLdsBytes;
#define SL_HY_SEG(H, DM, LT, NW)                                                                                  \
    hipLaunchKernelGGL((hy::hy_seg<H, DM, 16, LT, NW>), dim3(blocks), dim3(64 * NW), LT ? slice : 0, cs, h->d_pairs, \
                       items, nit, h->d_xoff, h->d_roff, h->d_tab, h->d_xbuf, epoch, h->d_ctl + c, h->spin_limit,   \
                       h->d_llcs, is_seg_class(c) ? 1 : 0)
#define SL_HY_SEG_NW(H, NW)                  \
    do {                                     \
        if (lds) SL_HY_SEG(H, 4, true, NW);  \
        else SL_HY_SEG(H, 16, false, NW);    \
    } while (0)
#define SL_HY_SEG_H(H)                      \
    do {                                    \
        if (NW == 1) SL_HY_SEG_NW(H, 1);    \
        else if (NW == 2) SL_HY_SEG_NW(H, 2); \
        else SL_HY_SEG_NW(H, 4);            \
    } while (0)
        if (H == 1) SL_HY_SEG_H(1);
        else if (H == 2) SL_HY_SEG_H(2);
        else SL_HY_SEG_H(4);
#undef SL_HY_SEG_H
#undef SL_HY_SEG_NW
#undef SL_HY_SEG
        SL_HIP_TRY(hipGetLastError());
    }
    for (int k = 0; k < 2; ++k) {
        SL_HIP_TRY(hipEventRecord(h->lane_done[k], h->lane_stream[k]));
        SL_HIP_TRY(hipStreamWaitEvent(h->stream, h->lane_done[k], 0));
    }
    return kOk;
}

// After a run (stream synchronised): a segment's bounded wait ran out?
int hy_check(Hyyro *h) {
    LaunchCtl ctl[kClasses];
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
    h->cls_sigma.assign(kClasses, 0);
    std::vector<int32_t> rows;
    std::vector<uint32_t> roff;
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
        // table rows padded to whole segments of the class's widest form, so
        // a segment's lanes past K read zero words, and the rows' codes
        // padded on both sides with sigma (the zero row)
        const int cls = hy_class(P.K);
        int segw = 1;
        for (int H : class_wpl_options(cls)) segw = std::max(segw, class_segment_words(cls, H));
        P.ks = (P.K + segw - 1) / segw * segw;
        h->tab_words = (h->tab_words + 1) & ~(int64_t)1;   // 16-byte aligned tables
        P.tab_off = h->tab_words;
        h->tab_words += (int64_t)(sigma + 1) * P.ks;
        roff.insert(roff.end(), hy::kRowPadLo, (uint32_t)sigma);
        P.roff_off = (int64_t)roff.size();
        roff.insert(roff.end(), rows.begin() + P.rows_off, rows.begin() + P.rows_off + P.m);
        roff.insert(roff.end(), hy::kRowPadHi, (uint32_t)sigma);
        h->cls_sigma[(size_t)cls] = std::max(h->cls_sigma[(size_t)cls], sigma);
        h->word_rows += (int64_t)P.m * P.K;
        h->pairs.push_back(P);
        h->cls_list[(size_t)hy_class(P.K)].push_back(i);
    }
    // ticket-kernel items of every class and width: pairs by work (rows x
    // words) descending, each pair's segments in order; exchange entries
    // (segmented classes): per segment, 2 per group it runs
    std::vector<hy::SegItem> all_items;
    std::vector<int64_t> xoff((size_t)std::max(npairs, 1), 0);
    int64_t xtotal = 0;
    h->seg_items.assign(kWidths * kClasses, {});
    h->item_off.assign(kWidths * kClasses, 0);
    for (int c = 0; c < kClasses; ++c) {
        auto lst = h->cls_list[(size_t)c];
        std::stable_sort(lst.begin(), lst.end(), [&](int x, int y) {
            return (int64_t)h->pairs[(size_t)x].m * h->pairs[(size_t)x].K >
                   (int64_t)h->pairs[(size_t)y].m * h->pairs[(size_t)y].K;
        });
        if (is_seg_class(c))
            for (int pi : lst) {   // exchange entries for the narrowest (H = 1) segments, which are most
                const hy::HyPair &P = h->pairs[(size_t)pi];
                const int sw = class_segment_words(c, 1);
                const int nseg = (P.K + sw - 1) / sw;
                const int64_t ng = (P.m + 63) / 64 + 8;   // groups a segment runs (hy_seg: G1 - G0 + 1)
                xoff[(size_t)pi] = xtotal;
                xtotal += (int64_t)nseg * ng * 2;
            }
        for (int H : class_wpl_options(c)) {
            const int segw = class_segment_words(c, H);
            const size_t v = (size_t)(kWidths * c + width_index(H));
            auto &items = h->seg_items[v];
            for (int pi : lst) {
                const int nseg = (h->pairs[(size_t)pi].K + segw - 1) / segw;
                for (int sg = 0; sg < nseg; ++sg) items.push_back(hy::SegItem{pi, sg});
            }
            h->item_off[v] = (int64_t)all_items.size();
            all_items.insert(all_items.end(), items.begin(), items.end());
        }
    }
    h->wpl.assign(kClasses, 0);
    for (int c = 0; c < kClasses; ++c) h->wpl[(size_t)c] = class_wpl_options(c)[0];
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (auto &ev : h->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->fork, hipEventDisableTiming);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&h->lane_stream[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&h->lane_done[k], hipEventDisableTiming);
    }
    auto alloc = [&](void **p, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(p, std::max<size_t>(bytes, 8));
    };
    alloc((void **)&h->d_pairs, sizeof(hy::HyPair) * h->pairs.size());
    alloc((void **)&h->d_items, sizeof(hy::SegItem) * all_items.size());
    alloc((void **)&h->d_xoff, sizeof(int64_t) * xoff.size());
    alloc((void **)&h->d_xbuf, sizeof(uint64_t) * (size_t)xtotal);
    alloc((void **)&h->d_ctl, sizeof(LaunchCtl) * kClasses);
    alloc((void **)&h->d_roff, sizeof(uint32_t) * roff.size());
    alloc((void **)&h->d_cols, cols.size());
    alloc((void **)&h->d_tab, sizeof(uint64_t) * (size_t)h->tab_words);
    alloc((void **)&h->d_llcs, sizeof(int32_t) * (size_t)std::max(npairs, 1));
    if (e == hipSuccess && !h->pairs.empty())
        e = hipMemcpy(h->d_pairs, h->pairs.data(), sizeof(hy::HyPair) * h->pairs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !all_items.empty())
        e = hipMemcpy(h->d_items, all_items.data(), sizeof(hy::SegItem) * all_items.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_xoff, xoff.data(), sizeof(int64_t) * xoff.size(), hipMemcpyHostToDevice);
    // exchange entries start with tag 0, which no run's (nonzero) epoch matches
    if (e == hipSuccess && xtotal > 0) e = launch_fill_u32((uint32_t *)h->d_xbuf, 0u, 2u * (uint64_t)xtotal, h->stream);
    h->xbuf_entries = xtotal;
    h->cleared_at = run_serial_now();
    if (e == hipSuccess && !roff.empty())
        e = hipMemcpy(h->d_roff, roff.data(), sizeof(uint32_t) * roff.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !cols.empty()) e = hipMemcpy(h->d_cols, cols.data(), cols.size(), hipMemcpyHostToDevice);
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
