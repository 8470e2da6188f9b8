// bitpar.hip -- kernel (b): bit-parallel sticky-braid combing of two binary
// strings, returning the LLCS score.
//
// Restates prefix_lcs_via_semi_local::binary::llcs_2symbol_smart_combing
// (reference prefix_lcs/bitwise/transposition_network_binary_alphabet.h:598-663)
// with the "formula 2" cell routine (:252-323).  Strands are single bits: a
// word of l holds 32 horizontal strands (1 = entered at the left edge), a word
// of t holds 32 vertical strands (0 = entered at the top).  After combing,
// LLCS = |a| - popcount(all l words)  (:648-662).
//
// Word size: the reference packs into uint64_t (a reversed and MSB-first, b
// LSB-first, encoders_and_decoders.h:50-139).  Read as 32-bit little-endian
// words the same buffers are exactly the 32-bit packing of the same strings,
// so the engine combs 32x32 word-cells with native 32-bit VALU operations;
// the combing (and the score) does not depend on the word size.
//
// Mapping: lane k of a wave owns one word-row of a (32 rows of the grid); word
// columns of b stream through the lanes exactly as in kernel (a): lane 0 takes
// the next column's t word, each lane hands its t word to lane k+1 with one
// DPP move per step, and column symbols come from a wave-private LDS ring.
// Passes of 64 word-rows hand their bottom t words to the next pass as tagged
// entries (see comb.hip); here the hand-off is flushed every 8 columns because
// one step (one 32x32 cell, ~630 VALU ops) is long: the lag between passes is
// ~72 steps, close to the 64-lane pipeline depth.
#include "common.h"

namespace slhip {
namespace {

constexpr int kDppShr1 = 0x138;                  // lane k <- lane k-1 (lane 0 keeps `old`)
constexpr int kRing = 128;                       // column ring (power of two)
constexpr int kGroup = 8;                        // columns per load/flush group
constexpr int kBitWaveLds = 3 * kRing;           // b ring, injected-t ring, outgoing-t ring
constexpr int kBitWavesPerBlock = 4;

__device__ __forceinline__ uint32_t shr1(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppShr1, 0xf, 0xf, false);
}
__device__ __forceinline__ uint64_t ld_co(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One 32x32 word-cell, formula 2 (transposition_network_binary_alphabet.h:
// 252-323, for 32-bit words): 31 upper anti-diagonals, the centre, 31 lower
// ones.  `a` is the a-word (reversed string, MSB-first), `b` the b-word.
__device__ __forceinline__ void bit_cell(uint32_t &l, uint32_t &t, uint32_t a, uint32_t b) {
    uint32_t mask = 1u;
#pragma unroll
    for (int sh = 31; sh > 0; --sh) {
        const uint32_t lc = l >> sh;
        const uint32_t tc = t << sh;
        const uint32_t cond = ~((a >> sh) ^ b);
        t = (lc | ~mask) & (t | (cond & mask));
        l = tc ^ (t << sh) ^ l;
        mask = (mask << 1) | 1u;
    }
    {
        const uint32_t cond = ~(a ^ b);
        const uint32_t lc = l, tc = t;
        t = (lc | ~mask) & (t | (cond & mask));
        l = tc ^ t ^ l;
    }
    mask = 0xFFFFFFFFu;
#pragma unroll
    for (int sh = 1; sh < 32; ++sh) {
        mask <<= 1;
        const uint32_t lc = l << sh;
        const uint32_t tc = t >> sh;
        const uint32_t cond = ~((a << sh) ^ b);
        t = (lc | ~mask) & (t | (cond & mask));
        l = tc ^ (t >> sh) ^ l;
    }
}

}  // namespace

// Descriptor of one bit-parallel problem: M word-rows (a), N word-columns (b).
struct BitDesc {
    const uint32_t *a_rev;   // M words: word j = a_rev word j (32-bit view)
    const uint32_t *b;       // N words
    uint64_t *buf;           // N tagged entries: t word of each column between passes
    uint32_t *dis;           // popcount accumulator (dis-braid)
    int32_t M, N;
    int32_t npasses;         // ceil(M / 64)
    uint32_t epoch;
};

// One pass = 64 consecutive word-rows (grid order: row g uses a_rev word
// M-1-g, its left strands start as all ones, :619-620).  Top strands of pass
// 0 start as zeros (:621-622).
// Returns this wave's dis-braid count for the pass (uniform across the wave).
__device__ uint32_t bit_pass(const BitDesc &d, int pass, LaunchCtl *ctl, uint32_t spin_limit, uint32_t *W) {
    const int lane = threadIdx.x & 63;
    const uint32_t M = (uint32_t)d.M, N = (uint32_t)d.N;
    const int64_t g = (int64_t)pass * 64 + lane;
    const bool row_ok = g < (int64_t)M;
    const uint32_t a = row_ok ? d.a_rev[M - 1 - (uint32_t)g] : 0u;
    uint32_t l = row_ok ? 0xFFFFFFFFu : 0u;
    const bool first = pass == 0, last = pass == d.npasses - 1;
    const uint32_t want_tag = (d.epoch << 16) | ((uint32_t)pass & 0xFFFFu);          // written by pass-1
    const uint64_t my_tag = (uint64_t)((d.epoch << 16) | (((uint32_t)pass + 1u) & 0xFFFFu)) << 32;
    uint32_t *ringB = W, *ringT = W + kRing, *ringO = W + 2 * kRing;

    const uint32_t steps = N + 63u;
    const uint32_t ngroups = (steps + kGroup - 1) / kGroup;
    // lanes 0..7 carry the group's columns; values for absent columns never matter
    auto issue = [&](uint32_t grp, uint64_t &tv, uint32_t &bw) {
        const uint32_t col = grp * kGroup + (uint32_t)lane;
        tv = 0;
        bw = 0;
        if (lane < kGroup && col < N) {
            tv = first ? 0ull : ld_co(&d.buf[col]);
            bw = d.b[col];
        }
    };
    auto settle = [&](uint32_t grp, uint64_t &tv) -> bool {
        if (first) return true;
        const uint32_t col = grp * kGroup + (uint32_t)lane;
        uint32_t spins = 0;
        for (;;) {
            const bool ready = lane >= kGroup || col >= N || (uint32_t)(tv >> 32) == want_tag;
            if (__builtin_amdgcn_ballot_w64(!ready) == 0) return true;
            if (++spins >= spin_limit ||
                __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
            if (!ready) tv = ld_co(&d.buf[col]);
        }
    };

    uint64_t tv = 0;
    uint32_t bw = 0;
    issue(0, tv, bw);
    uint32_t t = 0;
    bool ok = true;
    for (uint32_t grp = 0; grp < ngroups; ++grp) {
        if (!settle(grp, tv)) {   // timed out (reported): leave the loop, no early return
            ok = false;
            break;
        }
        if (lane < kGroup) {
            const uint32_t slot = (grp * kGroup + (uint32_t)lane) & (kRing - 1);
            ringB[slot] = bw;
            ringT[slot] = (uint32_t)tv;
        }
        lds_sync();
        if (grp + 1 < ngroups) issue(grp + 1, tv, bw);
        for (int u = 0; u < kGroup; ++u) {
            const uint32_t s = grp * kGroup + (uint32_t)u;
            const int64_t col = (int64_t)s - lane;                  // column combed by this lane now
            const bool col_ok = col >= 0 && col < (int64_t)N;
            const uint32_t bword = ringB[(uint32_t)col & (kRing - 1)];
            t = shr1(ringT[s & (kRing - 1)], t);                    // lane 0 takes column s
            uint32_t nl = l, nt = t;
            bit_cell(nl, nt, a, bword);
            if (col_ok) l = nl;          // lanes outside the grid keep their strands
            if (row_ok) t = nt;          // padding rows pass the column through
            ringO[(uint32_t)col & (kRing - 1)] = t;
        }
        lds_sync();
        // lane 63 finished columns s-63 for the group's steps: flush them
        const int64_t fcol = (int64_t)grp * kGroup - 63 + lane;
        if (lane < kGroup && fcol >= 0 && fcol < (int64_t)N && !last)
            st_co(&d.buf[fcol], my_tag | ringO[(uint32_t)fcol & (kRing - 1)]);
        lds_sync();
    }
    // dis-braid: strands that entered at the left and exit at the right (:648-659)
    uint32_t cnt = (ok && row_ok) ? (uint32_t)__builtin_popcount(l) : 0u;
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
    return __builtin_amdgcn_readfirstlane(cnt);
}

__global__ __launch_bounds__(256) void bitpar_kernel(BitDesc d, LaunchCtl *ctl, uint32_t spin_limit) {
    __shared__ uint32_t lds[kBitWavesPerBlock * kBitWaveLds];
    const int lane = threadIdx.x & 63;
    uint32_t *W = lds + (threadIdx.x >> 6) * kBitWaveLds;
    // The wave's passes add their dis-braid counts here; one atomic after the
    // queue is drained keeps lane-0-only work out of the ticket loop, whose
    // readfirstlane needs the whole wave converged at the top of every trip.
    uint32_t dis = 0;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&ctl->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane(tk);
        if (tk >= (uint32_t)d.npasses) break;   // passes are handed out in order
        dis += bit_pass(d, (int)tk, ctl, spin_limit, W);
    }
    if (lane == 0 && dis != 0) atomicAdd(d.dis, dis);
}

hipError_t launch_bitpar(const BitDesc &d, unsigned waves, LaunchCtl *ctl, uint32_t spin_limit, hipStream_t stream) {
    if (d.npasses <= 0 || waves == 0) return hipSuccess;
    const unsigned blocks = (waves + kBitWavesPerBlock - 1) / kBitWavesPerBlock;
    hipLaunchKernelGGL(bitpar_kernel, dim3(blocks), dim3(kBitWavesPerBlock * 64), 0, stream, d, ctl, spin_limit);
    return hipGetLastError();
}

}  // namespace slhip
