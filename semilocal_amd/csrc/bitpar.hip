// bitpar.hip -- kernel (b): bit-parallel sticky-braid combing of two binary
// strings, returning the LLCS score.
//
// Computes the score of prefix_lcs_via_semi_local::binary::
// llcs_2symbol_smart_combing (reference prefix_lcs/bitwise/
// transposition_network_binary_alphabet.h:598-663): strands are single bits
// -- 1 for a strand that entered at the left edge, 0 for one that entered at
// the top (:619-622) -- and LLCS = |a| - #(left strands that leave at the
// right) (:648-662).  The per-bit cell is that of formula 2 (:252-323): with
// l, t the two strands and match = (a_i == b_j),
//     t' = l & (t | match),     l' = t | (l & ~match).
//
// Layout (the reference's packing, read as 32-bit little-endian words):
// 32-bit word M-1-g of the reversed a holds rows 32g .. 32g+31 of the grid,
// row 32g+i at bit 31-i; 32-bit word c of b holds columns 32c .. 32c+31,
// column 32c+j at bit j (encoders_and_decoders.h:50-139).
//
// Sliding bit-diagonals.  Lane k >= 1 of a wave owns one row-word: its 32
// left strands L, bit p = local row 31-p.  At step D (the lane's own diagonal
// counter) every row combs one cell: local row i meets column D-i, so bit p
// meets column D-31+p -- one aligned 32-bit operation per step.  The lane's
// column strands TS (bit p = column D-31+p) slide one bit per step: the
// column that finished its 32 rows leaves at bit 0 and enters lane k+1 at
// bit 31 (one DPP wave_shr:1, then one funnel shift, v_alignbit).  Lane k runs
// 33 steps behind lane k-1 rather than 32, so the column it takes in left
// lane k-1 two steps earlier: the DPP reads a register that is a step old and
// stays off the step's dependent chain, which is one funnel shift and one
// cell update (v_bitop3).  The column symbols come as a funnel shift of the
// lane's two current b words, pre-aligned once per 32 steps for the lane's
// skew.  Lane 0 owns no rows: it feeds lane 1.  Its strands are all 1 and
// its match 0, so its column register just slides; reloading it once per
// block with the incoming column word makes it deliver one new column per
// step, and the DPP rotates, so what slides into it meanwhile are the
// columns leaving lane 63 -- the pass's output, one word per block, with no
// per-step work.  A step is the DPP move, two funnel shifts, the match (one
// v_bitop3 of the row masks) and the two cell updates: 6 VALU ops for 32
// cells (a single wave issues about one VALU op per 6.5 cycles, so that is
// the step).  A pass's critical path is n + 64*33 steps for 63 row-words.
//
// Edges.  Columns before 0 carry strand 1 in both directions, which leaves a
// row's strand 1 -- its initial value -- unchanged whatever the symbols.
// Past column n the lanes keep combing whatever arrives (clamped b words,
// zero strands): a cell only feeds cells to its right and below, so this
// never reaches a column before n, though it does overwrite the rows'
// strands.  The score is therefore read at the bottom edge instead of the
// right one: every left strand leaves at the right or at the bottom, so
//     LLCS = |a| - #(leave right) = |a| - 32M + #(leave bottom),
// the popcount of the last pass's column words.  Rows past the 32M of the
// a words (the last pass's padding lanes) have both row masks 0: they pass
// t through.
//
// Passes of 63 row-words hand their finished column words to the next pass
// as tagged 64-bit entries (see comb.hip), one per 32 columns, read a group
// of blocks ahead by the next pass.
#include "common.h"

namespace slhip {
namespace {

constexpr int kDppRor1 = 0x13C;   // wave_ror:1: lane k <- lane k-1, lane 0 <- lane 63
constexpr int kGroup = 4;         // blocks of 32 steps per group: loads are issued one group ahead
constexpr int kSlack = 2;         // blocks a pass lets its predecessor get ahead before it starts

__device__ __forceinline__ uint32_t ror1(uint32_t src) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)src, kDppRor1, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {   // low word of (hi:lo) >> s
    return __builtin_amdgcn_alignbit(hi, lo, s);
}
__device__ __forceinline__ uint64_t ld_co(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_co(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// 32 steps (one block).  Vhi, Vlo: the lane's b bits for this block and the
// previous one (bit j of Vlo:Vhi meets bit j-e-1 of the step-e diagonal).
// The DPP rotates: lane k takes lane k-1's column and lane 0, the feeder,
// takes the column leaving lane 63.  The feeder's register is replaced at
// the block's start by `feed`, the incoming column word before this block's
// (bit j = column 32(w-1)+j: after step e its bit 0 is column 32w+e-31, the
// one lane 1 takes in two steps later); the bits it shifts in meanwhile are
// lane 63's leaving columns, so at the block's end it holds them in order.
// Returns that register as it was at the start (lane 0: the previous block's
// 32 leaving columns).
__device__ __forceinline__ uint32_t bit_block(uint32_t &L, uint32_t &tn, uint32_t &tn1, uint32_t Vhi, uint32_t Vlo,
                                              uint32_t feed, uint32_t M1, uint32_t M0, bool feeder) {
    const uint32_t collected = tn;
    if (feeder) tn = feed;
#pragma unroll
    for (int e = 0; e < 32; ++e) {
        const uint32_t x = ror1(tn1);                          // lane k-1's column from two steps back
        const uint32_t TS = funnel(x, tn, 1);                  // slide: that column enters at bit 31
        const uint32_t BS = e == 31 ? Vhi : funnel(Vhi, Vlo, (uint32_t)e + 1u);
        const uint32_t match = (BS & M1) | (~BS & M0);
        const uint32_t t2 = L & (TS | match);
        L = TS | (L & ~match);
        tn1 = tn;
        tn = t2;
    }
    return collected;
}

}  // namespace

// One pass = kBitRowLanes consecutive row-words.  The last pass returns the
// count of left strands leaving at the bottom (uniform across the wave); the
// others return 0 and hand their column words to the next pass.
//
// Memory is read one group of kGroup blocks ahead: at a group's start the
// wave checks the tags of the group's incoming entries (spinning, bounded,
// only if one is not there yet), then issues the next group's loads, and
// only the group's end waits for them -- so no wait sits inside the blocks.
__device__ uint32_t bit_pass(const BitDesc &d, int pass, LaunchCtl *ctl, uint32_t spin_limit) {
    const int lane = threadIdx.x & 63;
    const int32_t M = d.M, N = d.N;
    const bool feeder = lane == 0;
    const int32_t g = pass * kBitRowLanes + lane - 1;    // row-word of lane >= 1
    const bool row_ok = !feeder && g < M;
    const uint32_t A = row_ok ? d.a_rev[M - 1 - g] : 0u;
    const uint32_t M1 = row_ok ? A : 0u, M0 = row_ok ? ~A : 0u;
    const bool first = pass == 0, last = pass == d.npasses - 1;
    const uint32_t want_tag = (d.epoch << 16) | ((uint32_t)pass & 0xFFFFu);          // written by pass-1
    const uint64_t my_tag = (uint64_t)((d.epoch << 16) | (((uint32_t)pass + 1u) & 0xFFFFu)) << 32;
    uint32_t L = 0xFFFFFFFFu, tn = 0xFFFFFFFFu, tn1 = 0xFFFFFFFFu;
    // The lane's b bits of block w start at column 32w - 33 lane = 32 (w + cofs) + r0.
    const int32_t cofs = -lane - (lane + 31) / 32;
    const uint32_t r0 = (uint32_t)((32 - lane % 32) % 32);
    auto bword = [&](int32_t c) { return d.b[c < 0 ? 0 : (c >= N ? N - 1 : c)]; };
    // Loads of the group at block w0: this lane's b words cofs + w0 + 1 + j, and
    // (lanes j < kGroup) the incoming entry w0 + j; indices are clamped, and
    // values outside [0, N) are never used.
    auto load_group = [&](int32_t w0, uint32_t (&bw)[kGroup], uint64_t &tin) {
#pragma unroll
        for (int j = 0; j < kGroup; ++j) bw[j] = bword(cofs + w0 + 1 + j);
        const int32_t e = w0 + (lane & (kGroup - 1));
        tin = ld_co(&d.buf[e < N ? e : N - 1]);
    };
    uint32_t bcur[kGroup], bnxt[kGroup];
    uint64_t tcur, tnxt;
    load_group(0, bcur, tcur);
    uint32_t Wold = bword(cofs);                                    // b word cofs + w
    uint32_t Vprev = funnel(Wold, bword(cofs - 1), r0);            // the lane's b bits of block -1
    bool ok = true;
    uint32_t tin_prev = 0xFFFFFFFFu;   // incoming word of block -1: columns before 0 carry strand 1
    uint32_t bottom = 0;   // feeder: left strands leaving at the bottom (last pass)
    auto group = [&](int32_t w0) {
        // A pass first lets its predecessor get kSlack blocks ahead of what it
        // needs (bounded wait, sleeps growing while the predecessor is far):
        // a pass that starts right at its predecessor's heels keeps catching
        // up with it and waiting, and each wait's overshoot is passed down
        // the whole chain of passes.
        if (!first && w0 == 0) {
            const int32_t ws = N - 1 < kSlack ? N - 1 : kSlack;
            uint32_t spins = 0, nap = 1;
            while (__builtin_amdgcn_readfirstlane((uint32_t)(ld_co(&d.buf[ws]) >> 32)) != want_tag) {
                if (++spins >= spin_limit ||
                    __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    break;
                }
                for (uint32_t k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(8);
                nap = nap < 32 ? 2 * nap : nap;
            }
        }
        // incoming column words of the group's blocks (the feeder's input), tags checked
        uint32_t tin[kGroup];
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const int32_t w = w0 + j;
            uint32_t lo = __builtin_amdgcn_readlane((uint32_t)tcur, j);
            uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(tcur >> 32), j);
            if (!first && w < N && hi != want_tag && ok) {   // not there yet: bounded wait
                for (uint32_t spins = 0; hi != want_tag;) {
                    if (++spins >= spin_limit ||
                        ((spins & 255u) == 0 &&
                         __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                        __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ok = false;
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    const uint64_t v = ld_co(&d.buf[w]);
                    lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
                    hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
                }
            }
            tin[j] = (!first && w < N) ? lo : 0u;
        }
        load_group(w0 + kGroup, bnxt, tnxt);
#pragma unroll
        for (int j = 0; j < kGroup; ++j) {
            const int32_t w = w0 + j;
            const uint32_t Wnew = bcur[j];
            const uint32_t V = funnel(Wnew, Wold, r0);
            const uint32_t feed = j == 0 ? tin_prev : tin[j - 1];   // incoming word of block w-1
            const uint32_t out = bit_block(L, tn, tn1, V, Vprev, feed, M1, M0, feeder);
            Vprev = V;
            Wold = Wnew;
            // the feeder held lane 63's columns of block w-1: column word w - 67
            const int32_t wo = w - 67;
            if (feeder && wo >= 0 && wo < N) {
                if (last) bottom += (uint32_t)__builtin_popcount(out);
                else st_co(&d.buf[wo], my_tag | out);
            }
        }
        tin_prev = tin[kGroup - 1];
#pragma unroll
        for (int j = 0; j < kGroup; ++j) bcur[j] = bnxt[j];
        tcur = tnxt;
    };
    const int32_t n_all = ((N + 67 + kGroup - 1) / kGroup) * kGroup;      // through the last column word out
    for (int32_t w0 = 0; w0 < n_all && ok; w0 += kGroup) group(w0);
    return ok && last ? __builtin_amdgcn_readfirstlane(bottom) : 0u;   // lane 0 holds the count
}

__global__ __launch_bounds__(64) void bitpar_kernel(BitDesc d, LaunchCtl *ctl, uint32_t spin_limit) {
    const int lane = threadIdx.x & 63;
    // The last pass's bottom count is added here; one atomic after the queue
    // is drained keeps lane-0-only work out of the ticket loop, whose
    // readfirstlane needs the whole wave converged at the top of every trip.
    uint32_t dis = 0;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&ctl->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane(tk);
        if (tk >= (uint32_t)d.npasses) break;   // passes are handed out in order
        dis += bit_pass(d, (int)tk, ctl, spin_limit);
    }
    if (lane == 0 && dis != 0) atomicAdd(d.dis, dis);
}

hipError_t launch_bitpar(const BitDesc &d, unsigned waves, LaunchCtl *ctl, uint32_t spin_limit, hipStream_t stream) {
    if (d.npasses <= 0 || waves == 0) return hipSuccess;
    hipLaunchKernelGGL(bitpar_kernel, dim3(waves), dim3(64), 0, stream, d, ctl, spin_limit);
    return hipGetLastError();
}

}  // namespace slhip
