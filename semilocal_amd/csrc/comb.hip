// comb.hip -- kernel (a): sticky-braid combing of a ragged batch of pairs.
//
// Restates the cell rule of semi_local::sticky_braid_mpi (reference
// semi_local.h:35-67, 212-276) as a systolic wave64 pipeline:
//
//  * one work item = one "pass" of a pair = 64*R consecutive rows (grid rows
//    of the shorter string).  Lane k of the wave owns rows
//    [pass*64R + kR, pass*64R + kR + R) and keeps their left strands and row
//    symbols in VGPRs for the whole pass.
//  * columns stream through the lanes: at step s lane k combs column s-k for
//    its R rows (top to bottom), then hands the column's top strand and
//    column symbol to lane k+1 with one DPP `wave_shr:1` move each.  Lane 0
//    takes the next column from a 64-column chunk held in a VGPR that is
//    rotated with `wave_shl:1`.  Lane 63's outgoing strands are collected
//    into a VGPR the same way and written back coalesced every 64 steps.
//  * passes of one pair run concurrently on different waves: pass p+1 reads
//    the bottom strands pass p wrote to the pair's strand buffer, in place,
//    after pass p's per-pass progress counter says the chunk is there
//    (release/acquire at agent scope).  Queue tickets are handed out in
//    (pair, pass) order, so a pass only waits on a pass that is already
//    running; every wait is bounded (spin limit) and a timeout poisons the
//    counter so that the grid still drains.
//
// Output: row->col (and optionally col->row) of the (m+n)-permutation,
// written straight into the caller's layout, including the reverse mapping of
// semi_local.h:98-108 when the caller's first string was the longer one.
#include "common.h"

namespace slhip {

namespace {

constexpr uint32_t kNoStrand = 0xFFFFFFFFu;   // top strand of an absent column: l > t never holds
constexpr uint32_t kPadRowSym = 0x10000u;     // symbol of a padding row: equals no column symbol
constexpr uint32_t kPadColSym = 0x10001u;     // symbol of an absent column: equals no row symbol
constexpr int kDppWaveShr1 = 0x138;           // lane k <- lane k-1 (lane 0 keeps `old`)
constexpr int kDppWaveShl1 = 0x130;           // lane k <- lane k+1 (lane 63 keeps `old`)

__device__ __forceinline__ uint32_t dpp_shr1(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppWaveShr1, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t dpp_shl1(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppWaveShl1, 0xf, 0xf, false);
}
__device__ __forceinline__ uint32_t dpp_rot_down(uint32_t src) {
    // lane k <- lane k+1; lane 63's value is never used
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)src, kDppWaveShl1, 0xf, 0xf, false);
}

// Write strand `s` exiting at position `c` of the oriented problem into the
// pair's output slice.  A strand id outside [0, size) can only come from a
// defect; it is reported through ctl->error instead of being written.
__device__ __forceinline__ void emit(int32_t *rc, int32_t *cr, uint32_t size, bool reverse, uint32_t s, uint32_t c,
                                     LaunchCtl *ctl) {
    if (s >= size || c >= size) {
        __hip_atomic_fetch_or(&ctl->error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    if (reverse) {   // P_ab[x] = size-1 - P_ba[size-1-x]  (semi_local.h:98-108)
        s = size - 1 - s;
        c = size - 1 - c;
    }
    rc[s] = (int32_t)c;
    if (cr) cr[c] = (int32_t)s;
}

// Bounded wait until progress[idx] >= need.  Returns false on timeout (or when
// another wave already reported one), after flagging the error.
__device__ __forceinline__ bool wait_progress(const uint32_t *progress, uint32_t idx, uint32_t need,
                                              LaunchCtl *ctl, uint32_t spin_limit) {
    uint32_t spins = 0;
    while (__hip_atomic_load(&progress[idx], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
        __builtin_amdgcn_s_sleep(2);
        if (++spins >= spin_limit ||
            __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
    }
    return true;
}

__device__ __forceinline__ void publish(uint32_t *progress, uint32_t idx, uint32_t value, int lane) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) __hip_atomic_store(&progress[idx], value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int R, typename Sym>
__device__ void comb_pass(const PairDesc &d, int pass, const Sym *__restrict__ pool, int32_t *__restrict__ rc,
                          int32_t *__restrict__ cr, uint32_t *__restrict__ buf, uint32_t *progress,
                          LaunchCtl *ctl, uint32_t spin_limit) {
    const int lane = threadIdx.x & 63;
    const uint32_t m = (uint32_t)d.m, n = (uint32_t)d.n, size = m + n;
    const bool reverse = d.reverse != 0;
    const Sym *rows = pool + d.rows_off;
    const Sym *cols = pool + d.cols_off;
    uint32_t *sbuf = buf + d.buf_off;
    int32_t *prc = rc + d.perm_off;
    int32_t *pcr = cr ? cr + d.perm_off : nullptr;
    const bool first = pass == 0;
    const bool last = pass == d.npasses - 1;
    const uint32_t item = (uint32_t)(d.item0 + pass);
    const int64_t row0 = (int64_t)pass * 64 * R + (int64_t)lane * R;

    // Left strand of grid row i is m-1-i (semi_local.h:70-82, 113-119).
    uint32_t l[R], av[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t i = row0 + q;
        const bool ok = i < (int64_t)m;
        l[q] = ok ? (uint32_t)(m - 1 - i) : 0u;
        av[q] = ok ? (uint32_t)rows[i] : kPadRowSym;
    }

    const uint32_t nchunks = (n + 63u + 63u) / 64u;   // steps 0 .. n+62 inclusive
    auto load_chunk = [&](uint32_t j, uint32_t &tin, uint32_t &bin) {
        const uint32_t col = 64u * j + (uint32_t)lane;
        if (col < n) {
            tin = first ? (m + col) : sbuf[col];   // top strand ids m+j (semi_local.h:77-80)
            bin = (uint32_t)cols[col];
        } else {
            tin = kNoStrand;
            bin = kPadColSym;
        }
    };
    auto chunk_ready = [&](uint32_t j) -> bool {
        if (first) return true;
        const uint32_t need = (j + 2u < nchunks) ? j + 2u : nchunks;
        return wait_progress(progress, item - 1u, need, ctl, spin_limit);
    };

    uint32_t tin_nx = kNoStrand, bin_nx = kPadColSym;
    bool ok = chunk_ready(0);
    if (ok) load_chunk(0, tin_nx, bin_nx);
    uint32_t t = kNoStrand, bs = kPadColSym, acc = 0;
    for (uint32_t j = 0; j < nchunks && ok; ++j) {
        uint32_t tin = tin_nx, bin = bin_nx;
        if (j + 1u < nchunks) {
            ok = chunk_ready(j + 1u);
            if (ok) load_chunk(j + 1u, tin_nx, bin_nx);
        }
#pragma unroll 4
        for (int u = 0; u < 64; ++u) {
            const uint32_t tin_rot = dpp_rot_down(tin);
            const uint32_t bin_rot = dpp_rot_down(bin);
            t = dpp_shr1(tin, t);      // lane 0 takes column s, lane k>0 takes lane k-1's
            bs = dpp_shr1(bin, bs);
            tin = tin_rot;
            bin = bin_rot;
#pragma unroll
            for (int q = 0; q < R; ++q) {
                // combing rule: swap iff symbols match or the strands already crossed
                const bool r = (av[q] == bs) | (l[q] > t);
                const uint32_t nl = r ? t : l[q];
                t = r ? l[q] : t;
                l[q] = nl;
            }
            acc = dpp_shl1(t, acc);    // lane 63's outgoing strand enters the collector
        }
        // acc lane k holds the final top strand of column 64j + k - 63
        const int64_t col = 64 * (int64_t)j + lane - 63;
        if (col >= 0 && col < (int64_t)n) {
            if (last) emit(prc, pcr, size, reverse, acc, (uint32_t)col, ctl);
            else sbuf[col] = acc;
        }
        if (!last) publish(progress, item, j + 1u, lane);
    }
    if (!ok) {
        // timed out: let any successor drain immediately (results are void)
        if (!last) publish(progress, item, 0xFFFFFFFFu, lane);
        return;
    }
    // right-edge exits: row i leaves at column n + (m-1-i)  (semi_local.h:93-95)
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t i = row0 + q;
        if (i < (int64_t)m) emit(prc, pcr, size, reverse, l[q], n + (uint32_t)(m - 1 - i), ctl);
    }
}

}  // namespace

template <typename Sym>
__global__ __launch_bounds__(256) void comb_batch_kernel(const PairDesc *__restrict__ pairs,
                                                         const WorkItem *__restrict__ items, uint32_t nitems,
                                                         const Sym *__restrict__ pool, int32_t *__restrict__ rc,
                                                         int32_t *__restrict__ cr, uint32_t *__restrict__ buf,
                                                         uint32_t *progress, LaunchCtl *ctl, uint32_t spin_limit) {
    const int lane = threadIdx.x & 63;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&ctl->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane(tk);
        if (tk >= nitems) break;     // every wave leaves once the queue is empty
        const WorkItem it = items[tk];
        const PairDesc d = pairs[it.pair];
        switch (d.rows_per_lane) {
            case 1: comb_pass<1, Sym>(d, it.pass, pool, rc, cr, buf, progress, ctl, spin_limit); break;
            case 2: comb_pass<2, Sym>(d, it.pass, pool, rc, cr, buf, progress, ctl, spin_limit); break;
            case 4: comb_pass<4, Sym>(d, it.pass, pool, rc, cr, buf, progress, ctl, spin_limit); break;
            case 8: comb_pass<8, Sym>(d, it.pass, pool, rc, cr, buf, progress, ctl, spin_limit); break;
            default: comb_pass<16, Sym>(d, it.pass, pool, rc, cr, buf, progress, ctl, spin_limit); break;
        }
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t value, uint64_t count) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    uint64_t blocks = (count + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, value, count);
    return hipGetLastError();
}

hipError_t launch_comb_batch(int sym_bytes, unsigned waves, hipStream_t stream, const PairDesc *pairs,
                             const WorkItem *items, uint32_t nitems, const void *pool, int32_t *rc, int32_t *cr,
                             uint32_t *buf, uint32_t *progress, LaunchCtl *ctl, uint32_t spin_limit) {
    if (nitems == 0 || waves == 0) return hipSuccess;
    const unsigned blocks = (waves + 3u) / 4u;
    if (sym_bytes == 1)
        hipLaunchKernelGGL(comb_batch_kernel<uint8_t>, dim3(blocks), dim3(256), 0, stream, pairs, items, nitems,
                           (const uint8_t *)pool, rc, cr, buf, progress, ctl, spin_limit);
    else
        hipLaunchKernelGGL(comb_batch_kernel<uint16_t>, dim3(blocks), dim3(256), 0, stream, pairs, items, nitems,
                           (const uint16_t *)pool, rc, cr, buf, progress, ctl, spin_limit);
    return hipGetLastError();
}

}  // namespace slhip
