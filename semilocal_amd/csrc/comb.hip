// comb.hip -- kernel (a): sticky-braid combing of a ragged batch of pairs.
//
// Restates the cell rule of semi_local::sticky_braid_mpi (reference
// semi_local.h:35-67, 212-276) as a systolic wave64 pipeline:
//
//  * one work item = one "pass" of a pair = 64*R consecutive rows (grid rows
//    of the shorter string).  Lane k of the wave owns rows
//    [pass*64R + kR, pass*64R + kR + R) and keeps their left strands and row
//    symbol values in VGPRs for the whole pass.
//  * columns stream through the lanes: at step s lane k combs column s-k for
//    its R rows (top to bottom) and hands the column's top strand to lane k+1
//    with one DPP `wave_shr:1` move; lane 0 takes the next column's strand.
//    Column symbol values and the injected strands come from a wave-private
//    LDS window filled 64 columns at a time (coalesced loads); every lane
//    writes its outgoing strand into an LDS slot indexed by column, so the
//    slot ends up holding lane 63's value, which is written back coalesced.
//    Per step: R cells of 3 VALU ops (one-hot symbol tags, or a per-lane
//    table of row-match masks) or 4 (symbol codes), one DPP move, two LDS
//    reads (three with the table) and one LDS write.
//  * passes of one pair run concurrently on different waves: pass p+1 reads
//    the bottom strands pass p wrote to the pair's strand buffer, in place,
//    as tagged 64-bit entries (strand + writer tag) with device-coherent
//    accesses.  The queue hands out the passes of a pair in pass order, so a
//    pass only waits on a pass that is already running; every wait is bounded
//    (spin limit) and a timeout makes every wave of the launch leave.
//
// Output: row->col (and optionally col->row) of the (m+n)-permutation,
// written straight into the caller's layout, including the reverse mapping of
// semi_local.h:98-108 when the caller's first string was the longer one.
#include "common.h"

namespace slhip {

namespace {

constexpr uint32_t kNoStrand = 0xFFFFFFFFu;   // top strand of an absent column: l > t never holds
constexpr int kDppWaveShr1 = 0x138;           // lane k <- lane k-1 (lane 0 keeps `old`)

__device__ __forceinline__ uint32_t dpp_shr1(uint32_t old, uint32_t src) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, kDppWaveShr1, 0xf, 0xf, false);
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

// Inter-pass hand-off ("tagged strands").  Pass p+1 of a pair reads the
// bottom strands of pass p from the pair's strand buffer, in place.  Every
// entry is 64 bits: the strand in the low word and, in the high word, a tag =
// (run epoch << 16) | (index of the writing pass + 1).  Entries are written and
// read with device-coherent (agent-scope) relaxed atomics -- single-copy
// atomic 64-bit accesses that gfx950 performs with the sc1 bit, past the
// per-XCD L2s -- so a consumer that reads the tag it expects has the strand
// that goes with it.  No progress counter, no fence and no wait for store
// completion is needed; a consumer whose chunk is not there yet re-reads it
// (bounded).  Stale entries carry an older tag: every run rewrites every entry
// of a multi-pass pair, and a run's epoch differs from the previous run's.
__device__ __forceinline__ uint64_t ld_coherent64(const uint64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_coherent64(uint64_t *p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t strand_tag(uint32_t epoch, uint32_t writer_pass) {
    return (epoch << 16) | ((writer_pass + 1u) & 0xFFFFu);
}
__device__ __forceinline__ bool launch_failed(LaunchCtl *ctl) {
    return __hip_atomic_load(&ctl->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}
__device__ __forceinline__ void report_timeout(LaunchCtl *ctl) {
    __hip_atomic_fetch_or(&ctl->error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// LDS words of one wave's private staging area: column values of the
// 128-column window read by the lanes, the 64 top strands injected at lane 0,
// and the 128-slot window the lanes' outgoing strands are written into.  The
// row-match table (kCellTable pairs) follows it: (sigma + 1) rows of 64 words,
// word s*64 + k = mask of lane k's rows that hold symbol s; row sigma is zero
// (absent columns).  Symbol-major rows make every table read conflict-free:
// lane k always reads bank k.
constexpr int kWinB = 0;      // [0, 128):   column value of column 64j-64+x
constexpr int kInjT = 128;    // [128, 192): top strand entering column 64j+x
constexpr int kOutT = 192;    // [192, 320): strand leaving lane k at step u -> slot 64+u-k
constexpr int kWaveLds = 320;
constexpr int kWavesPerBlock = 4;

// Ordering point for lanes of one wave that exchange data through LDS.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One cell of the combing rule (semi_local.h:46-61): the strands swap iff the
// symbols match or the left strand already crossed the top one (l > t).
// A match lifts l above every strand id (a tag bit above the id bits), so a
// single unsigned min gives the new left strand; since {l', t'} = {l, t} as a
// multiset, t' = l ^ t ^ l', one v_bitop3_b32 (three-input XOR, truth table
// 0x96) on gfx950.
//  ONEHOT: A & B is zero, or a bit above the id bits iff the symbols match:
//          one-hot symbol masks shifted above the ids (kCellOneHot), or the
//          row's own mask bit and the column symbol's table word
//          (kCellTable).  v_and_or_b32, v_min_u32, v_bitop3_b32 (3 VALU).
//  else:   A and B are symbol codes; ((A ^ B) - 1) has bit 31 set iff they are
//          equal: v_xad_u32, v_and_or_b32, v_min_u32, v_bitop3_b32 (4 VALU).
template <bool ONEHOT>
__device__ __forceinline__ void comb_cell(uint32_t A, uint32_t B, uint32_t &l, uint32_t &t) {
    uint32_t x;
    if (ONEHOT) x = (A & B) | l;
    else x = (((A ^ B) - 1u) & 0x80000000u) | l;
    const uint32_t nl = x < t ? x : t;
    t = __builtin_amdgcn_bitop3_b32(l, t, nl, 0x96);
    l = nl;
}

// Column value in the LDS window: the one-hot tag, the code, or the byte
// offset of the symbol's table row (64 words).
template <int MODE>
__device__ __forceinline__ uint32_t sym_value(uint32_t code, uint32_t tag_shift) {
    return MODE == kCellOneHot ? (1u << (tag_shift + code)) : MODE == kCellTable ? code * 256u : code;
}

// Always inlined into the queue loop: the pair descriptor and the loop's
// uniform values then stay in scalar registers.
template <int R, int MODE, typename Sym>
__device__ __attribute__((always_inline)) void comb_pass(const PairDesc &d, int pass, const Sym *__restrict__ pool, int32_t *__restrict__ rc,
                          int32_t *__restrict__ cr, uint64_t *__restrict__ buf, uint32_t epoch,
                          LaunchCtl *ctl, uint32_t spin_limit, uint32_t *W, uint32_t *T, uint32_t *wait_flags) {
    constexpr bool ONEHOT = MODE != kCellCode;
    const int lane = threadIdx.x & 63;
    const uint32_t m = (uint32_t)d.m, n = (uint32_t)d.n, size = m + n;
    const uint32_t tag_shift = (uint32_t)d.tag_shift;
    const bool reverse = d.reverse != 0;
    const Sym *rows = pool + d.rows_off;
    const Sym *cols = pool + d.cols_off;
    uint64_t *sbuf = buf + d.buf_off;
    int32_t *prc = rc + d.perm_off;
    int32_t *pcr = cr ? cr + d.perm_off : nullptr;
    const bool first = pass == 0;
    const bool last = pass == d.npasses - 1;
    const uint32_t want_tag = strand_tag(epoch, (uint32_t)pass - 1u);   // entries written by pass-1
    const uint64_t my_tag = (uint64_t)strand_tag(epoch, (uint32_t)pass) << 32;
    // passes before the last hold 64 d.rows_per_lane rows each; the last may
    // run at a smaller height R (the fewest rows that cover the rest)
    const int64_t row0 = (int64_t)pass * 64 * d.rows_per_lane + (int64_t)lane * R;
    // values of padding rows / absent columns: never equal to anything
    const uint32_t pad_row = MODE == kCellCode ? 0x10000u : 0u;
    const uint32_t pad_col = MODE == kCellOneHot ? 0u : MODE == kCellTable ? (uint32_t)d.sigma * 256u : 0x20000u;

    // Left strand of grid row i is m-1-i (semi_local.h:70-82, 113-119).
    uint32_t l[R], A[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
        const int64_t i = row0 + q;
        const bool ok = i < (int64_t)m;
        l[q] = ok ? (uint32_t)(m - 1 - i) : 0u;
        A[q] = ok ? (MODE == kCellTable ? (uint32_t)rows[i] : sym_value<MODE>((uint32_t)rows[i], tag_shift)) : pad_row;
    }
    if (MODE == kCellTable) {
        // lane k's table words: bit 32-R+q set iff row q holds the symbol
        // (ids fit below bit 32-R); row sigma stays zero for absent columns
        const uint32_t sigma = (uint32_t)d.sigma;
        for (uint32_t sym = 0; sym <= sigma; ++sym) {
            uint32_t mask = 0;
#pragma unroll
            for (int q = 0; q < R; ++q)
                if (row0 + q < (int64_t)m && A[q] == sym) mask |= 1u << (32 - R + q);
            T[sym * 64u + (uint32_t)lane] = mask;
        }
#pragma unroll
        for (int q = 0; q < R; ++q) A[q] = 1u << (32 - R + q);
        wave_lds_sync();
    }

    const uint32_t nchunks = (n + 63u + 63u) / 64u;   // steps 0 .. n+62 inclusive
    // Chunk j = columns [64j, 64j+64): lane k holds column 64j+k.  For pass > 0
    // the top strands come from the strand buffer as tagged entries.
    auto issue_chunk = [&](uint32_t j, uint64_t &tv, uint32_t &bin) {
        const uint32_t col = 64u * j + (uint32_t)lane;
        if (col < n) {
            tv = first ? (uint64_t)(m + col) : ld_coherent64(&sbuf[col]);   // top strand ids m+j (semi_local.h:77-80)
            bin = sym_value<MODE>((uint32_t)cols[col], tag_shift);
        } else {
            tv = kNoStrand;
            bin = pad_col;
        }
    };
    // Bounded wait until every lane's entry of chunk j carries want_tag.
    // While a pass waits on its predecessor it raises the predecessor's wait
    // flag (entry item0 + pass - 1: (tag << 1) | 1), and lowers it when the
    // wait ends; the predecessor reads its flag once per chunk and runs at
    // raised wave priority while it is up.  A chain of passes then runs at
    // about the speed of its average pass instead of its slowest one: waves
    // of equal priority on a SIMD issue by wave age, so one pass of a chain
    // can be held back for its whole length.
    uint32_t *const my_flag = wait_flags && !first ? wait_flags + d.item0 + pass - 1 : nullptr;
    const uint32_t *const succ_flag = wait_flags && !last ? wait_flags + d.item0 + pass : nullptr;
    auto settle_chunk = [&](uint32_t j, uint64_t &tv) -> bool {
        if (first) return true;
        const uint32_t col = 64u * j + (uint32_t)lane;
        uint32_t spins = 0;
        for (;;) {
            const bool ready = col >= n || (uint32_t)(tv >> 32) == want_tag;
            if (__builtin_amdgcn_ballot_w64(!ready) == 0) {
                if (spins && my_flag && lane == 0)
                    __hip_atomic_store(my_flag, epoch << 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return true;
            }
            if (spins == 0 && my_flag && lane == 0)
                __hip_atomic_store(my_flag, (epoch << 1) | 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (++spins >= spin_limit || launch_failed(ctl)) {
                report_timeout(ctl);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
            if (!ready) tv = ld_coherent64(&sbuf[col]);
        }
    };

    uint64_t tv_nx = kNoStrand;
    uint32_t bin_nx = pad_col;
    issue_chunk(0, tv_nx, bin_nx);
    uint32_t t = kNoStrand, b_prev = pad_col;
    uint32_t succ_waiting = 0;   // the successor's wait flag, read half a chunk ahead of its use
    bool ok = true;
    for (uint32_t j = 0; j < nchunks; ++j) {
        if (!settle_chunk(j, tv_nx)) {
            ok = false;
            break;
        }
        if (succ_flag && !d.prio) {
            if (__builtin_amdgcn_readfirstlane(succ_waiting) == ((epoch << 1) | 1u)) __builtin_amdgcn_s_setprio(2);
            else __builtin_amdgcn_s_setprio(0);
        }
        const uint32_t tin = (uint32_t)tv_nx, bin = bin_nx;
        W[kWinB + lane] = b_prev;
        W[kWinB + 64 + lane] = bin;
        W[kInjT + lane] = tin;
        b_prev = bin;
        wave_lds_sync();
        const uint32_t *wb = W + kWinB + 64 - lane;   // wb[u]: column 64j+u-lane
        uint32_t *wo = W + kOutT + 64 - lane;         // wo[u]: slot of column 64j+u-lane
        const char *tl = (const char *)(T + lane);    // this lane's table column
        // Column values (row-match masks for kCellTable) and injected strands
        // of a block of 8 steps are in registers before the block starts: the
        // next block's reads are issued during the current one, so no LDS
        // round trip sits on the step chain.  (The compiler cannot move them
        // itself: it cannot tell these reads from the wo[] writes.)
        auto lookup = [&](uint32_t (&v)[8]) {
            if (MODE == kCellTable) {
#pragma unroll
                for (int uu = 0; uu < 8; ++uu) v[uu] = *(const uint32_t *)(tl + v[uu]);   // row-match mask
            }
        };
        // One block of 8 steps from (bv, inj); reads the block at `un` into
        // (bv_nx, inj_nx) on the way.
        auto block = [&](int u0, int un, const uint32_t (&bv)[8], const uint32_t (&inj)[8], uint32_t (&bv_nx)[8],
                         uint32_t (&inj_nx)[8]) {
#pragma unroll
            for (int uu = 0; uu < 8; ++uu) {
                bv_nx[uu] = wb[un + uu];
                inj_nx[uu] = W[kInjT + un + uu];
            }
#pragma unroll
            for (int uu = 0; uu < 8; ++uu) {
                if (uu == 4) lookup(bv_nx);
                t = dpp_shr1(inj[uu], t);   // lane 0 takes column 64j+u, lane k>0 lane k-1's strand
#pragma unroll
                for (int q = 0; q < R; ++q) comb_cell<ONEHOT>(A[q], bv[uu], l[q], t);
                wo[u0 + uu] = t;
            }
        };
        // Steps [u_begin, u_end), a multiple of 16, two blocks per iteration
        // with the register sets swapped (no copies between blocks).
        auto steps = [&](int u_begin, int u_end) {
            uint32_t b0[8], i0[8], b1[8], i1[8];
#pragma unroll
            for (int uu = 0; uu < 8; ++uu) {
                b0[uu] = wb[u_begin + uu];
                i0[uu] = W[kInjT + u_begin + uu];
            }
            lookup(b0);
            for (int u0 = u_begin; u0 < u_end; u0 += 16) {
                block(u0, u0 + 8, b0, i0, b1, i1);
                const int un = u0 + 16 < u_end ? u0 + 16 : u0;   // last pair of blocks: re-reads valid slots
                block(u0 + 8, un, b1, i1, b0, i0);
            }
        };
        steps(0, 32);
        // read chunk j+1 half-way through chunk j: it is needed when chunk j+1
        // starts, and reading it late shortens the lag behind the previous pass
        if (j + 1u < nchunks) issue_chunk(j + 1u, tv_nx, bin_nx);
        if (succ_flag && !d.prio)
            succ_waiting = __hip_atomic_load(succ_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        steps(32, 64);
        wave_lds_sync();
        // slot 1+k was written last by lane 63 (step k): final top strand of column 64j+k-63
        const uint32_t acc = W[kOutT + 1 + lane];
        const int64_t col = 64 * (int64_t)j + lane - 63;
        if (col >= 0 && col < (int64_t)n) {
            if (last) emit(prc, pcr, size, reverse, acc, (uint32_t)col, ctl);
            else st_coherent64(&sbuf[col], my_tag | acc);
        }
        wave_lds_sync();
    }
    // right-edge exits: row i leaves at column n + (m-1-i)  (semi_local.h:93-95).
    // Skipped after a reported timeout; no early return, so the wave reaches
    // the ticket loop's readfirstlane converged.
    if (ok) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
            const int64_t i = row0 + q;
            if (i < (int64_t)m) emit(prc, pcr, size, reverse, l[q], n + (uint32_t)(m - 1 - i), ctl);
        }
    }
}

// One pass at tile height R, with the pair's symbol test.
template <int R, typename Sym>
__device__ __attribute__((always_inline)) void comb_pass_cells(const PairDesc &d, int pass, const Sym *pool,
                                                               int32_t *rc, int32_t *cr, uint64_t *buf,
                                                               uint32_t epoch, LaunchCtl *ctl, uint32_t spin_limit,
                                                               uint32_t table_words, uint32_t *W, uint32_t *T,
                                                               uint32_t *wait_flags) {
    if (d.cell == kCellOneHot)
        comb_pass<R, kCellOneHot, Sym>(d, pass, pool, rc, cr, buf, epoch, ctl, spin_limit, W, T, wait_flags);
    else if (d.cell == kCellTable && table_words != 0)
        comb_pass<R, kCellTable, Sym>(d, pass, pool, rc, cr, buf, epoch, ctl, spin_limit, W, T, wait_flags);
    else
        comb_pass<R, kCellCode, Sym>(d, pass, pool, rc, cr, buf, epoch, ctl, spin_limit, W, T, wait_flags);
}

}  // namespace

// Persistent queue consumer: every wave takes (pair, pass) tickets until the
// queue is empty.  RSET = 4, 8 or 16: every pass of the launch has that tile
// height (one instantiation per height); RSET = 0: each pair has its own
// (d.rows_per_lane, any of 4, 8, 16; its last pass d.last_rows, no larger),
// dispatched per ticket like the cell mode.  A pair marked `prio` (a chain of passes on the batch's critical path)
// runs its passes at raised wave priority, so on a SIMD shared with bulk
// passes its instructions issue first.  wait_flags (one word per queue item,
// or null) carry the wait feedback between consecutive passes of a pair.
template <typename Sym, int RSET>
__global__ __launch_bounds__(256) void comb_batch_kernel(const PairDesc *__restrict__ pairs,
                                                         const WorkItem *__restrict__ items, uint32_t nitems,
                                                         const Sym *__restrict__ pool, int32_t *__restrict__ rc,
                                                         int32_t *__restrict__ cr, uint64_t *__restrict__ buf,
                                                         uint32_t epoch, LaunchCtl *ctl, uint32_t spin_limit,
                                                         uint32_t table_words, uint32_t *wait_flags) {
    extern __shared__ uint32_t lds[];   // kWavesPerBlock x (kWaveLds + table_words) words
    const int lane = threadIdx.x & 63;
    uint32_t *W = lds + (threadIdx.x >> 6) * (kWaveLds + table_words);   // this wave's private staging area
    uint32_t *T = W + kWaveLds;                                          // its row-match table
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0) tk = atomicAdd(&ctl->ticket, 1u);
        tk = __builtin_amdgcn_readfirstlane(tk);
        if (tk >= nitems) break;     // every wave leaves once the queue is empty
        const WorkItem it = items[tk];
        const PairDesc d = pairs[it.pair];
        const int Rm = d.rows_per_lane;
        const int R = it.pass == d.npasses - 1 ? d.last_rows : Rm;
        const bool height_ok = (Rm == 4 || Rm == 8 || Rm == 16) && (R == 4 || R == 8 || R == 16) && R <= Rm;
        if (RSET != 0 ? (R != RSET || Rm != RSET) : !height_ok) {
            // a descriptor this instantiation cannot comb (a planning defect):
            // reported, never run with the wrong row layout
            if (lane == 0) __hip_atomic_fetch_or(&ctl->error, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            continue;
        }
        if (d.prio) __builtin_amdgcn_s_setprio(3);
        if (RSET == 4 || (RSET == 0 && R == 4))
            comb_pass_cells<4, Sym>(d, it.pass, pool, rc, cr, buf, epoch, ctl, spin_limit, table_words, W, T,
                                    wait_flags);
        else if (RSET == 8 || (RSET == 0 && R == 8))
            comb_pass_cells<8, Sym>(d, it.pass, pool, rc, cr, buf, epoch, ctl, spin_limit, table_words, W, T,
                                    wait_flags);
        else
            comb_pass_cells<16, Sym>(d, it.pass, pool, rc, cr, buf, epoch, ctl, spin_limit, table_words, W, T,
                                     wait_flags);
        __builtin_amdgcn_s_setprio(0);
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

template <typename Sym, int RSET>
static void launch_one(unsigned blocks, hipStream_t stream, const PairDesc *pairs, const WorkItem *items,
                       uint32_t nitems, const void *pool, int32_t *rc, int32_t *cr, uint64_t *buf, uint32_t epoch,
                       LaunchCtl *ctl, uint32_t spin_limit, uint32_t table_words, uint32_t *wait_flags) {
    const size_t lds_bytes = sizeof(uint32_t) * kWavesPerBlock * (kWaveLds + table_words);
    hipLaunchKernelGGL((comb_batch_kernel<Sym, RSET>), dim3(blocks), dim3(kWavesPerBlock * 64), lds_bytes, stream, pairs,
                       items, nitems, (const Sym *)pool, rc, cr, buf, epoch, ctl, spin_limit, table_words,
                       wait_flags);
}

// rows_per_lane: 4, 8 or 16 when every pair of the launch has that tile
// height, 0 when the pairs carry their own.  table_words: (largest sigma of
// the batch's kCellTable pairs + 1) x 64, or 0: then kCellTable pairs take
// the code cell (one fewer LDS read per step, the lower-latency choice for
// launches with few waves).
hipError_t launch_comb_batch(int sym_bytes, int rows_per_lane, unsigned waves, hipStream_t stream,
                             const PairDesc *pairs, const WorkItem *items, uint32_t nitems, const void *pool,
                             int32_t *rc, int32_t *cr, uint64_t *buf, uint32_t epoch, LaunchCtl *ctl,
                             uint32_t spin_limit, uint32_t table_words, uint32_t *wait_flags) {
    if (nitems == 0 || waves == 0) return hipSuccess;
    if (table_words > (uint32_t)(kTableMaxSymbols + 1) * 64u) return hipErrorInvalidValue;
    const unsigned blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
#define SL_LAUNCH(SYM, RR) \
    launch_one<SYM, RR>(blocks, stream, pairs, items, nitems, pool, rc, cr, buf, epoch, ctl, spin_limit, table_words, \
                        wait_flags)
    if (rows_per_lane != 0 && rows_per_lane != 4 && rows_per_lane != 8 && rows_per_lane != 16)
        return hipErrorInvalidValue;
    if (sym_bytes == 1) {
        if (rows_per_lane == 0) SL_LAUNCH(uint8_t, 0);
        else if (rows_per_lane == 4) SL_LAUNCH(uint8_t, 4);
        else if (rows_per_lane == 8) SL_LAUNCH(uint8_t, 8);
        else SL_LAUNCH(uint8_t, 16);
    } else {
        if (rows_per_lane == 0) SL_LAUNCH(uint16_t, 0);
        else if (rows_per_lane == 4) SL_LAUNCH(uint16_t, 4);
        else if (rows_per_lane == 8) SL_LAUNCH(uint16_t, 8);
        else SL_LAUNCH(uint16_t, 16);
    }
#undef SL_LAUNCH
    return hipGetLastError();
}

}  // namespace slhip
