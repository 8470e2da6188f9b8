// mult_api.hip -- host side of kernel (c): the level planner of the batched
// braid multiplication, staggered multiplication (glue of two block braids)
// and the hybrid block-decomposition scheduler, with their C ABI.
//
// Reference entry points replaced:
//   steady_ant_parallel_wrapper / steady_ant_optimized_seq
//       (unit_monge_mult/steady_ant.h:811-812, :378-453),
//   staggered_sticky_multiplication<RowGlue, UseParallel> (steady_ant.h:609-612),
//   hybrid_iterative<Input, WithIf> (semi_local.h:577-580).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/semilocal_hip.h"
#include "steady_ant.h"

namespace slhip {

// Internal interface of the combing batch (api.hip).
struct Batch;
int batch_create_internal(Batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                          const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes);
int batch_launch_internal(Batch *bt);
int batch_launch_replayable_internal(Batch *bt);
int batch_wait_internal(Batch *bt);
void batch_destroy_internal(Batch *bt);
hipStream_t batch_stream_internal(const Batch *bt);
const int32_t *batch_rc_internal(const Batch *bt);
int64_t batch_perm_off_internal(const Batch *bt, int pair);
int64_t batch_cells_internal(const Batch *bt);

namespace sa {

// Kernels (steady_ant.hip).
__global__ void load_pq(const int32_t *, const int32_t *, const int64_t *, const int32_t *, const SaProb *,
                        int32_t *, int32_t *);
__global__ void store_r(const int32_t *, const int32_t *, const SaProb *, const int64_t *, int32_t *);
__global__ void split_count(const int32_t *, const int32_t *, const int32_t *, const SaProb *, int, int32_t *,
                            int32_t *);
__global__ void split_scatter(const int32_t *, const int32_t *, const int32_t *, const SaProb *, int,
                              const int32_t *, const int32_t *, int32_t *, int32_t *, int32_t *, int32_t *);
__global__ void solve_small(const int32_t *, const int32_t *, const int32_t *, const SaProb *, int, int32_t *);
__global__ void merge_map(const int32_t *, const int32_t *, const int32_t *, const int32_t *, const SaProb *,
                          const int32_t *, int, int64_t, uint32_t *, uint32_t *, int32_t *, int32_t *);
__global__ void table_rows(const uint32_t *, const SaSeg *, const int2 *, int32_t *, int32_t *);
__global__ void table_small(const uint32_t *, const SaSeg *, int32_t *, int32_t *);
__global__ void table_cols(const SaSeg *, const int2 *, int, int32_t *, int32_t *);
__global__ void seg_search_walk(const uint32_t *, const uint32_t *, const int32_t *, const int32_t *, const SaSeg *,
                                const int2 *, int, const int32_t *, int32_t *);
__global__ void split_count2(const int32_t *, const int32_t *, const int32_t *, const SaProb *, int, int32_t *);
__global__ void split_scatter2(const int32_t *, const int32_t *, const int32_t *, const SaProb *, int, const int32_t *,
                               int32_t *, int32_t *, int32_t *, int32_t *, int32_t *, int32_t *);
__global__ void fill_i32(int32_t *, int32_t, int64_t);
__global__ void stag_qinv(const int32_t *, const StagJob *, const int32_t *, int32_t *);
__global__ void stag_count(const int32_t *, const int32_t *, const StagJob *, const int32_t *, int32_t *);
__global__ void stag_scan(const StagJob *, int32_t *);
__global__ void stag_scatter(const int32_t *, const int32_t *, const StagJob *, const int32_t *, const int32_t *,
                             const SaProb *, int32_t *, int32_t *, int32_t *, int32_t *);
__global__ void stag_glue(const int32_t *, const StagJob *, const SaProb *, const int32_t *, const int32_t *,
                          const int32_t *, int32_t *);
__global__ void copy_blocks(const int32_t *, const int64_t *, int32_t *);

namespace {

template <class T>
hipError_t dalloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hipMalloc((void **)p, sizeof(T) * count);
}

template <class T>
hipError_t upload(T **p, const std::vector<T> &v) {
    hipError_t e = dalloc(p, v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

hipError_t fill(int32_t *p, int32_t v, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int blocks = (int)std::min<int64_t>(ceil_div(count, 256), 4096);
    hipLaunchKernelGGL(fill_i32, dim3(blocks), dim3(256), 0, s, p, v, count);
    return hipGetLastError();
}

}  // namespace

// ---------------------------------------------------------------------------
// Multiply plan: the level structure and device buffers of one batch of
// products with fixed sizes.  Reusable: load inputs, run, read R.
// ---------------------------------------------------------------------------
struct Level {
    int S = 0;
    int nseg = 0, nwalk = 0, nrow = 0, ncol = 0;
    int64_t tab_total = 0;
    size_t seg0 = 0, walk0 = 0, row0 = 0, col0 = 0, sf0 = 0;   // offsets into the packed arrays
};

struct MulPlan {
    int nprob = 0;
    std::vector<int32_t> n;
    std::vector<int64_t> base;
    int64_t ntot = 0;
    int ntiles = 0;
    int S_top = 1, S0 = 1;
    std::vector<Level> lv;   // large levels, lv[i].S = S_top >> i
    int max_seg = 0;
    int64_t max_tab = 0;
    // device
    SaProb *d_probs = nullptr;
    int32_t *d_tile_prob = nullptr;
    SaSeg *d_segs = nullptr;
    int2 *d_walks = nullptr, *d_rows = nullptr, *d_cols = nullptr;
    int32_t *d_seg_first = nullptr;
    int32_t *d_P[2] = {nullptr, nullptr}, *d_Qi[2] = {nullptr, nullptr};
    int32_t *d_PR = nullptr, *d_QC = nullptr;
    int32_t *d_R[2] = {nullptr, nullptr};
    uint32_t *d_U = nullptr, *d_Ui = nullptr;
    int32_t *d_cntP = nullptr, *d_cntQ = nullptr, *d_cnt8 = nullptr, *d_cend = nullptr;
    int32_t *d_cntHi = nullptr, *d_cntLo = nullptr;
    int32_t *d_map_row = nullptr, *d_map_col = nullptr;   // staggered: core -> block rows / cols
    int result = 0;                                        // index of the R buffer holding the product
    int64_t elements_levels = 0;                           // sum over levels of the elements processed
};

static void plan_free(MulPlan *p) {
    if (!p) return;
    void *bufs[] = {p->d_probs, p->d_tile_prob, p->d_segs, p->d_walks, p->d_rows, p->d_cols, p->d_seg_first,
                    p->d_P[0], p->d_P[1], p->d_Qi[0], p->d_Qi[1], p->d_PR, p->d_QC, p->d_R[0], p->d_R[1],
                    p->d_U, p->d_Ui, p->d_cntP, p->d_cntQ, p->d_cnt8, p->d_cend, p->d_cntHi, p->d_cntLo, p->d_map_row, p->d_map_col};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    delete p;
}

#define PLAN_TRY(expr)                                                                    \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess) {                                                           \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));                 \
            plan_free(p);                                                                 \
            return _e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;                       \
        }                                                                                 \
    } while (0)

static int plan_create(MulPlan **out, const std::vector<int32_t> &sizes, bool staggered) {
    *out = nullptr;
    for (int32_t ni : sizes)
        if ((int64_t)ni > (int64_t)kMaxNb * kBlk) {
            set_error("braid multiply: size " + std::to_string(ni) + " exceeds the supported maximum " +
                      std::to_string((int64_t)kMaxNb * kBlk));
            return kErrUnsupported;
        }
    MulPlan *p = new MulPlan();
    p->nprob = (int)sizes.size();
    p->n = sizes;
    p->base.resize(sizes.size());
    std::vector<int32_t> tile_prob;
    int kmax = 0;
    for (int i = 0; i < p->nprob; ++i) {
        const int32_t ni = sizes[(size_t)i];
        p->base[(size_t)i] = p->ntot;
        const int64_t tiles = ceil_div(ni, kTile);
        for (int64_t t = 0; t < tiles; ++t) tile_prob.push_back(i);
        p->ntot += tiles * kTile;
        int k = 0;
        while ((int64_t(1) << k) < ni) ++k;
        kmax = std::max(kmax, k);
    }
    p->ntiles = (int)tile_prob.size();
    p->S_top = 1 << kmax;
    p->S0 = std::min(p->S_top, kTile);
    // large levels
    std::vector<SaSeg> segs;
    std::vector<int2> walks, rows, cols;
    std::vector<int32_t> seg_first;
    for (int S = p->S_top; S > kTile; S >>= 1) {
        Level L;
        L.S = S;
        L.seg0 = segs.size();
        L.walk0 = walks.size();
        L.row0 = rows.size();
        L.col0 = cols.size();
        L.sf0 = seg_first.size();
        for (int i = 0; i < p->nprob; ++i) {
            seg_first.push_back(L.nseg);
            const int32_t ni = sizes[(size_t)i];
            for (int64_t j = 0; j * S < ni; ++j) {
                SaSeg sg{};
                sg.x0 = p->base[(size_t)i] + j * S;
                sg.s = (int32_t)std::min<int64_t>(S, ni - j * S);
                sg.tile0 = (int32_t)(sg.x0 / kTile);
                sg.ntiles = (int32_t)ceil_div(sg.s, kTile);
                const bool split = sg.s > S / 2;
                sg.nb = split ? (int32_t)ceil_div(sg.s, kBlk) : 0;
                sg.tab_off2 = L.tab_total;
                if (split) {
                    L.tab_total += (int64_t)(sg.nb + 1) * (sg.nb + 1);
                    const int nw = (int)ceil_div(2 * (int64_t)sg.s, kSegSteps);
                    for (int k = 0; k < nw; ++k) walks.push_back(make_int2(L.nseg, k));
                    for (int r = 0; r < sg.nb; ++r) rows.push_back(make_int2(L.nseg, r));
                    for (int c = 0; c <= sg.nb; ++c) cols.push_back(make_int2(L.nseg, c));
                    L.nwalk += nw;
                    L.nrow += sg.nb;
                    L.ncol += sg.nb + 1;
                }
                segs.push_back(sg);
                ++L.nseg;
            }
        }
        p->max_seg = std::max(p->max_seg, L.nseg);
        p->max_tab = std::max(p->max_tab, L.tab_total);
        p->lv.push_back(L);
    }
    for (int S = p->S_top; S > 1; S >>= 1) p->elements_levels += p->ntot;
    std::vector<SaProb> probs((size_t)p->nprob);
    for (int i = 0; i < p->nprob; ++i) probs[(size_t)i] = SaProb{p->base[(size_t)i], sizes[(size_t)i], 0};
    const size_t nt = (size_t)p->ntot, nl = p->lv.size();
    PLAN_TRY(upload(&p->d_probs, probs));
    PLAN_TRY(upload(&p->d_tile_prob, tile_prob));
    PLAN_TRY(upload(&p->d_segs, segs));
    PLAN_TRY(upload(&p->d_walks, walks));
    PLAN_TRY(upload(&p->d_rows, rows));
    PLAN_TRY(upload(&p->d_cols, cols));
    PLAN_TRY(upload(&p->d_seg_first, seg_first));
    for (int b = 0; b < 2; ++b) {
        PLAN_TRY(dalloc(&p->d_P[b], nt));
        PLAN_TRY(dalloc(&p->d_Qi[b], nt));
        PLAN_TRY(dalloc(&p->d_R[b], nt));
    }
    PLAN_TRY(dalloc(&p->d_PR, nt * std::max<size_t>(nl, 1)));
    PLAN_TRY(dalloc(&p->d_QC, nt * std::max<size_t>(nl, 1)));
    // slack past the last tile: the segment searches read aligned 16-byte
    // groups that can reach a few entries past a subproblem's end
    PLAN_TRY(dalloc(&p->d_U, nt + kWalkSlack));
    PLAN_TRY(dalloc(&p->d_Ui, nt + kWalkSlack));
    PLAN_TRY(dalloc(&p->d_cntP, (size_t)p->ntiles));
    PLAN_TRY(dalloc(&p->d_cntQ, (size_t)p->ntiles));
    PLAN_TRY(dalloc(&p->d_cnt8, 8 * (size_t)p->ntiles));
    PLAN_TRY(dalloc(&p->d_cend, std::max<size_t>(segs.size(), 1)));   // one entry per segment of every level
    PLAN_TRY(dalloc(&p->d_cntHi, (size_t)p->max_tab));
    PLAN_TRY(dalloc(&p->d_cntLo, (size_t)p->max_tab));
    if (staggered) {
        PLAN_TRY(dalloc(&p->d_map_row, nt));
        PLAN_TRY(dalloc(&p->d_map_col, nt));
    }
    // the uploads above go through the null stream, which is not ordered with
    // the non-blocking streams the plan later runs on: finish them here
    PLAN_TRY(hipDeviceSynchronize());
    *out = p;
    return kOk;
}

// Runs the levels on `s`; inputs are in d_P[0] (row -> col) and d_Qi[0]
// (col -> row).  The product ends in d_R[p->result].
static int plan_launch(MulPlan *p, hipStream_t s) {
    if (p->ntiles == 0) return kOk;
    const int nt = p->ntiles;
    const int64_t ntot = p->ntot;
    int cur = 0;
    size_t li = 0;
    for (; li + 1 < p->lv.size(); li += 2) {   // two split levels per pass (S and S/2)
        const Level &L = p->lv[li];
        hipLaunchKernelGGL(split_count2, dim3(nt), dim3(kSplitThreads), 0, s, p->d_P[cur], p->d_Qi[cur],
                           p->d_tile_prob, p->d_probs, L.S, p->d_cnt8);
        hipLaunchKernelGGL(split_scatter2, dim3(nt), dim3(kSplitThreads), 0, s, p->d_P[cur], p->d_Qi[cur],
                           p->d_tile_prob, p->d_probs, L.S, p->d_cnt8, p->d_P[cur ^ 1], p->d_Qi[cur ^ 1],
                           p->d_PR + li * ntot, p->d_QC + li * ntot, p->d_PR + (li + 1) * ntot,
                           p->d_QC + (li + 1) * ntot);
        cur ^= 1;
    }
    for (; li < p->lv.size(); ++li) {
        const Level &L = p->lv[li];
        hipLaunchKernelGGL(split_count, dim3(nt), dim3(kSplitThreads), 0, s, p->d_P[cur], p->d_Qi[cur],
                           p->d_tile_prob, p->d_probs, L.S, p->d_cntP, p->d_cntQ);
        hipLaunchKernelGGL(split_scatter, dim3(nt), dim3(kSplitThreads), 0, s, p->d_P[cur], p->d_Qi[cur], p->d_tile_prob,
                           p->d_probs, L.S, p->d_cntP, p->d_cntQ, p->d_P[cur ^ 1], p->d_Qi[cur ^ 1],
                           p->d_PR + li * ntot, p->d_QC + li * ntot);
        cur ^= 1;
    }
    hipLaunchKernelGGL(solve_small, dim3(nt), dim3(kTile), 0, s, p->d_P[cur], p->d_Qi[cur], p->d_tile_prob,
                       p->d_probs, p->S0, p->d_R[0]);
    int rc = 0;
    if (!p->lv.empty()) {   // every level's c_end entries, cleared by one launch
        const Level &L0 = p->lv.back();
        SL_HIP_TRY(fill(p->d_cend, 0, (int64_t)(L0.seg0 + L0.nseg), s));
    }
    for (size_t k = p->lv.size(); k-- > 0;) {
        const Level &L = p->lv[k];
        const SaSeg *segs = p->d_segs + L.seg0;
        const int2 *walks = p->d_walks + L.walk0;
        int32_t *cend = p->d_cend + L.seg0;
        hipLaunchKernelGGL(merge_map, dim3(nt), dim3(kSplitThreads), 0, s, p->d_R[rc], p->d_PR + k * ntot,
                           p->d_QC + k * ntot, p->d_tile_prob, p->d_probs, p->d_seg_first + L.sf0, L.S, ntot,
                           p->d_U, p->d_Ui, p->d_R[rc ^ 1], cend);
        rc ^= 1;
        if (L.nwalk == 0) continue;
        if (L.S > kDirectNb * kBlk && L.S <= kSmallTabNb * kBlk) {   // small tables: one launch
            hipLaunchKernelGGL(table_small, dim3((unsigned)L.nseg), dim3(256), 0, s, p->d_U, segs, p->d_cntHi,
                               p->d_cntLo);
        } else if (L.S > kDirectNb * kBlk) {   // smaller subproblems are searched by direct scans
            hipLaunchKernelGGL(table_rows, dim3((unsigned)L.nrow), dim3(256), 0, s, p->d_U, segs, p->d_rows + L.row0,
                               p->d_cntHi, p->d_cntLo);
            hipLaunchKernelGGL(table_cols, dim3((unsigned)ceil_div(L.ncol, 4)), dim3(256), 0, s, segs,
                               p->d_cols + L.col0, L.ncol, p->d_cntHi, p->d_cntLo);
        }
        hipLaunchKernelGGL(seg_search_walk, dim3((unsigned)(8 * ceil_div(ceil_div(L.nwalk, 4), 8))), dim3(256), 0, s,
                           p->d_U, p->d_Ui, p->d_cntHi, p->d_cntLo, segs, walks, L.nwalk, cend, p->d_R[rc]);
    }
    p->result = rc;
    SL_HIP_TRY(hipGetLastError());
    return kOk;
}

// ---------------------------------------------------------------------------
// Standalone products: inputs and outputs on the host, concatenated.
// ---------------------------------------------------------------------------
struct MulHandle {
    MulPlan *plan = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    int32_t *d_in_p = nullptr, *d_in_q = nullptr, *d_out = nullptr;
    int64_t *d_off = nullptr;
    int64_t total = 0;
    std::vector<int64_t> off;
};

static void mul_free(MulHandle *h) {
    if (!h) return;
    plan_free(h->plan);
    (void)hipFree(h->d_in_p);
    (void)hipFree(h->d_in_q);
    (void)hipFree(h->d_out);
    (void)hipFree(h->d_off);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

static bool is_permutation(const int32_t *v, int32_t n, std::vector<uint8_t> &seen) {
    seen.assign((size_t)n, 0);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t x = v[i];
        if (x < 0 || x >= n || seen[(size_t)x]) return false;
        seen[(size_t)x] = 1;
    }
    return true;
}

static int mul_create(MulHandle **out, int nprob, const int32_t *sizes) {
    *out = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    MulHandle *h = new MulHandle();
    std::vector<int32_t> n(sizes, sizes + nprob);
    h->off.resize((size_t)nprob);
    for (int i = 0; i < nprob; ++i) {
        h->off[(size_t)i] = h->total;
        h->total += n[(size_t)i];
    }
    rc = plan_create(&h->plan, n, false);
    if (rc) {
        mul_free(h);
        return rc;
    }
#define MUL_TRY(expr)                                                          \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
            mul_free(h);                                                       \
            return _e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;            \
        }                                                                      \
    } while (0)
    MUL_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    MUL_TRY(hipEventCreate(&h->ev[0]));
    MUL_TRY(hipEventCreate(&h->ev[1]));
    MUL_TRY(dalloc(&h->d_in_p, (size_t)h->total));
    MUL_TRY(dalloc(&h->d_in_q, (size_t)h->total));
    MUL_TRY(dalloc(&h->d_out, (size_t)h->total));
    MUL_TRY(upload(&h->d_off, h->off));
#undef MUL_TRY
    *out = h;
    return kOk;
}

static int mul_upload(MulHandle *h, const int32_t *p_cat, const int32_t *q_cat) {
    std::vector<uint8_t> seen;
    for (int i = 0; i < h->plan->nprob; ++i) {
        const int32_t ni = h->plan->n[(size_t)i];
        if (!is_permutation(p_cat + h->off[(size_t)i], ni, seen) ||
            !is_permutation(q_cat + h->off[(size_t)i], ni, seen)) {
            set_error("braid multiply: input " + std::to_string(i) + " is not a full permutation of [0, n)");
            return kErrArg;
        }
    }
    if (h->total == 0) return kOk;
    SL_HIP_TRY(hipMemcpyAsync(h->d_in_p, p_cat, sizeof(int32_t) * h->total, hipMemcpyHostToDevice, h->stream));
    SL_HIP_TRY(hipMemcpyAsync(h->d_in_q, q_cat, sizeof(int32_t) * h->total, hipMemcpyHostToDevice, h->stream));
    SL_HIP_TRY(hipStreamSynchronize(h->stream));
    return kOk;
}

static int mul_launch(MulHandle *h) {
    MulPlan *p = h->plan;
    if (p->ntiles == 0) return kOk;
    hipLaunchKernelGGL(load_pq, dim3(p->ntiles), dim3(kTile), 0, h->stream, h->d_in_p, h->d_in_q, h->d_off,
                       p->d_tile_prob, p->d_probs, p->d_P[0], p->d_Qi[0]);
    int rc = plan_launch(p, h->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(store_r, dim3(p->ntiles), dim3(kTile), 0, h->stream, p->d_R[p->result], p->d_tile_prob,
                       p->d_probs, h->d_off, h->d_out);
    SL_HIP_TRY(hipGetLastError());
    return kOk;
}

static int mul_download(MulHandle *h, int32_t *r_cat) {
    SL_HIP_TRY(hipStreamSynchronize(h->stream));
    if (h->total) SL_HIP_TRY(hipMemcpy(r_cat, h->d_out, sizeof(int32_t) * h->total, hipMemcpyDeviceToHost));
    return kOk;
}

// ---------------------------------------------------------------------------
// Hybrid scheduler (semi_local.h:577-738): comb small_m x small_n blocks in
// one batched launch, then glue them level by level with staggered
// multiplications; all merges of a level run as one batched multiply.
// ---------------------------------------------------------------------------
struct BlockRef {
    int64_t off;   // offset in the pool holding it
    int32_t size, bm, bn;
};

struct MergeLevel {
    MulPlan *plan = nullptr;
    std::vector<StagJob> jobs;
    std::vector<int64_t> moves;   // (src, dst, len) triples of blocks passed through
    StagJob *d_jobs = nullptr;
    int32_t *d_tile_job = nullptr, *d_cnt = nullptr, *d_qcr = nullptr;
    int64_t *d_moves = nullptr;
    int ntiles = 0;
    int max_len = 0;
    int64_t max_move = 0;
    int src_pool = 0, dst_pool = 0;   // 0: the combing batch's output, 1/2: the hybrid's pools
};

struct Hybrid {
    Batch *batch = nullptr;
    int m = 0, n = 0;
    std::vector<MergeLevel> levels;
    int32_t *d_pool[2] = {nullptr, nullptr};
    int final_pool = 0;
    int64_t final_off = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int64_t cells = 0;
    hipGraph_t graph = nullptr;          // the whole run (combing + merge levels), captured once
    hipGraphExec_t graph_exec = nullptr;
    bool use_graph = false;
};

static void hybrid_free(Hybrid *h) {
    if (!h) return;
    for (auto &L : h->levels) {
        plan_free(L.plan);
        (void)hipFree(L.d_jobs);
        (void)hipFree(L.d_tile_job);
        (void)hipFree(L.d_cnt);
        (void)hipFree(L.d_qcr);
        (void)hipFree(L.d_moves);
    }
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    (void)hipFree(h->d_pool[0]);
    (void)hipFree(h->d_pool[1]);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->batch) batch_destroy_internal(h->batch);
    delete h;
}

static int hybrid_create(Hybrid **out, const void *a, int m, const void *b, int n, int sym_bytes, int small_m,
                         int small_n) {
    *out = nullptr;
    if (m < 0 || n < 0 || (m > 0 && !a) || (n > 0 && !b) || small_m < 1 || small_n < 1 ||
        (sym_bytes != 1 && sym_bytes != 2 && sym_bytes != 4)) {
        set_error("hybrid_iterative: bad argument");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    Hybrid *h = new Hybrid();
    h->m = m;
    h->n = n;
    // blocks (semi_local.h:586-627): floor-sized, the last row / column of
    // blocks takes the remainder
    const int cols_per_block = n / small_n, rows_per_block = m / small_m;
    const int ntasks = small_m * small_n;
    std::vector<int64_t> a_off((size_t)ntasks), b_off((size_t)ntasks);
    std::vector<int32_t> a_len((size_t)ntasks), b_len((size_t)ntasks);
    std::vector<BlockRef> cur((size_t)ntasks);
    for (int i = 0; i < ntasks; ++i) {
        const int sc = (i % small_n) * cols_per_block;
        int ec = std::min(sc + cols_per_block, n);
        const int sr = (i / small_n) * rows_per_block;
        int er = std::min(sr + rows_per_block, m);
        if (i % small_n == small_n - 1) ec = n;
        if (i / small_n == small_m - 1) er = m;
        a_off[(size_t)i] = sr;
        a_len[(size_t)i] = er - sr;
        b_off[(size_t)i] = sc;
        b_len[(size_t)i] = ec - sc;
        cur[(size_t)i] = BlockRef{0, (er - sr) + (ec - sc), er - sr, ec - sc};
    }
    static const uint8_t dummy[4] = {0, 0, 0, 0};
    rc = batch_create_internal(&h->batch, ntasks, m ? a : dummy, a_off.data(), a_len.data(), n ? b : dummy,
                               b_off.data(), b_len.data(), sym_bytes);
    if (rc) {
        hybrid_free(h);
        return rc;
    }
    h->cells = batch_cells_internal(h->batch);
    for (int i = 0; i < ntasks; ++i) cur[(size_t)i].off = batch_perm_off_internal(h->batch, i);
    // merge levels (semi_local.h:629-725)
    int steps = 0;
    for (int t = 1; t < small_n; t *= 2) ++steps;
    for (int t = 1; t < small_m; t *= 2) ++steps;
    int block_rows = rows_per_block, block_cols = cols_per_block;
    int sm = small_m, sn = small_n;
    int src_pool = 0;
    int64_t pool_cap[2] = {0, 0};
    for (int st = 0; st < steps; ++st) {
        bool in_row = sn > sm;
        if (sn > 1 && sm > 1) in_row = block_rows >= 2 * block_cols;   // :646-650
        if (in_row) block_cols *= 2;
        else block_rows *= 2;
        const int ncols = in_row ? (sn + 1) / 2 : sn;
        const int nrows = !in_row ? (sm + 1) / 2 : sm;
        MergeLevel L;
        L.src_pool = src_pool;
        L.dst_pool = src_pool == 1 ? 2 : 1;
        std::vector<BlockRef> nxt((size_t)ncols * nrows);
        std::vector<int32_t> ksizes;
        int64_t out_off = 0, qcr_off = 0;
        std::vector<int32_t> tile_job;
        for (int i = 0; i < ncols * nrows; ++i) {
            const int r = i / ncols, c = i % ncols;
            const BlockRef *P, *Q;
            if (in_row) {
                P = &cur[(size_t)(r * sn + 2 * c)];
                Q = 2 * c + 1 < sn ? &cur[(size_t)(r * sn + 2 * c + 1)] : nullptr;
            } else {
                P = &cur[(size_t)(2 * r * sn + c)];
                Q = 2 * r + 1 < sm ? &cur[(size_t)((2 * r + 1) * sn + c)] : nullptr;
            }
            if (!Q) {   // passes through unmerged
                L.moves.push_back(P->off);
                L.moves.push_back(out_off);
                L.moves.push_back(P->size);
                L.max_move = std::max<int64_t>(L.max_move, P->size);
                nxt[(size_t)i] = BlockRef{out_off, P->size, P->bm, P->bn};
                out_off += P->size;
                continue;
            }
            StagJob jb{};
            jb.p_off = P->off;
            jb.q_off = Q->off;
            jb.pr = P->size;
            jb.qr = Q->size;
            jb.row_glue = in_row ? 0 : 1;
            jb.k = in_row ? P->bm : P->bn;   // shared strands: rows (horizontal glue) / cols (vertical)
            jb.prob = (int32_t)ksizes.size();
            ksizes.push_back(jb.k);
            jb.qcr_off = qcr_off;
            qcr_off += jb.qr;
            jb.out_off = out_off;
            const int32_t N = jb.pr + jb.qr - jb.k;
            BlockRef nb{out_off, N, in_row ? P->bm : P->bm + Q->bm, in_row ? P->bn + Q->bn : P->bn};
            out_off += N;
            jb.tileP0 = (int32_t)tile_job.size();
            jb.ntilesP = (int32_t)ceil_div(jb.pr, kTile);
            jb.tileQ0 = jb.tileP0 + jb.ntilesP;
            jb.ntilesQ = (int32_t)ceil_div(jb.qr, kTile);
            for (int t = 0; t < jb.ntilesP + jb.ntilesQ; ++t) tile_job.push_back((int32_t)L.jobs.size());
            L.max_len = std::max(L.max_len, jb.k + jb.pr + jb.qr);
            L.jobs.push_back(jb);
            nxt[(size_t)i] = nb;
        }
        L.ntiles = (int)tile_job.size();
        pool_cap[L.dst_pool - 1] = std::max(pool_cap[L.dst_pool - 1], out_off);
#define HY_TRY(expr)                                                           \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            set_error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
            hybrid_free(h);                                                    \
            return _e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;            \
        }                                                                      \
    } while (0)
        h->levels.push_back(L);
        MergeLevel &LL = h->levels.back();
        rc = plan_create(&LL.plan, ksizes, true);
        if (rc) {
            hybrid_free(h);
            return rc;
        }
        HY_TRY(upload(&LL.d_jobs, LL.jobs));
        HY_TRY(upload(&LL.d_tile_job, tile_job));
        HY_TRY(dalloc(&LL.d_cnt, (size_t)LL.ntiles));
        HY_TRY(dalloc(&LL.d_qcr, (size_t)qcr_off));
        HY_TRY(upload(&LL.d_moves, LL.moves));
        cur.swap(nxt);
        sn = ncols;
        sm = nrows;
        src_pool = LL.dst_pool;
    }
    HY_TRY(dalloc(&h->d_pool[0], (size_t)pool_cap[0]));
    HY_TRY(dalloc(&h->d_pool[1], (size_t)pool_cap[1]));
    for (auto &e : h->ev) HY_TRY(hipEventCreate(&e));
#undef HY_TRY
    h->final_pool = src_pool;
    h->final_off = cur[0].off;
    *out = h;
    return kOk;
}

static const int32_t *hybrid_pool(const Hybrid *h, int which) {
    return which == 0 ? batch_rc_internal(h->batch) : h->d_pool[which - 1];
}

// Launches the merge levels on the batch stream (after the combing launch).
static int hybrid_merges(Hybrid *h) {
    hipStream_t s = batch_stream_internal(h->batch);
    for (auto &L : h->levels) {
        const int32_t *src = hybrid_pool(h, L.src_pool);
        int32_t *dst = h->d_pool[L.dst_pool - 1];
        MulPlan *p = L.plan;
        if (!L.jobs.empty()) {
            if (L.ntiles > 0) {
                hipLaunchKernelGGL(stag_qinv, dim3(L.ntiles), dim3(kTile), 0, s, src, L.d_jobs, L.d_tile_job,
                                   L.d_qcr);
                hipLaunchKernelGGL(stag_count, dim3(L.ntiles), dim3(kTile), 0, s, src, L.d_qcr, L.d_jobs,
                                   L.d_tile_job, L.d_cnt);
                hipLaunchKernelGGL(stag_scan, dim3(2 * (unsigned)L.jobs.size()), dim3(kTile), 0, s, L.d_jobs,
                                   L.d_cnt);
                hipLaunchKernelGGL(stag_scatter, dim3(L.ntiles), dim3(kTile), 0, s, src, L.d_qcr, L.d_jobs,
                                   L.d_tile_job, L.d_cnt, p->d_probs, p->d_P[0], p->d_Qi[0], p->d_map_row,
                                   p->d_map_col);
            }
            int rc = plan_launch(p, s);
            if (rc) return rc;
            if (L.max_len > 0)
                hipLaunchKernelGGL(stag_glue, dim3((unsigned)ceil_div(L.max_len, 256), (unsigned)L.jobs.size()),
                                   dim3(256), 0, s, src, L.d_jobs, p->d_probs, p->d_R[p->result], p->d_map_row,
                                   p->d_map_col, dst);
        }
        if (!L.moves.empty() && L.max_move > 0) {
            const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(L.max_move, 256), 1024);
            hipLaunchKernelGGL(copy_blocks, dim3(gx, (unsigned)(L.moves.size() / 3)), dim3(256), 0, s, src,
                               L.d_moves, dst);
        }
        SL_HIP_TRY(hipGetLastError());
    }
    return kOk;
}

// Captures one run (the combing launch and every merge level) into a HIP
// graph on the batch's stream; later runs replay it with one launch.
static int hybrid_capture(Hybrid *h) {
    hipStream_t s = batch_stream_internal(h->batch);
    SL_HIP_TRY(hipStreamSynchronize(s));
    SL_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    int rc = batch_launch_replayable_internal(h->batch);
    if (!rc) rc = hybrid_merges(h);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);   // ends the capture on every path
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return kErrHip;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        return kErrHip;
    }
    h->graph = g;
    h->graph_exec = ex;
    return kOk;
}

static int hybrid_run(Hybrid *h) {
    if (h->use_graph) {
        if (!h->graph_exec) {
            const int rc = hybrid_capture(h);
            if (rc) return rc;
        }
        SL_HIP_TRY(hipGraphLaunch(h->graph_exec, batch_stream_internal(h->batch)));
        return kOk;
    }
    int rc = batch_launch_internal(h->batch);
    if (!rc) rc = hybrid_merges(h);
    return rc;
}

}  // namespace sa
}  // namespace slhip

using namespace slhip;
using namespace slhip::sa;

extern "C" {

int sl_braid_multiply_batch(int nprob, const int32_t *p_cat, const int32_t *q_cat, const int32_t *sizes,
                            int32_t *r_cat) {
    if (nprob < 0 || (nprob > 0 && (!sizes || !p_cat || !q_cat || !r_cat))) {
        set_error("sl_braid_multiply_batch: bad argument");
        return kErrArg;
    }
    for (int i = 0; i < nprob; ++i)
        if (sizes[i] < 0) {
            set_error("sl_braid_multiply_batch: negative size");
            return kErrArg;
        }
    MulHandle *h = nullptr;
    int rc = mul_create(&h, nprob, sizes);
    if (rc) return rc;
    rc = mul_upload(h, p_cat, q_cat);
    if (!rc) rc = mul_launch(h);
    if (!rc) rc = mul_download(h, r_cat);
    mul_free(h);
    return rc;
}

int sl_braid_multiply(const int32_t *p, const int32_t *q, int n, int32_t *r) {
    if (n < 0 || (n > 0 && (!p || !q || !r))) {
        set_error("sl_braid_multiply: bad argument");
        return kErrArg;
    }
    return sl_braid_multiply_batch(1, p, q, &n, r);
}

int sl_braid_multiply_time(int nprob, const int32_t *p_cat, const int32_t *q_cat, const int32_t *sizes, int warmup,
                           int iters, double *ms_per_run, int64_t *elements_levels) {
    if (nprob < 0 || iters <= 0 || warmup < 0 || !ms_per_run || (nprob > 0 && (!sizes || !p_cat || !q_cat))) {
        set_error("sl_braid_multiply_time: bad argument");
        return kErrArg;
    }
    MulHandle *h = nullptr;
    int rc = mul_create(&h, nprob, sizes);
    if (rc) return rc;
    rc = mul_upload(h, p_cat, q_cat);
    for (int i = 0; !rc && i < warmup; ++i) rc = mul_launch(h);
    if (!rc) {
        if (hipEventRecord(h->ev[0], h->stream) != hipSuccess) rc = kErrHip;
        for (int i = 0; !rc && i < iters; ++i) rc = mul_launch(h);
        if (!rc && hipEventRecord(h->ev[1], h->stream) != hipSuccess) rc = kErrHip;
        if (!rc && hipEventSynchronize(h->ev[1]) != hipSuccess) rc = kErrHip;
        float ms = 0.f;
        if (!rc && hipEventElapsedTime(&ms, h->ev[0], h->ev[1]) != hipSuccess) rc = kErrHip;
        if (!rc) *ms_per_run = ms / iters;
        if (rc == kErrHip) set_error("sl_braid_multiply_time: HIP event failure");
    }
    if (!rc && elements_levels) *elements_levels = h->plan->elements_levels;
    mul_free(h);
    return rc;
}

int sl_staggered_multiply(const int32_t *p, int pr, const int32_t *q, int qr, int k, int row_glue, int32_t *prod) {
    if (pr < 0 || qr < 0 || k < 0 || k > pr || k > qr || (pr > 0 && !p) || (qr > 0 && !q) ||
        (pr + qr - k > 0 && !prod)) {
        set_error("sl_staggered_multiply: bad argument");
        return kErrArg;
    }
    if (k == pr && k == qr) {
        set_error("sl_staggered_multiply: k equals both sizes; the reference leaves the product unset for this "
                  "case (steady_ant.h:616-619)");
        return kErrArg;
    }
    std::vector<uint8_t> seen;
    if (!is_permutation(p, pr, seen) || !is_permutation(q, qr, seen)) {
        set_error("sl_staggered_multiply: p and q must be full permutations");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    // one job: pool = [p | q], output pool of pr + qr - k
    MergeLevel L;
    StagJob jb{};
    jb.p_off = 0;
    jb.q_off = pr;
    jb.pr = pr;
    jb.qr = qr;
    jb.k = k;
    jb.row_glue = row_glue ? 1 : 0;
    jb.prob = 0;
    jb.qcr_off = 0;
    jb.out_off = 0;
    jb.tileP0 = 0;
    jb.ntilesP = (int32_t)ceil_div(pr, kTile);
    jb.tileQ0 = jb.ntilesP;
    jb.ntilesQ = (int32_t)ceil_div(qr, kTile);
    std::vector<int32_t> tile_job((size_t)(jb.ntilesP + jb.ntilesQ), 0);
    std::vector<int32_t> pool((size_t)pr + qr);
    if (pr) std::memcpy(pool.data(), p, sizeof(int32_t) * pr);
    if (qr) std::memcpy(pool.data() + pr, q, sizeof(int32_t) * qr);
    MulPlan *plan = nullptr;
    rc = plan_create(&plan, std::vector<int32_t>{k}, true);
    if (rc) return rc;
    int32_t *d_pool = nullptr, *d_out = nullptr, *d_tile_job = nullptr, *d_cnt = nullptr, *d_qcr = nullptr;
    StagJob *d_jobs = nullptr;
    hipStream_t s = nullptr;
    const int N = pr + qr - k;
    auto cleanup = [&]() {
        plan_free(plan);
        (void)hipFree(d_pool);
        (void)hipFree(d_out);
        (void)hipFree(d_tile_job);
        (void)hipFree(d_cnt);
        (void)hipFree(d_qcr);
        (void)hipFree(d_jobs);
        if (s) (void)hipStreamDestroy(s);
    };
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = upload(&d_pool, pool);
    if (e == hipSuccess) e = dalloc(&d_out, (size_t)N);
    if (e == hipSuccess) e = upload(&d_tile_job, tile_job);
    if (e == hipSuccess) e = dalloc(&d_cnt, tile_job.size());
    if (e == hipSuccess) e = dalloc(&d_qcr, (size_t)qr);
    if (e == hipSuccess) e = upload(&d_jobs, std::vector<StagJob>{jb});
    if (e != hipSuccess) {
        set_error(std::string("sl_staggered_multiply: ") + hipGetErrorString(e));
        cleanup();
        return e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;
    }
    const int nt = (int)tile_job.size();
    if (nt > 0) {
        hipLaunchKernelGGL(stag_qinv, dim3(nt), dim3(kTile), 0, s, d_pool, d_jobs, d_tile_job, d_qcr);
        hipLaunchKernelGGL(stag_count, dim3(nt), dim3(kTile), 0, s, d_pool, d_qcr, d_jobs, d_tile_job, d_cnt);
        hipLaunchKernelGGL(stag_scan, dim3(2), dim3(kTile), 0, s, d_jobs, d_cnt);
        hipLaunchKernelGGL(stag_scatter, dim3(nt), dim3(kTile), 0, s, d_pool, d_qcr, d_jobs, d_tile_job, d_cnt,
                           plan->d_probs, plan->d_P[0], plan->d_Qi[0], plan->d_map_row, plan->d_map_col);
    }
    rc = plan_launch(plan, s);
    if (!rc && k + pr + qr > 0)
        hipLaunchKernelGGL(stag_glue, dim3((unsigned)ceil_div(k + pr + qr, 256), 1), dim3(256), 0, s, d_pool, d_jobs,
                           plan->d_probs, plan->d_R[plan->result], plan->d_map_row, plan->d_map_col, d_out);
    if (!rc) {
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e == hipSuccess && N > 0) e = hipMemcpy(prod, d_out, sizeof(int32_t) * N, hipMemcpyDeviceToHost);
        if (e != hipSuccess) {
            set_error(std::string("sl_staggered_multiply: ") + hipGetErrorString(e));
            rc = kErrHip;
        }
    }
    cleanup();
    return rc;
}

int sl_hybrid_create(sl_hybrid **out, const void *a, int m, const void *b, int n, int sym_bytes, int small_m,
                     int small_n) {
    if (!out) {
        set_error("sl_hybrid_create: null handle");
        return kErrArg;
    }
    Hybrid *h = nullptr;
    const int rc = hybrid_create(&h, a, m, b, n, sym_bytes, small_m, small_n);
    *out = reinterpret_cast<sl_hybrid *>(h);
    return rc;
}

int sl_hybrid_run(sl_hybrid *hh) {
    Hybrid *h = reinterpret_cast<Hybrid *>(hh);
    if (!h) {
        set_error("sl_hybrid_run: null handle");
        return kErrArg;
    }
    int rc = hybrid_run(h);
    if (!rc) rc = batch_wait_internal(h->batch);
    return rc;
}

int sl_hybrid_download(sl_hybrid *hh, int32_t *row_to_col) {
    Hybrid *h = reinterpret_cast<Hybrid *>(hh);
    if (!h || (!row_to_col && h->m + h->n > 0)) {
        set_error("sl_hybrid_download: bad argument");
        return kErrArg;
    }
    if (h->m + h->n == 0) return kOk;
    SL_HIP_TRY(hipStreamSynchronize(batch_stream_internal(h->batch)));
    SL_HIP_TRY(hipMemcpy(row_to_col, hybrid_pool(h, h->final_pool) + h->final_off, sizeof(int32_t) * (h->m + h->n),
                         hipMemcpyDeviceToHost));
    return kOk;
}

int sl_hybrid_time(sl_hybrid *hh, int warmup, int iters, double *ms_per_run, double *comb_ms, double *merge_ms) {
    Hybrid *h = reinterpret_cast<Hybrid *>(hh);
    if (!h || iters <= 0 || warmup < 0 || !ms_per_run) {
        set_error("sl_hybrid_time: bad argument");
        return kErrArg;
    }
    hipStream_t s = batch_stream_internal(h->batch);
    for (int i = 0; i < warmup; ++i) {
        int rc = hybrid_run(h);
        if (!rc) rc = batch_wait_internal(h->batch);
        if (rc) return rc;
    }
    double tot = 0, tc = 0, tm = 0;
    if (h->use_graph) {   // one graph launch per run: no split into phases
        SL_HIP_TRY(hipEventRecord(h->ev[0], s));
        for (int i = 0; i < iters; ++i) {
            const int rc = hybrid_run(h);
            if (rc) return rc;
        }
        SL_HIP_TRY(hipEventRecord(h->ev[2], s));
        const int rc = batch_wait_internal(h->batch);
        if (rc) return rc;
        float a = 0;
        SL_HIP_TRY(hipEventElapsedTime(&a, h->ev[0], h->ev[2]));
        *ms_per_run = a / iters;
        if (comb_ms) *comb_ms = -1.0;
        if (merge_ms) *merge_ms = -1.0;
        return kOk;
    }
    for (int i = 0; i < iters; ++i) {
        SL_HIP_TRY(hipEventRecord(h->ev[0], s));
        int rc = batch_launch_internal(h->batch);
        if (rc) return rc;
        SL_HIP_TRY(hipEventRecord(h->ev[1], s));
        rc = hybrid_merges(h);
        if (rc) return rc;
        SL_HIP_TRY(hipEventRecord(h->ev[2], s));
        rc = batch_wait_internal(h->batch);
        if (rc) return rc;
        float a = 0, b = 0;
        SL_HIP_TRY(hipEventElapsedTime(&a, h->ev[0], h->ev[1]));
        SL_HIP_TRY(hipEventElapsedTime(&b, h->ev[1], h->ev[2]));
        tc += a;
        tm += b;
        tot += a + b;
    }
    *ms_per_run = tot / iters;
    if (comb_ms) *comb_ms = tc / iters;
    if (merge_ms) *merge_ms = tm / iters;
    return kOk;
}

int sl_hybrid_stats(const sl_hybrid *hh, int64_t *cells, int *merge_levels) {
    const Hybrid *h = reinterpret_cast<const Hybrid *>(hh);
    if (!h) {
        set_error("sl_hybrid_stats: null handle");
        return kErrArg;
    }
    if (cells) *cells = h->cells;
    if (merge_levels) *merge_levels = (int)h->levels.size();
    return kOk;
}

int sl_hybrid_set_graph(sl_hybrid *hh, int enable) {
    Hybrid *h = reinterpret_cast<Hybrid *>(hh);
    if (!h) {
        set_error("sl_hybrid_set_graph: null handle");
        return kErrArg;
    }
    h->use_graph = enable != 0;
    return kOk;
}

int sl_hybrid_autotune(sl_hybrid *hh, int iters, int *rows_per_lane, int *waves_per_cu) {
    Hybrid *h = reinterpret_cast<Hybrid *>(hh);
    if (!h || iters <= 0) {
        set_error("sl_hybrid_autotune: bad argument");
        return kErrArg;
    }
    SL_HIP_TRY(hipStreamSynchronize(batch_stream_internal(h->batch)));
    // a new shape can move the queue's device buffers: drop the captured graph
    if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph) (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    return sl_batch_autotune(reinterpret_cast<sl_batch *>(h->batch), iters, rows_per_lane, waves_per_cu);
}

void sl_hybrid_destroy(sl_hybrid *hh) { hybrid_free(reinterpret_cast<Hybrid *>(hh)); }

int sl_hybrid_iterative(const void *a, int m, const void *b, int n, int sym_bytes, int small_m, int small_n,
                        int32_t *row_to_col) {
    sl_hybrid *h = nullptr;
    int rc = sl_hybrid_create(&h, a, m, b, n, sym_bytes, small_m, small_n);
    if (!rc) rc = sl_hybrid_run(h);
    if (!rc) rc = sl_hybrid_download(h, row_to_col);
    sl_hybrid_destroy(h);
    return rc;
}

// Block split per length bucket.  The reference sizes the hybrid's blocks
// by a fixed rule (hybrid_iterative_wrapper, semi_local.h:740-790); here the
// grid is measured.  Bucket of a pair: floor(log2(max(m, n))).
static std::mutex g_grid_mu;
static int32_t g_grid[32][2];   // bucket -> (small_m, small_n), 0 x 0: not tuned

static int grid_bucket(int m, int n) {
    int v = std::max(std::max(m, n), 1), b = 0;
    while (v >>= 1) ++b;
    return b;
}

int sl_hybrid_autotune_blocks(const void *a, int m, const void *b, int n, int sym_bytes, int ncand,
                              const int32_t *cand_m, const int32_t *cand_n, int iters, int32_t *best_m,
                              int32_t *best_n, double *ms) {
    if (ncand < 1 || !cand_m || !cand_n || iters < 1 || !best_m || !best_n) {
        set_error("sl_hybrid_autotune_blocks: bad argument");
        return kErrArg;
    }
    for (int i = 0; i < ncand; ++i)
        if (cand_m[i] < 1 || cand_n[i] < 1) {
            set_error("sl_hybrid_autotune_blocks: a candidate grid has fewer than 1 x 1 blocks");
            return kErrArg;
        }
    double best = 1e300;
    int pick = -1;
    for (int i = 0; i < ncand; ++i) {
        sl_hybrid *h = nullptr;
        int rc = sl_hybrid_create(&h, a, m, b, n, sym_bytes, cand_m[i], cand_n[i]);
        double t = 0;
        if (!rc) rc = sl_hybrid_time(h, 1, iters, &t, nullptr, nullptr);
        sl_hybrid_destroy(h);
        if (rc) return rc;
        if (ms) ms[i] = t;
        if (t < best) {
            best = t;
            pick = i;
        }
    }
    *best_m = cand_m[pick];
    *best_n = cand_n[pick];
    std::lock_guard<std::mutex> lk(g_grid_mu);
    g_grid[grid_bucket(m, n)][0] = cand_m[pick];
    g_grid[grid_bucket(m, n)][1] = cand_n[pick];
    return kOk;
}

int sl_hybrid_bucket_blocks(int m, int n, int32_t *small_m, int32_t *small_n) {
    if (m < 0 || n < 0 || !small_m || !small_n) {
        set_error("sl_hybrid_bucket_blocks: bad argument");
        return kErrArg;
    }
    std::lock_guard<std::mutex> lk(g_grid_mu);
    *small_m = g_grid[grid_bucket(m, n)][0];
    *small_n = g_grid[grid_bucket(m, n)][1];
    return kOk;
}

}  // extern "C"
