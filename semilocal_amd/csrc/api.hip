// api.hip -- host side of the C ABI (include/semilocal_hip.h): device and
// stream management, ragged-batch planning (orientation, length-bucket tile
// table, LPT queue order), uploads, launches and timing.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <queue>
#include <unordered_map>
#include <vector>

#include "../../include/semilocal_hip.h"
#include "common.h"

namespace slhip {

hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t stream);
hipError_t launch_comb_batch(int sym_bytes, int rows_per_lane, unsigned waves, hipStream_t stream,
                             const PairDesc *pairs, const WorkItem *items, uint32_t nitems, const void *pool,
                             int32_t *rc, int32_t *cr, uint64_t *buf, uint32_t epoch, LaunchCtl *ctl,
                             uint32_t spin_limit, uint32_t table_words, uint32_t *wait_flags);

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }
const char *last_error() { return g_last_error.c_str(); }

// ---------------------------------------------------------------------------
// Device state
// ---------------------------------------------------------------------------
struct DeviceState {
    bool ready = false;
    int device = 0;
    int cus = 0;
    int clock_khz = 0;
    size_t hbm = 0;
    char name[256] = {0};
    char arch[256] = {0};
};
static DeviceState g_dev;
static std::mutex g_dev_mu;

int ensure_device(int device) {
    std::lock_guard<std::mutex> lk(g_dev_mu);
    if (g_dev.ready && (device < 0 || device == g_dev.device)) {
        SL_HIP_TRY(hipSetDevice(g_dev.device));
        return kOk;
    }
    int count = 0;
    SL_HIP_TRY(hipGetDeviceCount(&count));
    if (count <= 0) {
        set_error("no HIP device visible");
        return kErrHip;
    }
    if (device < 0) device = 0;
    if (device >= count) {
        set_error("device index out of range");
        return kErrArg;
    }
    SL_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    SL_HIP_TRY(hipGetDeviceProperties(&prop, device));
    g_dev.device = device;
    g_dev.cus = prop.multiProcessorCount;
    g_dev.clock_khz = prop.clockRate;
    g_dev.hbm = prop.totalGlobalMem;
    // the marketing name can be empty on some boxes; the ISA name is always set
    std::snprintf(g_dev.arch, sizeof g_dev.arch, "%s", prop.gcnArchName);
    std::snprintf(g_dev.name, sizeof g_dev.name, "%s", prop.name[0] ? prop.name : g_dev.arch);
    g_dev.ready = true;
    return kOk;
}

int device_cus() { return g_dev.cus > 0 ? g_dev.cus : 256; }

// ---------------------------------------------------------------------------
// Alphabet packing: symbols only matter through equality, so any injective
// relabeling gives the same kernel.  uint8 input is used as is; wider input
// is remapped to dense codes (uint8 if <= 256 distinct, else uint16).
// ---------------------------------------------------------------------------
struct PackedPool {
    int sym_bytes = 1;
    std::vector<uint8_t> bytes;   // pool in the chosen width
};

template <typename T>
static int pack_pool_from(const T *a_cat, const int64_t *a_off, const int32_t *a_len, const T *b_cat,
                          const int64_t *b_off, const int32_t *b_len, int npairs, bool identity_u8,
                          std::vector<int64_t> &a_pos, std::vector<int64_t> &b_pos, PackedPool &out) {
    int64_t total = 0;
    for (int i = 0; i < npairs; ++i) total += (int64_t)a_len[i] + b_len[i];
    a_pos.resize(npairs);
    b_pos.resize(npairs);
    if (identity_u8) {
        out.sym_bytes = 1;
        out.bytes.resize((size_t)std::max<int64_t>(total, 1));
        int64_t p = 0;
        for (int i = 0; i < npairs; ++i) {
            a_pos[i] = p;
            std::memcpy(out.bytes.data() + p, (const uint8_t *)a_cat + a_off[i], (size_t)a_len[i]);
            p += a_len[i];
            b_pos[i] = p;
            std::memcpy(out.bytes.data() + p, (const uint8_t *)b_cat + b_off[i], (size_t)b_len[i]);
            p += b_len[i];
        }
        return kOk;
    }
    std::unordered_map<int64_t, uint32_t> code;
    code.reserve(1024);
    std::vector<uint32_t> tmp((size_t)std::max<int64_t>(total, 1));
    int64_t p = 0;
    for (int i = 0; i < npairs; ++i) {
        for (int side = 0; side < 2; ++side) {
            const T *src = side == 0 ? a_cat + a_off[i] : b_cat + b_off[i];
            const int32_t len = side == 0 ? a_len[i] : b_len[i];
            (side == 0 ? a_pos : b_pos)[i] = p;
            for (int32_t k = 0; k < len; ++k) {
                auto it = code.find((int64_t)src[k]);
                uint32_t c;
                if (it == code.end()) {
                    c = (uint32_t)code.size();
                    code.emplace((int64_t)src[k], c);
                } else {
                    c = it->second;
                }
                tmp[(size_t)p++] = c;
            }
        }
    }
    if (code.size() > 65535) {
        set_error("more than 65535 distinct symbols");
        return kErrUnsupported;
    }
    out.sym_bytes = code.size() <= 256 ? 1 : 2;
    out.bytes.resize((size_t)std::max<int64_t>(total, 1) * out.sym_bytes);
    if (out.sym_bytes == 1)
        for (int64_t k = 0; k < total; ++k) out.bytes[(size_t)k] = (uint8_t)tmp[(size_t)k];
    else
        for (int64_t k = 0; k < total; ++k) ((uint16_t *)out.bytes.data())[k] = (uint16_t)tmp[(size_t)k];
    return kOk;
}

// Rewrites each pair's symbols (rows and columns share one map) to dense codes
// 0..sigma_pair-1 in first-seen order (only equality matters, so relabeling
// keeps the kernel) and records sigma and the strand-id bits.
static void remap_pairs_dense(std::vector<PairDesc> &pairs, PackedPool &pool) {
    std::vector<int32_t> table(pool.sym_bytes == 1 ? 256 : 65536, -1);
    std::vector<uint32_t> touched;
    for (auto &d : pairs) {
        touched.clear();
        int32_t sigma = 0;
        auto remap = [&](int64_t off, int32_t len) {
            for (int32_t k = 0; k < len; ++k) {
                uint32_t v = pool.sym_bytes == 1 ? pool.bytes[(size_t)(off + k)]
                                                 : ((uint16_t *)pool.bytes.data())[off + k];
                if (table[v] < 0) {
                    table[v] = sigma++;
                    touched.push_back(v);
                }
                if (pool.sym_bytes == 1) pool.bytes[(size_t)(off + k)] = (uint8_t)table[v];
                else ((uint16_t *)pool.bytes.data())[off + k] = (uint16_t)table[v];
            }
        };
        remap(d.rows_off, d.m);
        remap(d.cols_off, d.n);
        for (uint32_t v : touched) table[v] = -1;
        const uint32_t max_id = (uint32_t)(d.m + d.n > 1 ? d.m + d.n - 1 : 1);
        int bits = 0;
        while (bits < 32 && (max_id >> bits) != 0) ++bits;
        d.tag_shift = bits;
        d.sigma = sigma;
        d.cell = kCellCode;
    }
}

// ---------------------------------------------------------------------------
// Tile height (rows per lane, R) for a whole batch.  One pass = 64R rows; the
// passes of a pair run concurrently, each ~2.5 chunks (160 steps) behind the
// previous one.  Larger R costs less per cell (one DPP move and three LDS
// accesses per step are shared by R cells) but lengthens the critical path of
// the largest pair (fewer, longer passes).  The default minimises
// max(throughput time, critical path) with per-R rates measured on MI355X;
// the critical path's step time counts the waves that share a SIMD (the
// batch's passes per SIMD, 1 to 3), so one large pair with a single wave per
// SIMD gets tall tiles.  sl_batch_autotune replaces the model with a
// measurement on the batch itself.
// ---------------------------------------------------------------------------
static const int kTileHeights[3] = {4, 8, 16};

// Per-pair tile model (assign_tiles).  Instructions per step beside the R
// cells: the DPP hand-off, the LDS write of the outgoing strand and the
// block's LDS reads (measured on MI355X: 3.66 VALU lane-ops per cell at R = 8,
// 6.7 % of them padding rows).
constexpr double kStepOverheadOps = 3.4;
constexpr double kValuEfficiency = 0.75;     // VALU busy a full batch sustains
constexpr double kBulkCyclesPerOp = 8.0;     // two waves share a SIMD, 4 issue cycles each
constexpr double kCriticalCyclesPerOp = 6.5; // a wave at raised priority: its own issue rate
constexpr double kDefaultCriticalFraction = 0.6;
constexpr double kCriticalLag = 256.0;       // queue lag (steps) between passes of a critical chain

// `passes_at(R)`: the batch's total passes at tile height R.
template <typename PassesAt>
static int model_rows_core(PassesAt passes_at, int64_t total_cells, int32_t m_max, int32_t n_max, int cus) {
    const double rate[3] = {6.2e12, 7.2e12, 8.0e12};   // single-pass cells/s on 256 CUs
    const double scale = cus > 0 ? cus / 256.0 : 1.0;
    const double simds = 4.0 * (cus > 0 ? cus : 256);
    int best = 16;
    double best_t = 1e30;
    for (int k = 2; k >= 0; --k) {
        const int R = kTileHeights[k];
        const double total_passes = passes_at(R);
        const double sharing = std::min(3.0, std::max(1.0, total_passes / simds));
        const double passes = std::ceil(m_max / (64.0 * R));
        const double steps = n_max + passes * 160.0;
        const double clk_per_step = (R * 4.1 + 2.0) * 4.0 * sharing;
        const double t_crit = steps * clk_per_step / 2.4e9;
        const double t_thru = (double)total_cells / (rate[k] * scale);
        const double t = std::max(t_crit, t_thru);
        if (t < best_t * 0.97) {   // a smaller R must be clearly faster
            best_t = t;
            best = R;
        }
    }
    return best;
}

int model_rows_per_lane(const std::vector<PairDesc> &pairs, int64_t total_cells, int32_t m_max, int32_t n_max,
                        int cus) {
    auto passes_at = [&](int R) {
        double total = 0.0;
        for (const auto &d : pairs)
            if (d.m > 0 && d.n > 0) total += std::ceil(d.m / (64.0 * R));
        return total;
    };
    return model_rows_core(passes_at, total_cells, m_max, n_max, cus);
}

// The same model for one oriented pair (m <= n), with no allocation.
int model_rows_one_pair(int32_t m, int32_t n, int cus) {
    auto passes_at = [&](int R) { return std::ceil(m / (64.0 * R)); };
    return model_rows_core(passes_at, (int64_t)m * n, m, n, cus);
}

// ---------------------------------------------------------------------------
// Batch
// ---------------------------------------------------------------------------
struct Batch {
    int npairs = 0;
    int sym_bytes = 1;
    bool with_cr = false;
    std::vector<PairDesc> pairs;       // caller order
    std::vector<WorkItem> items;       // passes pair by pair (largest pairs first)
    std::vector<WorkItem> order;       // ticket order of the batched launch
    int schedule = 2;                  // 0: `items` as they are; s >= 1: simulated list schedule, lag 128 s (schedule_items)
    std::vector<int64_t> perm_off;     // caller order
    int64_t perm_total = 0;
    int64_t buf_total = 0;
    int64_t cells = 0;
    uint32_t spin_limit = 1u << 23;
    int rows_per_lane = 0;             // tile policy: 4, 8 or 16 for every pair, 0 = per pair (assign_tiles)
    double critical_fraction = kDefaultCriticalFraction;   // per-pair policy: see assign_tiles
    int launch_rows = 16;              // tile height the kernel is instantiated for (0: per pair)
    unsigned waves_per_cu = 8;         // with schedule 2: the best default shape measured on MI355X
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    PairDesc *d_pairs = nullptr;
    WorkItem *d_items = nullptr;
    uint32_t *d_flags = nullptr;       // wait flags, one word per entry of d_items (see comb_pass)
    void *d_pool = nullptr;
    int32_t *d_rc = nullptr;
    int32_t *d_cr = nullptr;
    uint64_t *d_buf = nullptr;         // tagged inter-pass strands, one entry per column of each pair
    uint32_t epoch = 0;                // this run's tag (next_run_tag); part of the strand tags
    uint64_t cleared_at = 0;           // run serial when d_buf was last cleared
    LaunchCtl *d_ctl = nullptr;    // [0] batched launch, [1+i] per-pair launch of pair i
    unsigned waves_cap = 0;
    int cus = 0;                       // CUs the queue is planned for; 0: the device's
    uint32_t table_words = 0;          // LDS words per wave of the row-match table (kCellTable pairs)
    size_t items_cap = 0;
    DeviceArena *arena = nullptr;      // null: hipMalloc / hipFree
    bool own_stream = true;
    // Host sources of the stream-ordered uploads: kept for the batch's life
    // (the copies run on the batch's stream, after the host returns).
    std::vector<uint8_t> host_pool;
    std::vector<int32_t> host_ident;
};

static hipError_t bt_alloc(Batch *bt, void **p, size_t bytes) {
    *p = nullptr;
    return bt->arena ? bt->arena->alloc(p, bytes) : hipMalloc(p, bytes);
}

static void bt_free(Batch *bt, void *p) {
    if (!p) return;
    if (bt->arena) bt->arena->release(p);
    else (void)hipFree(p);
}

static void batch_free(Batch *b) {
    if (!b) return;
    bt_free(b, b->d_pairs);
    bt_free(b, b->d_items);
    bt_free(b, b->d_flags);
    bt_free(b, b->d_pool);
    bt_free(b, b->d_rc);
    bt_free(b, b->d_cr);
    bt_free(b, b->d_buf);
    bt_free(b, b->d_ctl);
    for (auto &e : b->ev)
        if (e) (void)hipEventDestroy(e);
    if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

constexpr int kMaxSchedule = 8;

// Queue order of the persistent launch.  Waves take tickets in order, and a
// pass can only run some way behind its predecessor (it reads the strands
// the previous pass leaves, at least 128 steps behind), so handing out the
// ~200 passes of a large pair back to back would park hundreds of waves on
// waits.  The order comes from a list-scheduling simulation instead: W
// waves; a pass lasts n + 64 steps, each step costing in proportion to the
// instructions of its tile height; pass p of a pair becomes ready `lag`
// steps after pass p-1 started; a free wave takes the ready pass of the
// pair with the longest remaining chain (its remaining passes' lags plus one
// pass: critical-path list scheduling, so a skinny pair of one long pass is
// not left for the end of the launch).  Passes of a pair keep their order, so
// a pass only ever waits on a pass handed out before it.  A longer lag interleaves more pairs
// and leaves a pass more slack behind its predecessor (a pass that catches
// up with a slower predecessor waits), at the price of a longer chain per
// pair: priority pairs (the critical chains) keep the short kCriticalLag,
// the bulk takes `lag` (128 x schedule, chosen by the autotuner).

// VALU instructions per cell of the pair's symbol test at tile height R
// (choose_cells makes the same choice).
static int cell_ops(const PairDesc &d, int R) {
    if (d.tag_shift + d.sigma <= 32) return 3;                                      // one-hot tags
    if (d.sigma <= kTableMaxSymbols && d.tag_shift <= 32 - R) return 3;             // row-match table
    return 4;                                                                       // symbol codes
}
static int64_t passes_at(const PairDesc &d, int R) { return ((int64_t)d.m + 64LL * R - 1) / (64LL * R); }
static double step_ops(const PairDesc &d, int R) { return cell_ops(d, R) * (double)R + kStepOverheadOps; }
// Tile height of the last pass at main height R: the smallest of 4, 8, 16
// (no larger than R) whose 64 x height rows cover the rows left for it.
static int last_height(const PairDesc &d, int R) {
    const int64_t rest = (int64_t)d.m - (passes_at(d, R) - 1) * 64LL * R;
    for (int h : kTileHeights)
        if (h <= R && 64LL * h >= rest) return h;
    return R;
}
// Instructions of the pair's passes at main height R, the last at its own.
static double pair_ops(const PairDesc &d, int R) {
    return ((double)(passes_at(d, R) - 1) * step_ops(d, R) + step_ops(d, last_height(d, R))) * ((double)d.n + 63.0);
}
// Cycles from the start of the pair's first pass to the end of its last: the
// last pass starts (P - 1) lags after the first and runs n + 64 steps.
static double chain_cycles(const PairDesc &d, int R, double lag, double cycles_per_op) {
    return ((double)d.n + 64.0 + (double)(passes_at(d, R) - 1) * lag) * step_ops(d, R) * cycles_per_op;
}

// Tile height (d.rows_per_lane) and priority (d.prio) of every pair.
//  * Policy 4, 8 or 16: every pair at that height, no priority.
//  * Policy 0, per pair (the default):
//    - a batch whose passes at R = 16 cannot give every SIMD two waves is
//      latency-bound: one height for all from model_rows_per_lane;
//    - otherwise each pair takes the height with the fewest instructions for
//      its padded rows: ceil(m / 64R) passes x (n + 63) steps x step_ops,
//      the last pass at the smallest height that covers its rows (a pair of
//      1,100 rows at R = 16: a pass of 1,024 rows and one of 256, not two of
//      1,024), and the batch's modelled throughput time is those
//      instructions x 4 cycles / (SIMDs x kValuEfficiency);
//    - a pair whose chain of passes at that height and the bulk lag would
//      take longer than critical_fraction x that time is a critical chain:
//      it takes the height with the shortest chain at kCriticalLag (short
//      passes: R = 4 for the largest pairs, whatever it costs in
//      instructions) and runs at raised wave priority.  critical_fraction 0
//      marks none.
// The reference sizes its blocks for the same two forces -- enough parallel
// work, few row pieces -- in hybrid_iterative_wrapper (semi_local.h:740-790).
static void assign_tiles(Batch *bt) {
    for (auto &d : bt->pairs) d.prio = 0;
    if (bt->rows_per_lane != 0) {
        for (auto &d : bt->pairs) d.rows_per_lane = bt->rows_per_lane;
        bt->launch_rows = bt->rows_per_lane;
        return;
    }
    const int cus = bt->cus > 0 ? bt->cus : device_cus();
    const double simds = 4.0 * std::max(cus, 1);
    double passes16 = 0.0;
    int32_t m_max = 0, n_max = 0;
    for (const auto &d : bt->pairs) {
        if (d.m <= 0 || d.n <= 0) continue;
        passes16 += (double)passes_at(d, 16);
        m_max = std::max(m_max, d.m);
        n_max = std::max(n_max, d.n);
    }
    if (passes16 < 2.0 * simds) {
        const int R = model_rows_per_lane(bt->pairs, bt->cells, m_max, n_max, cus);
        for (auto &d : bt->pairs) d.rows_per_lane = R;
        bt->launch_rows = R;
        return;
    }
    double instr = 0.0;
    for (auto &d : bt->pairs) {
        d.rows_per_lane = 16;
        if (d.m <= 0 || d.n <= 0) continue;
        double best = 1e300;
        for (int R : kTileHeights) {
            const double c = pair_ops(d, R);
            if (c < best) {
                best = c;
                d.rows_per_lane = R;
            }
        }
        instr += best;
    }
    const double t_thru = instr * 4.0 / (simds * kValuEfficiency);
    const double lag = 128.0 * std::max(bt->schedule, 1);
    if (bt->critical_fraction > 0.0) {
        for (auto &d : bt->pairs) {
            if (d.m <= 0 || d.n <= 0) continue;
            if (chain_cycles(d, d.rows_per_lane, lag, kBulkCyclesPerOp) <= bt->critical_fraction * t_thru) continue;
            double best = 1e300;
            for (int R : kTileHeights) {
                const double c = chain_cycles(d, R, kCriticalLag, kCriticalCyclesPerOp);
                if (c < best) {
                    best = c;
                    d.rows_per_lane = R;
                }
            }
            d.prio = 1;
        }
    }
    int common = -1;
    for (const auto &d : bt->pairs) {
        if (d.m <= 0 || d.n <= 0) continue;
        if (common == -1) common = d.rows_per_lane;
        else if (common != d.rows_per_lane) common = 0;
    }
    bt->launch_rows = common == -1 ? 16 : common;
    if (bt->launch_rows != 0)
        for (auto &d : bt->pairs) d.rows_per_lane = bt->launch_rows;
}

static double step_cost(const PairDesc &d) {   // relative to a tile height of 8
    return (3.0 * d.rows_per_lane + kStepOverheadOps) / (3.0 * 8 + kStepOverheadOps);
}

static void schedule_items(Batch *bt, double lag) {
    const size_t nitems = bt->items.size();
    bt->order.clear();
    bt->order.reserve(nitems);
    if (nitems == 0) return;
    struct Ready {
        double chain;   // steps (at tile height 8) from the pass's start to the end of the pair's last pass
        int pair;
        bool operator<(const Ready &o) const { return chain < o.chain || (chain == o.chain && pair > o.pair); }
    };
    auto chain = [&](int i, int done) {
        const PairDesc &d = bt->pairs[(size_t)i];
        const double rest = (double)(d.npasses - done - 1);
        return ((double)d.n + 64.0 + rest * (d.prio ? kCriticalLag : lag)) * step_cost(d);
    };
    using Pending = std::pair<double, int>;   // (ready time, pair)
    std::priority_queue<Ready> ready;
    std::priority_queue<Pending, std::vector<Pending>, std::greater<Pending>> pending;
    std::priority_queue<double, std::vector<double>, std::greater<double>> waves;
    std::vector<int> next((size_t)bt->npairs, 0);
    for (int i = 0; i < bt->npairs; ++i)
        if (bt->pairs[i].npasses > 0) ready.push(Ready{chain(i, 0), i});
    const size_t w = std::max<size_t>(1, std::min<size_t>(nitems, bt->waves_cap));
    for (size_t k = 0; k < w; ++k) waves.push(0.0);
    while (bt->order.size() < nitems) {
        double t = waves.top();
        waves.pop();
        if (ready.empty() && !pending.empty() && pending.top().first > t) t = pending.top().first;
        while (!pending.empty() && pending.top().first <= t) {
            const int i = pending.top().second;
            pending.pop();
            ready.push(Ready{chain(i, next[i]), i});
        }
        const Ready r = ready.top();
        ready.pop();
        const PairDesc &d = bt->pairs[r.pair];
        bt->order.push_back(WorkItem{r.pair, next[r.pair]});
        const double sc = step_cost(d);
        waves.push(t + ((double)d.n + 64.0) * sc);
        if (++next[r.pair] < d.npasses) pending.push(Pending{t + (d.prio ? kCriticalLag : lag) * sc, r.pair});
    }
}

// Symbol test of each pair's cell for its tile height R: one-hot
// tags when bits(m+n-1) + sigma <= 32; else the per-lane row-match table when
// sigma <= kTableMaxSymbols and the R mask bits fit above the ids
// (bits(m+n-1) <= 32 - R); else symbol codes.  The table takes
// (sigma + 1) x 64 LDS words per wave, sized for the batch's largest sigma.
static void choose_cells(Batch *bt) {
    int32_t table_sigma = 0;
    for (auto &d : bt->pairs) {
        if (d.tag_shift + d.sigma <= 32) {
            d.cell = kCellOneHot;
        } else if (d.sigma <= kTableMaxSymbols && d.tag_shift <= 32 - d.rows_per_lane) {
            d.cell = kCellTable;
            table_sigma = std::max(table_sigma, d.sigma);
        } else {
            d.cell = kCellCode;
        }
    }
    bt->table_words = table_sigma ? (uint32_t)(table_sigma + 1) * 64u : 0u;
}

// Splits every pair into passes of 64R rows at its tile height (assign_tiles).
// `items` lists the passes pair by pair (largest pairs first; the
// one-launch-per-pair form launches a pair's consecutive run), `order` is the
// ticket order of the batched launch: `items` itself
// (longest-processing-time first) or the simulated schedule; the autotuner
// measures both.
static void plan_items(Batch *bt) {
    assign_tiles(bt);
    std::vector<int> order(bt->npairs);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return (int64_t)bt->pairs[x].m * bt->pairs[x].n > (int64_t)bt->pairs[y].m * bt->pairs[y].n;
    });
    bt->items.clear();
    for (int i : order) {
        PairDesc &d = bt->pairs[i];
        const int64_t rpp = 64LL * d.rows_per_lane;
        d.npasses = (d.m > 0 && d.n > 0) ? (int32_t)((d.m + rpp - 1) / rpp) : 0;
        // a launch instantiated for one height runs every pass at it
        d.last_rows = bt->launch_rows != 0 || d.npasses == 0 ? d.rows_per_lane : last_height(d, d.rows_per_lane);
        d.item0 = (int32_t)bt->items.size();
        for (int p = 0; p < d.npasses; ++p) bt->items.push_back(WorkItem{i, p});
    }
    bt->waves_cap = (unsigned)(bt->cus > 0 ? bt->cus : device_cus()) * bt->waves_per_cu;
    choose_cells(bt);
    if (bt->schedule >= 1) schedule_items(bt, 128.0 * bt->schedule);
    else bt->order = bt->items;
}

// (Re)uploads the pair table and the queue (device copy: `items`, then
// `order`); grows the device copies as needed.  Every upload of a batch runs
// on the batch's own stream, so it is ordered before the batch's next launch
// (copies through the null stream are not ordered with a non-blocking
// stream).  The host vectors are only rewritten after a sync of that stream.
static hipError_t upload_items(Batch *bt) {
    hipError_t e = hipSuccess;
    if (!bt->d_pairs) e = bt_alloc(bt, (void **)&bt->d_pairs, sizeof(PairDesc) * std::max<size_t>(bt->pairs.size(), 1));
    if (e == hipSuccess && !bt->pairs.empty())
        e = hipMemcpyAsync(bt->d_pairs, bt->pairs.data(), sizeof(PairDesc) * bt->pairs.size(), hipMemcpyHostToDevice,
                           bt->stream);
    // sized once for the most passes any tile height gives (R = 4), so a
    // change of tile shape never reallocates the queue
    size_t most = 0;
    for (const auto &d : bt->pairs)
        if (d.m > 0 && d.n > 0) most += (size_t)((d.m + 255) / 256);
    const size_t need = 2 * std::max(bt->items.size(), most);
    if (e == hipSuccess && (!bt->d_items || need > bt->items_cap)) {
        bt_free(bt, bt->d_items);
        bt_free(bt, bt->d_flags);
        bt->d_items = nullptr;
        bt->d_flags = nullptr;
        bt->items_cap = std::max<size_t>(need, 1);
        e = bt_alloc(bt, (void **)&bt->d_items, sizeof(WorkItem) * bt->items_cap);
        if (e == hipSuccess) e = bt_alloc(bt, (void **)&bt->d_flags, sizeof(uint32_t) * bt->items_cap);
        // new memory may hold any word: cleared before the first launch
        if (e == hipSuccess) e = launch_fill_u32(bt->d_flags, 0u, bt->items_cap, bt->stream);
    }
    if (e == hipSuccess && !bt->items.empty())
        e = hipMemcpyAsync(bt->d_items, bt->items.data(), sizeof(WorkItem) * bt->items.size(), hipMemcpyHostToDevice,
                           bt->stream);
    if (e == hipSuccess && !bt->order.empty())
        e = hipMemcpyAsync(bt->d_items + bt->items.size(), bt->order.data(), sizeof(WorkItem) * bt->order.size(),
                           hipMemcpyHostToDevice, bt->stream);
    return e;
}

// Allocates and uploads on the batch's stream; `src` must outlive the copy.
template <typename T>
static hipError_t upload(Batch *bt, T **dst, const T *src, size_t count) {
    hipError_t e = bt_alloc(bt, (void **)dst, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess) return e;
    if (count) e = hipMemcpyAsync(*dst, src, sizeof(T) * count, hipMemcpyHostToDevice, bt->stream);
    return e;
}

static int batch_create(Batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                        const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes, int flags,
                        int rows_per_lane, DeviceArena *arena = nullptr, hipStream_t ext_stream = nullptr) {
    if (!out || npairs < 0 || (npairs > 0 && (!a_cat || !a_off || !a_len || !b_cat || !b_off || !b_len))) {
        set_error("sl_batch_create: null argument");
        return kErrArg;
    }
    if (sym_bytes != 1 && sym_bytes != 2 && sym_bytes != 4) {
        set_error("sym_bytes must be 1, 2 or 4");
        return kErrArg;
    }
    if (rows_per_lane != 0 && rows_per_lane != 4 && rows_per_lane != 8 && rows_per_lane != 16) {
        set_error("rows_per_lane must be 0 (per pair), 4, 8 or 16");
        return kErrArg;
    }
    for (int i = 0; i < npairs; ++i) {
        if (a_len[i] < 0 || b_len[i] < 0 || a_off[i] < 0 || b_off[i] < 0) {
            set_error("negative length or offset");
            return kErrArg;
        }
        if ((int64_t)a_len[i] + b_len[i] > 0x7fffffffLL) {
            set_error("pair too large: m + n must fit int32");
            return kErrUnsupported;
        }
    }
    int rc = ensure_device();
    if (rc) return rc;

    Batch *bt = new Batch();
    bt->npairs = npairs;
    bt->with_cr = (flags & SL_BATCH_WITH_COL_TO_ROW) != 0;
    bt->arena = arena;

    std::vector<int64_t> a_pos, b_pos;
    PackedPool pool;
    if (sym_bytes == 1)
        rc = pack_pool_from<uint8_t>((const uint8_t *)a_cat, a_off, a_len, (const uint8_t *)b_cat, b_off, b_len,
                                     npairs, true, a_pos, b_pos, pool);
    else if (sym_bytes == 2)
        rc = pack_pool_from<uint16_t>((const uint16_t *)a_cat, a_off, a_len, (const uint16_t *)b_cat, b_off, b_len,
                                      npairs, false, a_pos, b_pos, pool);
    else
        rc = pack_pool_from<int32_t>((const int32_t *)a_cat, a_off, a_len, (const int32_t *)b_cat, b_off, b_len,
                                     npairs, false, a_pos, b_pos, pool);
    if (rc) {
        delete bt;
        return rc;
    }
    bt->sym_bytes = pool.sym_bytes;

    // orientation and output/buffer offsets (caller order)
    bt->pairs.resize(npairs);
    bt->perm_off.resize(npairs);
    for (int i = 0; i < npairs; ++i) {
        PairDesc &d = bt->pairs[i];
        const bool rev = a_len[i] > b_len[i];
        d.m = rev ? b_len[i] : a_len[i];
        d.n = rev ? a_len[i] : b_len[i];
        d.reverse = rev ? 1 : 0;
        d.rows_off = rev ? b_pos[i] : a_pos[i];
        d.cols_off = rev ? a_pos[i] : b_pos[i];
        d.perm_off = bt->perm_total;
        bt->perm_off[i] = bt->perm_total;
        bt->perm_total += (int64_t)d.m + d.n;
        bt->cells += (int64_t)d.m * d.n;
        d.buf_off = bt->buf_total;   // strand buffer for multi-pass tiling (sized for any R)
        bt->buf_total += d.n;
    }
    remap_pairs_dense(bt->pairs, pool);
    bt->rows_per_lane = rows_per_lane;   // 0: per pair (assign_tiles)
    plan_items(bt);

    hipError_t e = hipSuccess;
    if (ext_stream) {
        bt->stream = ext_stream;
        bt->own_stream = false;
    } else {
        e = hipStreamCreateWithFlags(&bt->stream, hipStreamNonBlocking);
    }
    for (int k = 0; k < 4 && e == hipSuccess; ++k) e = hipEventCreate(&bt->ev[k]);
    if (e == hipSuccess) e = upload_items(bt);
    bt->host_pool = std::move(pool.bytes);
    if (e == hipSuccess) e = upload(bt, (uint8_t **)&bt->d_pool, bt->host_pool.data(), bt->host_pool.size());
    if (e == hipSuccess)
        e = bt_alloc(bt, (void **)&bt->d_rc, sizeof(int32_t) * std::max<int64_t>(bt->perm_total, 1));
    if (e == hipSuccess && bt->with_cr)
        e = bt_alloc(bt, (void **)&bt->d_cr, sizeof(int32_t) * std::max<int64_t>(bt->perm_total, 1));
    if (e == hipSuccess)
        e = bt_alloc(bt, (void **)&bt->d_buf, sizeof(uint64_t) * std::max<int64_t>(bt->buf_total, 1));
    // cleared on the batch's own stream, so the clear is ordered before every
    // combing launch (a null-stream memset is not ordered with a non-blocking
    // stream, and the memory may hold another batch's tagged entries)
    if (e == hipSuccess)
        e = launch_fill_u32((uint32_t *)bt->d_buf, 0u, 2u * (uint64_t)std::max<int64_t>(bt->buf_total, 1), bt->stream);
    bt->cleared_at = run_serial_now();
    if (e == hipSuccess) e = bt_alloc(bt, (void **)&bt->d_ctl, sizeof(LaunchCtl) * (size_t)(npairs + 1));
    // the control blocks hold the error bits batch_sync reports: cleared
    // here, so a sync before the first run reads no memory left by others
    if (e == hipSuccess)
        e = launch_fill_u32((uint32_t *)bt->d_ctl, 0u, (uint64_t)(npairs + 1) * 2u, bt->stream);
    if (e != hipSuccess) {
        set_error(std::string("sl_batch_create: ") + hipGetErrorString(e));
        batch_free(bt);
        return e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;
    }
    // degenerate pairs (m == 0 or n == 0) have no tiles: their kernel is the
    // identity, copied from one host array that lives as long as the batch
    int64_t ident_len = 0;
    for (const auto &d : bt->pairs)
        if (d.npasses == 0) ident_len = std::max<int64_t>(ident_len, (int64_t)d.m + d.n);
    bt->host_ident.resize((size_t)ident_len);
    std::iota(bt->host_ident.begin(), bt->host_ident.end(), 0);
    for (int i = 0; i < npairs && e == hipSuccess; ++i) {
        const PairDesc &d = bt->pairs[i];
        if (d.npasses == 0 && d.m + d.n > 0) {
            const size_t bytes = sizeof(int32_t) * (size_t)(d.m + d.n);
            e = hipMemcpyAsync(bt->d_rc + d.perm_off, bt->host_ident.data(), bytes, hipMemcpyHostToDevice, bt->stream);
            if (e == hipSuccess && bt->with_cr)
                e = hipMemcpyAsync(bt->d_cr + d.perm_off, bt->host_ident.data(), bytes, hipMemcpyHostToDevice,
                                   bt->stream);
        }
    }
    // a batch on its own stream is complete when created; one on a caller's
    // stream is ordered on that stream
    if (e == hipSuccess && bt->own_stream) e = hipStreamSynchronize(bt->stream);
    if (e != hipSuccess) {
        set_error(std::string("sl_batch_create: ") + hipGetErrorString(e));
        batch_free(bt);
        return kErrHip;
    }
    *out = bt;
    return kOk;
}

// Host-side guard: every device buffer a launch touches must exist.
static int batch_check_ready(const Batch *bt) {
    if (!bt->d_pairs || !bt->d_items || !bt->d_flags || !bt->d_pool || !bt->d_rc || !bt->d_buf || !bt->d_ctl ||
        (bt->with_cr && !bt->d_cr) || !bt->stream) {
        set_error("batch is missing a device buffer (internal defect); nothing launched");
        return kErrHip;
    }
    return kOk;
}

// Run tags come from one serial for the whole process, so no two runs of any
// two handles within 65,535 runs of each other tag their hand-offs alike;
// memory a handle inherits from another is cleared at creation.  A handle
// whose buffer was last cleared 65,535 or more runs ago (a resident batch in
// a long-lived process) clears it again before its next run, so a stale entry
// can never carry the tag a consumer waits for.  Tag 0 is a cleared entry.
static std::atomic<uint64_t> g_run_serial{0};
uint64_t run_serial_now() { return g_run_serial.load(); }
bool next_run_tag(uint64_t &cleared_at, uint32_t *tag) {
    const uint64_t s = g_run_serial.fetch_add(1u) + 1u;
    *tag = (uint32_t)(s % 65535u) + 1u;
    if (s - cleared_at >= 65535u) {
        cleared_at = s;
        return true;
    }
    return false;
}

static int batch_reset(Batch *bt) {
    const int rc = batch_check_ready(bt);
    if (rc) return rc;
    if (next_run_tag(bt->cleared_at, &bt->epoch)) {   // new tags for this run's strand hand-offs
        SL_HIP_TRY(launch_fill_u32((uint32_t *)bt->d_buf, 0u, (uint64_t)std::max<int64_t>(bt->buf_total, 1) * 2u,
                                   bt->stream));
        SL_HIP_TRY(launch_fill_u32(bt->d_flags, 0u, bt->items_cap, bt->stream));
    }
    SL_HIP_TRY(launch_fill_u32((uint32_t *)bt->d_ctl, 0u, (uint64_t)(bt->npairs + 1) * 2u, bt->stream));
    return kOk;
}

// One persistent launch over `nitems` queue entries.  A launch with fewer
// waves than CUs (one pair's passes, a small batch) is latency-bound: it gives
// kCellTable pairs the code cell, whose steps carry one LDS read fewer.
static hipError_t launch_items(Batch *bt, const WorkItem *items, uint32_t nitems, LaunchCtl *ctl) {
    const unsigned waves = std::min<unsigned>(nitems, bt->waves_cap);
    const void *pool = bt->d_pool;
    const uint32_t table_words = waves >= (unsigned)device_cus() ? bt->table_words : 0u;
    // per-pair pointers are resolved inside the kernel from PairDesc offsets
    return launch_comb_batch(bt->sym_bytes, bt->launch_rows, waves, bt->stream, bt->d_pairs, items, nitems, pool,
                             bt->d_rc, bt->d_cr, bt->d_buf, bt->epoch, ctl, bt->spin_limit, table_words,
                             bt->d_flags);
}

static int batch_run(Batch *bt) {
    int rc = batch_reset(bt);
    if (rc) return rc;
    SL_HIP_TRY(launch_items(bt, bt->d_items + bt->items.size(), (uint32_t)bt->order.size(), bt->d_ctl));
    return kOk;
}

// Launch meant for replay from a captured HIP graph.  The epoch is captured
// into the kernel arguments and stays fixed, so the strand buffer is cleared
// on the stream first (a fill kernel, not a memset node): a tag written by an
// earlier replay can then never be taken for this replay's.
static int batch_run_replayable(Batch *bt) {
    const int rc = batch_check_ready(bt);
    if (rc) return rc;
    if (bt->epoch == 0) (void)next_run_tag(bt->cleared_at, &bt->epoch);
    SL_HIP_TRY(launch_fill_u32((uint32_t *)bt->d_buf, 0u, (uint64_t)bt->buf_total * 2u, bt->stream));
    SL_HIP_TRY(launch_fill_u32(bt->d_flags, 0u, bt->items_cap, bt->stream));
    bt->cleared_at = run_serial_now();
    SL_HIP_TRY(launch_fill_u32((uint32_t *)bt->d_ctl, 0u, (uint64_t)(bt->npairs + 1) * 2u, bt->stream));
    SL_HIP_TRY(launch_items(bt, bt->d_items + bt->items.size(), (uint32_t)bt->order.size(), bt->d_ctl));
    return kOk;
}

static int batch_run_per_pair(Batch *bt) {
    int rc = batch_reset(bt);
    if (rc) return rc;
    for (int i = 0; i < bt->npairs; ++i) {
        const PairDesc &d = bt->pairs[i];
        if (d.npasses == 0) continue;
        SL_HIP_TRY(launch_items(bt, bt->d_items + d.item0, (uint32_t)d.npasses, bt->d_ctl + 1 + i));
    }
    return kOk;
}

static int batch_sync(Batch *bt) {
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    std::vector<LaunchCtl> ctl((size_t)bt->npairs + 1);
    SL_HIP_TRY(hipMemcpy(ctl.data(), bt->d_ctl, sizeof(LaunchCtl) * ctl.size(), hipMemcpyDeviceToHost));
    uint32_t err = 0;
    for (const auto &c : ctl) err |= c.error;
    if (err & 2u) {
        set_error("combing kernel: strand id out of range (internal defect); results are void");
        return kErrHip;
    }
    if (err & 1u) {
        set_error("combing kernel: a bounded wait timed out; results are void");
        return kErrTimeout;
    }
    return kOk;
}

// Internal interface used by the hybrid scheduler (hybrid.hip).
int batch_create_internal(Batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                          const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes) {
    return batch_create(out, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes, 0, 0);
}
int batch_create_arena(Batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                       const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes,
                       DeviceArena *arena, hipStream_t stream) {
    return batch_create(out, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes, 0, 0, arena, stream);
}
int batch_create_flags_internal(Batch **out, int npairs, const void *a_cat, const int64_t *a_off,
                                const int32_t *a_len, const void *b_cat, const int64_t *b_off, const int32_t *b_len,
                                int sym_bytes, int flags) {
    return batch_create(out, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes, flags, 0);
}
int batch_launch_internal(Batch *bt) { return batch_run(bt); }
int batch_launch_replayable_internal(Batch *bt) { return batch_run_replayable(bt); }
int batch_wait_internal(Batch *bt) { return batch_sync(bt); }
void batch_destroy_internal(Batch *bt) { batch_free(bt); }
hipStream_t batch_stream_internal(const Batch *bt) { return bt->stream; }
const int32_t *batch_rc_internal(const Batch *bt) { return bt->d_rc; }
const int32_t *batch_cr_internal(const Batch *bt) { return bt->d_cr; }
int64_t batch_perm_off_internal(const Batch *bt, int pair) { return bt->perm_off[(size_t)pair]; }
int64_t batch_cells_internal(const Batch *bt) { return bt->cells; }

}  // namespace slhip

using namespace slhip;

extern "C" {

int sl_version(void) { return 100; }

const char *sl_last_error(void) { return last_error(); }

int sl_init(int device) { return ensure_device(device); }

int sl_device_info(char *buf, int len) {
    if (!buf || len <= 0) {
        set_error("sl_device_info: null buffer");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    std::snprintf(buf, (size_t)len, "%s;%d;%d;%zu;%s", g_dev.name, g_dev.cus, g_dev.clock_khz, g_dev.hbm, g_dev.arch);
    return kOk;
}

// Plain streaming copy (16 B per lane per access), the byte-counter scale.
__global__ void probe_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n16) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int sl_probe_copy(int64_t bytes, double *ms) {
    if (bytes <= 0 || (bytes & 15)) {
        set_error("sl_probe_copy: bytes must be a positive multiple of 16");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    void *src = nullptr, *dst = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&src, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
    if (e == hipSuccess) e = hipMemset(src, 1, (size_t)bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(probe_copy_kernel, dim3(8192), dim3(256), 0, nullptr, (const uint4 *)src, (uint4 *)dst,
                           (uint64_t)bytes / 16);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (ms) *ms = t;
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        set_error(std::string("sl_probe_copy: ") + hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

int sl_llcs_from_kernel(const int32_t *row_to_col, int m, int n) {
    if (!row_to_col || m < 0 || n < 0) return -1;
    int acc = 0;
    for (int s = 0; s < m; ++s)
        if (row_to_col[s] >= n) acc++;
    return m - acc;
}

int sl_batch_create(sl_batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                    const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes, int flags,
                    int rows_per_lane) {
    Batch *b = nullptr;
    int rc = batch_create(&b, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes, flags, rows_per_lane);
    if (rc) return rc;
    // the opaque handle is the Batch itself
    *out = reinterpret_cast<sl_batch *>(b);
    return kOk;
}

int sl_batch_run(sl_batch *bt) {
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    return batch_run(reinterpret_cast<Batch *>(bt));
}

int sl_batch_run_per_pair(sl_batch *bt) {
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    return batch_run_per_pair(reinterpret_cast<Batch *>(bt));
}

int sl_batch_sync(sl_batch *bt) {
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    return batch_sync(reinterpret_cast<Batch *>(bt));
}

int sl_batch_download(sl_batch *h, int32_t *rc_cat, int32_t *cr_cat) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || !rc_cat) {
        set_error("sl_batch_download: null argument");
        return kErrArg;
    }
    if (cr_cat && !bt->with_cr) {
        set_error("batch was created without SL_BATCH_WITH_COL_TO_ROW");
        return kErrArg;
    }
    if (bt->perm_total) {   // ordered after the batch's launches on its stream
        SL_HIP_TRY(hipMemcpyAsync(rc_cat, bt->d_rc, sizeof(int32_t) * bt->perm_total, hipMemcpyDeviceToHost,
                                  bt->stream));
        if (cr_cat)
            SL_HIP_TRY(hipMemcpyAsync(cr_cat, bt->d_cr, sizeof(int32_t) * bt->perm_total, hipMemcpyDeviceToHost,
                                      bt->stream));
    }
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    return kOk;
}

int sl_batch_time(sl_batch *h, int mode, int warmup, int iters, double *ms_per_run, double *kernel_ms_out) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || !ms_per_run || iters <= 0 || warmup < 0) {
        set_error("sl_batch_time: bad argument");
        return kErrArg;
    }
    auto run = [&]() { return mode == 0 ? batch_run(bt) : batch_run_per_pair(bt); };
    for (int w = 0; w < warmup; ++w) {
        int rc = run();
        if (rc) return rc;
    }
    // the warm-up runs' errors; without warm-up runs only the stream is idle
    int rc = warmup > 0 ? batch_sync(bt) : kOk;
    if (rc) return rc;
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    SL_HIP_TRY(hipEventRecord(bt->ev[0], bt->stream));
    for (int it = 0; it < iters; ++it) {
        rc = run();
        if (rc) return rc;
    }
    SL_HIP_TRY(hipEventRecord(bt->ev[1], bt->stream));
    rc = batch_sync(bt);
    if (rc) return rc;
    float ms = 0.f;
    SL_HIP_TRY(hipEventElapsedTime(&ms, bt->ev[0], bt->ev[1]));
    *ms_per_run = ms / iters;
    if (kernel_ms_out) {
        // kernels only: reset outside the bracket, events around the launches
        double tot = 0.0;
        for (int it = 0; it < iters; ++it) {
            rc = batch_reset(bt);
            if (rc) return rc;
            SL_HIP_TRY(hipEventRecord(bt->ev[2], bt->stream));
            if (mode == 0) {
                SL_HIP_TRY(launch_items(bt, bt->d_items + bt->items.size(), (uint32_t)bt->order.size(), bt->d_ctl));
            } else {
                for (int i = 0; i < bt->npairs; ++i) {
                    const PairDesc &d = bt->pairs[i];
                    if (d.npasses == 0) continue;
                    SL_HIP_TRY(launch_items(bt, bt->d_items + d.item0, (uint32_t)d.npasses, bt->d_ctl + 1 + i));
                }
            }
            SL_HIP_TRY(hipEventRecord(bt->ev[3], bt->stream));
            rc = batch_sync(bt);
            if (rc) return rc;
            float kms = 0.f;
            SL_HIP_TRY(hipEventElapsedTime(&kms, bt->ev[2], bt->ev[3]));
            tot += kms;
        }
        *kernel_ms_out = tot / iters;
    }
    return kOk;
}

int sl_batch_set_tiles(sl_batch *h, int rows_per_lane, int waves_per_cu) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || (rows_per_lane != 0 && rows_per_lane != 4 && rows_per_lane != 8 && rows_per_lane != 16) ||
        waves_per_cu < 1 || waves_per_cu > 64) {
        set_error("sl_batch_set_tiles: rows_per_lane must be 0 (per pair), 4, 8 or 16 and waves_per_cu 1..64");
        return kErrArg;
    }
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    bt->rows_per_lane = rows_per_lane;
    bt->waves_per_cu = (unsigned)waves_per_cu;
    plan_items(bt);
    SL_HIP_TRY(upload_items(bt));
    return kOk;
}

int sl_batch_get_tiles(const sl_batch *h, int *rows_per_lane, int *waves_per_cu) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    if (rows_per_lane) *rows_per_lane = bt->rows_per_lane;
    if (waves_per_cu) *waves_per_cu = (int)bt->waves_per_cu;
    return kOk;
}

int sl_batch_set_critical_fraction(sl_batch *h, double fraction) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || !(fraction >= 0.0) || fraction > 1e6) {
        set_error("sl_batch_set_critical_fraction: fraction must be >= 0 (0: no critical chains)");
        return kErrArg;
    }
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    bt->critical_fraction = fraction;
    plan_items(bt);
    SL_HIP_TRY(upload_items(bt));
    return kOk;
}

int sl_batch_get_critical_fraction(const sl_batch *h, double *fraction) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt || !fraction) {
        set_error("sl_batch_get_critical_fraction: bad argument");
        return kErrArg;
    }
    *fraction = bt->critical_fraction;
    return kOk;
}

int sl_batch_pair_tiles(const sl_batch *h, int32_t *rows_per_lane, int32_t *priority) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt || (bt->npairs > 0 && !rows_per_lane && !priority)) {
        set_error("sl_batch_pair_tiles: bad argument");
        return kErrArg;
    }
    for (int i = 0; i < bt->npairs; ++i) {
        const PairDesc &d = bt->pairs[(size_t)i];
        const bool empty = d.m <= 0 || d.n <= 0;
        if (rows_per_lane) rows_per_lane[i] = empty ? 0 : d.rows_per_lane;
        if (priority) priority[i] = empty ? 0 : d.prio;
    }
    return kOk;
}

int sl_batch_set_schedule(sl_batch *h, int schedule) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || schedule < 0 || schedule > kMaxSchedule) {
        set_error("sl_batch_set_schedule: schedule must be 0..8");
        return kErrArg;
    }
    SL_HIP_TRY(hipStreamSynchronize(bt->stream));
    bt->schedule = schedule;
    plan_items(bt);
    SL_HIP_TRY(upload_items(bt));
    return kOk;
}

int sl_batch_get_schedule(const sl_batch *h, int *schedule) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt || !schedule) {
        set_error("sl_batch_get_schedule: bad argument");
        return kErrArg;
    }
    *schedule = bt->schedule;
    return kOk;
}

int sl_batch_autotune(sl_batch *h, int iters, int *rows_per_lane, int *waves_per_cu) {
    Batch *bt = reinterpret_cast<Batch *>(h);
    if (!bt || iters <= 0) {
        set_error("sl_batch_autotune: bad argument");
        return kErrArg;
    }
    struct Shape {
        int sched, r, w;
        double crit;
        double ms;
    };
    auto time_shape = [&](Shape &sh, int warmup, int runs) -> int {
        bt->schedule = sh.sched;
        bt->critical_fraction = sh.crit;
        int rc = sl_batch_set_tiles(h, sh.r, sh.w);
        if (!rc) rc = sl_batch_time(h, 0, warmup, runs, &sh.ms, nullptr);
        if (rc)
            set_error("sl_batch_autotune: rows_per_lane " + std::to_string(sh.r) + ", waves_per_cu " +
                      std::to_string(sh.w) + ", order " + std::to_string(sh.sched) + ": " + last_error());
        return rc;
    };
    // every shape `iters` cold runs (each run is a whole launch), then the
    // four fastest again warm, as a caller's steady state runs them: shapes a
    // few percent apart trade places under the first pass's noise and cold
    // launches.  The finalists are timed in interleaved rounds (a warm-up run
    // and a few timed runs each, shape after shape) and ranked by their
    // median round, so a drift of the clock over the seconds of tuning does
    // not favour the shape timed first or last.
    const int wpc_cand[] = {4, 8, 12, 16};
    const int sched_cand[] = {0, 2, 3, 4, 6};
    const double crit_cand[] = {0.0, 0.4, 0.6, 0.8};
    std::vector<Shape> shapes;
    for (int sched : sched_cand)
        for (int w : wpc_cand) {
            for (int r : kTileHeights) shapes.push_back(Shape{sched, r, w, 0.0, 0.0});
            for (double c : crit_cand) shapes.push_back(Shape{sched, 0, w, c, 0.0});
        }
    for (Shape &sh : shapes) {
        const int rc = time_shape(sh, 0, iters);
        if (rc) return rc;
    }
    std::sort(shapes.begin(), shapes.end(), [](const Shape &x, const Shape &y) { return x.ms < y.ms; });
    const size_t finalists = std::min<size_t>(4, shapes.size());
    constexpr int kRounds = 7;
    std::vector<std::vector<double>> round_ms(finalists);
    for (int round = 0; round < kRounds; ++round)
        for (size_t i = 0; i < finalists; ++i) {
            Shape sh = shapes[i];
            const int rc = time_shape(sh, 1, std::max(3, iters));
            if (rc) return rc;
            round_ms[i].push_back(sh.ms);
        }
    for (size_t i = 0; i < finalists; ++i) {
        std::vector<double> &v = round_ms[i];
        std::nth_element(v.begin(), v.begin() + kRounds / 2, v.end());
        shapes[i].ms = v[kRounds / 2];
    }
    const Shape best = *std::min_element(shapes.begin(), shapes.begin() + (std::ptrdiff_t)finalists,
                                         [](const Shape &x, const Shape &y) { return x.ms < y.ms; });
    bt->schedule = best.sched;
    bt->critical_fraction = best.crit;
    int rc = sl_batch_set_tiles(h, best.r, best.w);
    if (rc) return rc;
    if (rows_per_lane) *rows_per_lane = best.r;
    if (waves_per_cu) *waves_per_cu = best.w;
    return kOk;
}

int sl_batch_stats(const sl_batch *h, int64_t *cells, int64_t *tiles, int64_t *pairs) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    if (cells) *cells = bt->cells;
    if (tiles) *tiles = (int64_t)bt->items.size();
    if (pairs) *pairs = bt->npairs;
    return kOk;
}

int sl_batch_cell_modes(const sl_batch *h, int64_t *code_pairs, int64_t *onehot_pairs, int64_t *table_pairs) {
    const Batch *bt = reinterpret_cast<const Batch *>(h);
    if (!bt) {
        set_error("null batch");
        return kErrArg;
    }
    int64_t count[3] = {0, 0, 0};
    for (const auto &d : bt->pairs)
        if (d.m > 0 && d.n > 0) ++count[d.cell];
    if (code_pairs) *code_pairs = count[kCellCode];
    if (onehot_pairs) *onehot_pairs = count[kCellOneHot];
    if (table_pairs) *table_pairs = count[kCellTable];
    return kOk;
}

void sl_batch_destroy(sl_batch *h) { batch_free(reinterpret_cast<Batch *>(h)); }

int sl_plan_queue(int npairs, const int32_t *a_len, const int32_t *b_len, int rows_per_lane, int waves_per_cu, int cus,
                  int schedule, int32_t *pair_out, int32_t *pass_out, int64_t *nitems) {
    if (npairs < 0 || (npairs > 0 && (!a_len || !b_len)) || !nitems ||
        (rows_per_lane != 0 && rows_per_lane != 4 && rows_per_lane != 8 && rows_per_lane != 16) ||
        waves_per_cu < 1 || waves_per_cu > 64 || cus < 1 || schedule < 0 || schedule > kMaxSchedule) {
        set_error("sl_plan_queue: bad argument");
        return kErrArg;
    }
    Batch bt;   // host planning only: no device state is touched
    bt.npairs = npairs;
    bt.pairs.resize((size_t)npairs);
    for (int i = 0; i < npairs; ++i) {
        if (a_len[i] < 0 || b_len[i] < 0) {
            set_error("sl_plan_queue: negative length");
            return kErrArg;
        }
        PairDesc &d = bt.pairs[(size_t)i];
        d = PairDesc{};
        d.m = std::min(a_len[i], b_len[i]);   // rows = the shorter string, as in a batch
        d.n = std::max(a_len[i], b_len[i]);
        d.sigma = 1;
        d.tag_shift = 1;
        bt.cells += (int64_t)d.m * d.n;
    }
    bt.rows_per_lane = rows_per_lane;
    bt.waves_per_cu = (unsigned)waves_per_cu;
    bt.cus = cus;
    bt.schedule = schedule;
    plan_items(&bt);
    *nitems = (int64_t)bt.order.size();
    for (size_t k = 0; k < bt.order.size(); ++k) {
        if (pair_out) pair_out[k] = bt.order[k].pair;
        if (pass_out) pass_out[k] = bt.order[k].pass;
    }
    return kOk;
}

int sl_comb_batch(int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len, const void *b_cat,
                  const int64_t *b_off, const int32_t *b_len, int sym_bytes, int32_t *rc_cat, int32_t *cr_cat) {
    if (!rc_cat && npairs > 0) {
        set_error("sl_comb_batch: null output");
        return kErrArg;
    }
    Batch *bt = nullptr;
    int rc = batch_create(&bt, npairs, a_cat, a_off, a_len, b_cat, b_off, b_len, sym_bytes,
                          cr_cat ? SL_BATCH_WITH_COL_TO_ROW : 0, 0);
    if (rc) return rc;
    rc = batch_run(bt);
    if (!rc) rc = batch_sync(bt);
    if (!rc) rc = sl_batch_download(reinterpret_cast<sl_batch *>(bt), rc_cat, cr_cat);
    batch_free(bt);
    return rc;
}

int sl_semi_local_comb(const void *a, int m, const void *b, int n, int sym_bytes, int32_t *row_to_col,
                       int32_t *col_to_row) {
    if (m < 0 || n < 0 || (m > 0 && !a) || (n > 0 && !b) || (!row_to_col && m + n > 0)) {
        set_error("sl_semi_local_comb: bad argument");
        return kErrArg;
    }
    static const uint8_t dummy = 0;
    const int64_t zero = 0;
    const int32_t ml = m, nl = n;
    return sl_comb_batch(1, a ? a : &dummy, &zero, &ml, b ? b : &dummy, &zero, &nl, sym_bytes, row_to_col,
                         col_to_row);
}

}  // extern "C"
