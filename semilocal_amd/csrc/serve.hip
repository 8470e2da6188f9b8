// serve.hip -- the engine's request loop: continuous batching of pair
// comparisons on one GPU, admission by free device memory, braids kept
// resident in HBM, and follow-up substring queries answered from them.
//
// No reference counterpart: the reference computes one comparison per call
// and answers queries on a host-side Permutation (approximate_matching/
// utils.h:268-359, SemiLocalStringSubstringWrapper).  Each result here is
// exactly what the one-shot entry points return for the same input.
//
// Structure.  Client threads submit requests (copied into the server) and
// collect results; one scheduler thread owns every launch.  Device memory is
// one allocation made at start-up, carved by a first-fit arena; a comparison
// is admitted when its batch buffers and its braid fit, otherwise the loop
// evicts least-recently-used idle braids or waits for in-flight work to
// return memory.  Admitted comparisons are packed into micro-batches, each
// combed by one persistent launch (kernel (a)) on one of several streams, so
// work from several micro-batches overlaps on the device; requests arriving
// meanwhile form the next micro-batch.  Queries wait until their braid is
// resident, then run batched: one block per query counts
// #{r >= i+m : P[r] < j} over the braid.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/semilocal_hip.h"
#include "braid_index.h"
#include "common.h"

namespace slhip {

struct Batch;
int batch_create_arena(Batch **out, int npairs, const void *a_cat, const int64_t *a_off, const int32_t *a_len,
                       const void *b_cat, const int64_t *b_off, const int32_t *b_len, int sym_bytes,
                       DeviceArena *arena, hipStream_t stream);
int batch_launch_internal(Batch *bt);
int batch_wait_internal(Batch *bt);
void batch_destroy_internal(Batch *bt);
const int32_t *batch_rc_internal(const Batch *bt);
int64_t batch_perm_off_internal(const Batch *bt, int pair);

namespace srv {

using Clock = std::chrono::steady_clock;

// ---------------------------------------------------------------------------
// Kernels.
// ---------------------------------------------------------------------------
struct ScoreDesc {     // one finished comparison of a micro-batch
    const int32_t *rc; // its row -> col in the batch's output
    int32_t *braid;    // its resident copy
    int32_t m, n;      // caller's lengths (a: rows, b: cols)
};

struct QueryDesc {
    const int32_t *rc;
    int32_t m, n;
    int32_t kind;      // SL_QUERY_*
    int32_t i, j;
    int32_t pad;
    BraidIndexView ix; // levels > 0: the braid's wavelet-matrix index
};

__device__ __forceinline__ int block_sum(int v, int *tmp) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) tmp[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += tmp[w];
    return t;
}

// LLCS = m - #{s < m : P[s] >= n} (include/LCS/LCS.h:223-233), and the
// braid's resident copy; one block per comparison.
__global__ __launch_bounds__(256) void score_and_keep(const ScoreDesc *__restrict__ d, int64_t *__restrict__ out) {
    __shared__ int tmp[4];
    const ScoreDesc sd = d[blockIdx.x];
    const int size = sd.m + sd.n;
    int cnt = 0;
    for (int s = threadIdx.x; s < size; s += 256) {
        const int32_t v = sd.rc[s];
        if (sd.braid) sd.braid[s] = v;
        cnt += (s < sd.m && v >= sd.n);
    }
    const int tot = block_sum(cnt, tmp);
    if (threadIdx.x == 0) out[blockIdx.x] = (int64_t)sd.m - tot;
}

// Semi-local queries over a resident braid (approximate_matching/utils.h:
// 281-293, 331-342):  H(i, j) = j - i - #{r >= i + m : P[r] < j}.
__global__ __launch_bounds__(256) void answer_queries(const QueryDesc *__restrict__ q, int64_t *__restrict__ out) {
    __shared__ int tmp[4];
    const QueryDesc qd = q[blockIdx.x];
    if (qd.ix.levels > 0) {   // indexed braid: O(log(m+n)) rank walk by one lane
        if (threadIdx.x == 0) out[blockIdx.x] = wm_query(qd.ix, qd.kind, qd.i, qd.j);
        return;
    }
    int hi_, hj;   // H arguments
    if (qd.kind == SL_QUERY_STRING_SUBSTRING) {         // LCS(a, b[i:j)) = H(i, j)
        hi_ = qd.i;
        hj = qd.j;
    } else if (qd.kind == SL_QUERY_SUBSTRING_STRING) {  // LCS(a[i:j), b) = H(-i, n+m-j) - i - (m-j)
        hi_ = -qd.i;
        hj = qd.n + qd.m - qd.j;
    } else {                                            // raw H(i, j)
        hi_ = qd.i;
        hj = qd.j;
    }
    const int size = qd.m + qd.n;
    int cnt = 0;
    for (int r = hi_ + qd.m + (int)threadIdx.x; r < size; r += 256) cnt += qd.rc[r] < hj;
    const int tot = block_sum(cnt, tmp);
    if (threadIdx.x == 0) {
        int64_t h = (int64_t)hj - hi_ - tot;
        if (qd.kind == SL_QUERY_SUBSTRING_STRING) h -= (int64_t)qd.i + (qd.m - qd.j);
        out[blockIdx.x] = h;
    }
}

// ---------------------------------------------------------------------------
// First-fit arena over one device allocation.
// ---------------------------------------------------------------------------
class Arena final : public DeviceArena {
  public:
    hipError_t init(size_t cap) {
        cap_ = cap & ~(size_t)255;
        hipError_t e = hipMalloc(&base_, cap_);
        if (e == hipSuccess) free_[0] = cap_;
        return e;
    }
    ~Arena() override {
        if (base_) (void)hipFree(base_);
    }
    hipError_t alloc(void **p, size_t bytes) override {
        std::lock_guard<std::mutex> g(mu_);
        bytes = std::max<size_t>((bytes + 255) & ~(size_t)255, 256);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, sz = it->second;
            free_.erase(it);
            if (sz > bytes) free_[off + bytes] = sz - bytes;
            used_[off] = bytes;
            in_use_ += bytes;
            peak_ = std::max(peak_, in_use_);
            *p = (char *)base_ + off;
            return hipSuccess;
        }
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    void release(void *p) override {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        const size_t off = (size_t)((char *)p - (char *)base_);
        auto u = used_.find(off);
        if (u == used_.end()) return;
        size_t sz = u->second, start = off;
        used_.erase(u);
        in_use_ -= sz;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && nx->first == off + sz) {   // merge with the next hole
            sz += nx->second;
            nx = free_.erase(nx);
        }
        if (nx != free_.begin()) {                           // merge with the previous hole
            auto pv = std::prev(nx);
            if (pv->first + pv->second == start) {
                start = pv->first;
                sz += pv->second;
                free_.erase(pv);
            }
        }
        free_[start] = sz;
    }
    size_t capacity() const { return cap_; }
    size_t in_use() {
        std::lock_guard<std::mutex> g(mu_);
        return in_use_;
    }
    size_t peak() {
        std::lock_guard<std::mutex> g(mu_);
        return peak_;
    }

  private:
    void *base_ = nullptr;
    size_t cap_ = 0, in_use_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
    std::unordered_map<size_t, size_t> used_;
    std::mutex mu_;
};

// ---------------------------------------------------------------------------
// Requests, braids, in-flight micro-batches.
// ---------------------------------------------------------------------------
enum ReqState { kQueued = 0, kRunning = 1, kDone = 2 };

struct Req {
    int64_t id = 0;
    bool is_query = false;
    // comparison
    std::vector<uint8_t> a, b;   // raw symbols, sym_bytes each
    int sym_bytes = 1, m = 0, n = 0;
    bool keep = true;
    // query
    int64_t braid = 0;
    int kind = 0, qi = 0, qj = 0;
    // outcome
    int state = kQueued;
    int status = SL_OK;
    int64_t result = 0;
    std::string error;
    Clock::time_point t_submit, t_done;
};

enum BraidState { kBraidPending = 0, kBraidResident = 1, kBraidGone = 2 };

struct Braid {
    int state = kBraidPending;
    int32_t *d_rc = nullptr;
    void *d_ix = nullptr;       // wavelet-matrix index (braids of at least index_min entries)
    BraidIndexView ix{};
    bool ix_ready = false;      // the index is built: queries use it
    int64_t ix_job = 0;         // nonzero while its build is in flight
    int m = 0, n = 0;
    int refs = 0;               // unfinished queries that read it
    bool released = false;      // the client gave it up
    uint64_t last_use = 0;
    std::string gone_reason;
};

struct Flight {
    bool busy = false;
    bool query = false;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    Batch *bt = nullptr;
    std::vector<Req *> reqs;
    ScoreDesc *h_sd = nullptr, *d_sd = nullptr;
    QueryDesc *h_qd = nullptr, *d_qd = nullptr;
    int64_t *h_out = nullptr, *d_out = nullptr;
    int64_t cells = 0;
};

// Index builds of one micro-batch, on the index stream after the batch's
// results.  Memory of a braid dropped while its build runs is returned when
// the build is done.
struct IndexJob {
    hipEvent_t done = nullptr;
    std::vector<int64_t> braids;
    std::vector<void *> release_when_done;   // scratch, and memory of braids dropped meanwhile
};

struct Server {
    Arena arena;
    int max_inflight = 4;
    int max_batch = 1024;
    int64_t max_batch_cells = 0;
    int64_t index_min = 0;           // kept braids with m+n >= index_min get an index (<= 0: none)
    std::vector<Flight> flights;
    hipStream_t ix_stream = nullptr;  // index builds, off the comparisons' streams
    std::map<int64_t, IndexJob> ix_jobs;
    int64_t next_ix_job = 1;
    std::mutex mu;
    std::condition_variable cv;      // wakes the scheduler
    std::condition_variable done_cv; // wakes waiting clients
    std::deque<Req *> cmp_queue, q_queue;
    std::unordered_map<int64_t, Req *> reqs;
    std::unordered_map<int64_t, Braid> braids;
    int64_t next_id = 1;
    uint64_t use_clock = 0;
    bool stop = false;
    std::thread worker;
    std::string fatal;   // a device-side failure that stopped the loop
    // statistics
    int64_t n_compares = 0, n_queries = 0, cells_done = 0, deferrals = 0, evictions = 0, launches = 0;
    int64_t n_indexed = 0, n_index_queries = 0;
    int64_t peak_inflight = 0, peak_inflight_compares = 0;   // micro-batches on the device at one time
    Clock::time_point t_first_submit, t_last_done;
    bool any_submit = false;
};

// Counts the micro-batches launched and not yet complete (the flights that
// are busy) right after a launch, and keeps the peaks.  Called with the lock.
static void note_inflight(Server *S) {
    int64_t all = 0, compares = 0;
    for (const auto &F : S->flights) {
        all += F.busy;
        compares += F.busy && !F.query;
    }
    S->peak_inflight = std::max(S->peak_inflight, all);
    S->peak_inflight_compares = std::max(S->peak_inflight_compares, compares);
}

static void finish(Server *S, Req *r, int status, int64_t result, const std::string &err) {
    r->status = status;
    r->result = result;
    r->error = err;
    r->state = kDone;
    r->t_done = Clock::now();
    S->t_last_done = r->t_done;
}

// Returns a braid's device memory (its kernel and its index) to the pool,
// or, while its index build still reads the kernel, once the build is done.
static void braid_drop(Server *S, Braid &b) {
    auto job = b.ix_job ? S->ix_jobs.find(b.ix_job) : S->ix_jobs.end();
    if (job != S->ix_jobs.end()) {
        job->second.release_when_done.push_back(b.d_rc);
        if (b.d_ix) job->second.release_when_done.push_back(b.d_ix);
    } else {
        S->arena.release(b.d_rc);
        if (b.d_ix) S->arena.release(b.d_ix);
    }
    b.d_rc = nullptr;
    b.d_ix = nullptr;
    b.ix = BraidIndexView{};
    b.ix_ready = false;
    b.ix_job = 0;
}

// Finished index builds: their braids' queries switch to the index.
static bool complete_index_jobs(Server *S) {
    bool any = false;
    for (auto it = S->ix_jobs.begin(); it != S->ix_jobs.end();) {
        const hipError_t q = hipEventQuery(it->second.done);
        if (q == hipErrorNotReady) {
            ++it;
            continue;
        }
        if (q != hipSuccess) {
            S->fatal = std::string("request loop: index build failed: ") + hipGetErrorString(q);
            return any;
        }
        for (int64_t id : it->second.braids) {
            auto bi = S->braids.find(id);
            if (bi != S->braids.end() && bi->second.ix_job == it->first) {
                bi->second.ix_ready = true;
                bi->second.ix_job = 0;
                ++S->n_indexed;
            }
        }
        for (void *p : it->second.release_when_done) S->arena.release(p);
        (void)hipEventDestroy(it->second.done);
        it = S->ix_jobs.erase(it);
        any = true;
    }
    return any;
}

static void braid_free_if_unused(Server *S, int64_t id, Braid &b) {
    if (b.state == kBraidResident && b.refs == 0 && b.released) {
        braid_drop(S, b);
        b.state = kBraidGone;
        b.gone_reason = "released";
    }
    (void)id;
}

// Evicts the least-recently-used idle braid; false if there is none.
static bool evict_one(Server *S) {
    int64_t victim = 0;
    uint64_t best = UINT64_MAX;
    for (auto &kv : S->braids) {
        Braid &b = kv.second;
        if (b.state == kBraidResident && b.refs == 0 && b.ix_job == 0 && b.last_use < best) {
            best = b.last_use;
            victim = kv.first;
        }
    }
    if (!victim) return false;
    Braid &b = S->braids[victim];
    braid_drop(S, b);
    b.state = kBraidGone;
    b.gone_reason = "evicted to admit new work";
    ++S->evictions;
    return true;
}

// Completion of one micro-batch (its event has fired).  Called with the lock.
static void complete_flight(Server *S, Flight &F) {
    if (!F.query) {
        int rc = batch_wait_internal(F.bt);   // kernel-side error bits
        batch_destroy_internal(F.bt);
        F.bt = nullptr;
        for (size_t k = 0; k < F.reqs.size(); ++k) {
            Req *r = F.reqs[k];
            Braid &b = S->braids[r->id];
            if (rc) {
                braid_drop(S, b);
                b.state = kBraidGone;
                b.gone_reason = "comparison failed";
                finish(S, r, rc, 0, last_error());
                continue;
            }
            b.state = kBraidResident;
            b.last_use = ++S->use_clock;
            finish(S, r, SL_OK, F.h_out[k], "");
            S->cells_done += (int64_t)r->m * r->n;
            ++S->n_compares;
            if (!r->keep) {
                b.released = true;
                braid_free_if_unused(S, r->id, b);
            }
            r->a.clear();
            r->a.shrink_to_fit();
            r->b.clear();
            r->b.shrink_to_fit();
        }
    } else {
        for (size_t k = 0; k < F.reqs.size(); ++k) {
            Req *r = F.reqs[k];
            Braid &b = S->braids[r->braid];
            --b.refs;
            b.last_use = ++S->use_clock;
            finish(S, r, SL_OK, F.h_out[k], "");
            ++S->n_queries;
            braid_free_if_unused(S, r->braid, b);
        }
    }
    F.reqs.clear();
    F.busy = false;
}

// Packs the head of the comparison queue into a micro-batch on flight F.
// Called with the lock held; releases it around host packing and uploads.
// Returns true if something was launched or resolved.
static bool launch_compares(Server *S, Flight &F, std::unique_lock<std::mutex> &lk) {
    if (S->cmp_queue.empty()) return false;
    // take a run of requests with the same symbol width, within the limits
    std::vector<Req *> take;
    int64_t cells = 0;
    const int sym = S->cmp_queue.front()->sym_bytes;
    for (Req *r : S->cmp_queue) {
        if ((int)take.size() >= S->max_batch || r->sym_bytes != sym) break;
        const int64_t c = (int64_t)r->m * r->n;
        if (!take.empty() && cells + c > S->max_batch_cells) break;
        take.push_back(r);
        cells += c;
    }
    for (;;) {
        // host packing (outside the lock: the requests are owned by the queue)
        lk.unlock();
        std::vector<int64_t> a_off(take.size()), b_off(take.size());
        std::vector<int32_t> a_len(take.size()), b_len(take.size());
        size_t atot = 0, btot = 0;
        for (size_t k = 0; k < take.size(); ++k) {
            a_off[k] = (int64_t)(atot / sym);
            b_off[k] = (int64_t)(btot / sym);
            a_len[k] = take[k]->m;
            b_len[k] = take[k]->n;
            atot += take[k]->a.size();
            btot += take[k]->b.size();
        }
        std::vector<uint8_t> acat(std::max<size_t>(atot, 4)), bcat(std::max<size_t>(btot, 4));
        size_t ao = 0, bo = 0;
        for (Req *r : take) {
            if (!r->a.empty()) std::memcpy(acat.data() + ao, r->a.data(), r->a.size());
            if (!r->b.empty()) std::memcpy(bcat.data() + bo, r->b.data(), r->b.size());
            ao += r->a.size();
            bo += r->b.size();
        }
        Batch *bt = nullptr;
        int rc = batch_create_arena(&bt, (int)take.size(), acat.data(), a_off.data(), a_len.data(), bcat.data(),
                                    b_off.data(), b_len.data(), sym, &S->arena, F.stream);
        // resident braid copies
        std::vector<int32_t *> keep(take.size(), nullptr);
        bool nomem = rc == kErrNoMem;
        if (!rc) {
            for (size_t k = 0; k < take.size() && !nomem; ++k) {
                void *p = nullptr;
                if (S->arena.alloc(&p, sizeof(int32_t) * (size_t)std::max(1, take[k]->m + take[k]->n)) !=
                    hipSuccess)
                    nomem = true;
                keep[k] = (int32_t *)p;
            }
        }
        if (nomem) {
            for (int32_t *p : keep) S->arena.release(p);
            if (bt) batch_destroy_internal(bt);
            lk.lock();
            if (evict_one(S)) continue;                    // retry with the memory of an idle braid
            if (take.size() > 1) {                          // retry with a smaller micro-batch
                take.resize(take.size() / 2);
                cells = 0;
                for (Req *r : take) cells += (int64_t)r->m * r->n;
                continue;
            }
            bool others = false;
            for (auto &G : S->flights) others |= G.busy;
            if (others) {                                   // in-flight work will return memory
                ++S->deferrals;
                return false;
            }
            Req *r = take[0];                               // cannot fit even on an idle device
            S->cmp_queue.pop_front();
            S->braids[r->id].state = kBraidGone;
            S->braids[r->id].gone_reason = "comparison does not fit the device memory pool";
            finish(S, r, SL_ERR_NOMEM, 0, "comparison needs more device memory than the pool holds");
            S->done_cv.notify_all();
            return true;
        }
        if (rc) {   // not a memory shortage: fail the micro-batch's requests
            lk.lock();
            const std::string msg = last_error();
            for (Req *r : take) {
                S->cmp_queue.pop_front();
                S->braids[r->id].state = kBraidGone;
                S->braids[r->id].gone_reason = "comparison failed";
                finish(S, r, rc, 0, msg);
            }
            S->done_cv.notify_all();
            return true;
        }
        // launch: combing, then score + braid copies, results to pinned host memory
        for (size_t k = 0; k < take.size(); ++k)
            F.h_sd[k] = ScoreDesc{batch_rc_internal(bt) + batch_perm_off_internal(bt, (int)k), keep[k], take[k]->m,
                                  take[k]->n};
        rc = batch_launch_internal(bt);
        hipError_t e = rc ? hipErrorUnknown : hipSuccess;
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.d_sd, F.h_sd, sizeof(ScoreDesc) * take.size(), hipMemcpyHostToDevice, F.stream);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(score_and_keep, dim3((unsigned)take.size()), dim3(256), 0, F.stream, F.d_sd, F.d_out);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(F.h_out, F.d_out, sizeof(int64_t) * take.size(), hipMemcpyDeviceToHost, F.stream);
        if (e == hipSuccess) e = hipEventRecord(F.done, F.stream);
        // indexes of the large braids: built from their resident copies on the
        // index stream once the micro-batch's results are out, so no result
        // waits for an index (a braid whose index does not fit keeps the scan)
        std::vector<void *> ixmem(take.size(), nullptr);
        std::vector<BraidIndexView> ixview(take.size());
        IndexJob job;
        bool waited = false;
        for (size_t k = 0; k < take.size() && e == hipSuccess; ++k) {
            const int64_t size = (int64_t)take[k]->m + take[k]->n;
            if (S->index_min <= 0 || size < S->index_min || !take[k]->keep) continue;
            void *mem = nullptr, *scratch = nullptr;
            if (S->arena.alloc(&mem, braid_index_bytes(size)) != hipSuccess) continue;
            if (S->arena.alloc(&scratch, braid_index_scratch_bytes(size)) != hipSuccess) {
                S->arena.release(mem);
                continue;
            }
            if (!waited) e = hipStreamWaitEvent(S->ix_stream, F.done, 0);
            waited = true;
            if (e == hipSuccess) e = braid_index_build(keep[k], take[k]->m, take[k]->n, mem, scratch, S->ix_stream,
                                                       &ixview[k]);
            ixmem[k] = mem;
            job.braids.push_back(take[k]->id);
            job.release_when_done.push_back(scratch);
        }
        if (e == hipSuccess && waited) e = hipEventCreateWithFlags(&job.done, hipEventDisableTiming);
        if (e == hipSuccess && waited) e = hipEventRecord(job.done, S->ix_stream);
        lk.lock();
        if (e != hipSuccess) {
            S->fatal = std::string("request loop launch failed: ") + hipGetErrorString(e);
            for (int32_t *p : keep) S->arena.release(p);
            for (void *p : ixmem) S->arena.release(p);
            for (void *p : job.release_when_done) S->arena.release(p);
            if (job.done) (void)hipEventDestroy(job.done);
            batch_destroy_internal(bt);
            return false;
        }
        int64_t job_id = 0;
        if (waited) {
            job_id = S->next_ix_job++;
            S->ix_jobs.emplace(job_id, std::move(job));
        }
        for (size_t k = 0; k < take.size(); ++k) {
            S->cmp_queue.pop_front();
            Braid &b = S->braids[take[k]->id];
            b.d_rc = keep[k];
            if (ixmem[k]) {
                b.d_ix = ixmem[k];
                b.ix = ixview[k];
                b.ix_ready = false;
                b.ix_job = job_id;
            }
            b.m = take[k]->m;
            b.n = take[k]->n;
            take[k]->state = kRunning;
        }
        F.busy = true;
        F.query = false;
        F.bt = bt;
        F.reqs = take;
        F.cells = cells;
        ++S->launches;
        note_inflight(S);
        return true;
    }
}

// Launches every query whose braid is resident (up to max_batch) on flight F.
static bool launch_queries(Server *S, Flight &F, std::unique_lock<std::mutex> &lk) {
    std::vector<Req *> take;
    bool resolved = false;
    for (auto it = S->q_queue.begin(); it != S->q_queue.end() && (int)take.size() < S->max_batch;) {
        Req *r = *it;
        auto bi = S->braids.find(r->braid);
        Braid &b = bi->second;
        if (b.state == kBraidGone) {
            --b.refs;
            finish(S, r, SL_ERR_ARG, 0, "braid " + std::to_string(r->braid) + " is not resident: " + b.gone_reason);
            it = S->q_queue.erase(it);
            resolved = true;
            continue;
        }
        if (b.state == kBraidResident) {
            take.push_back(r);
            it = S->q_queue.erase(it);
            continue;
        }
        ++it;   // its comparison is still queued or running
    }
    if (resolved) S->done_cv.notify_all();
    if (take.empty()) return resolved;
    for (size_t k = 0; k < take.size(); ++k) {
        const Braid &b = S->braids[take[k]->braid];
        F.h_qd[k] = QueryDesc{b.d_rc, b.m, b.n, take[k]->kind, take[k]->qi, take[k]->qj, 0,
                              b.ix_ready ? b.ix : BraidIndexView{}};
        if (b.ix_ready && b.ix.levels > 0) ++S->n_index_queries;
        take[k]->state = kRunning;
    }
    F.busy = true;
    F.query = true;
    F.reqs = take;
    lk.unlock();
    hipError_t e = hipMemcpyAsync(F.d_qd, F.h_qd, sizeof(QueryDesc) * take.size(), hipMemcpyHostToDevice, F.stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(answer_queries, dim3((unsigned)take.size()), dim3(256), 0, F.stream, F.d_qd, F.d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(F.h_out, F.d_out, sizeof(int64_t) * take.size(), hipMemcpyDeviceToHost, F.stream);
    if (e == hipSuccess) e = hipEventRecord(F.done, F.stream);
    lk.lock();
    if (e != hipSuccess) S->fatal = std::string("request loop launch failed: ") + hipGetErrorString(e);
    ++S->launches;
    note_inflight(S);
    return true;
}

static void run_loop(Server *S) {
    std::unique_lock<std::mutex> lk(S->mu);
    while (!S->stop && S->fatal.empty()) {
        bool progress = false;
        for (auto &F : S->flights) {
            if (!F.busy) continue;
            const hipError_t q = hipEventQuery(F.done);
            if (q == hipSuccess) {
                complete_flight(S, F);
                progress = true;
            } else if (q != hipErrorNotReady) {
                S->fatal = std::string("request loop: device failure: ") + hipGetErrorString(q);
                break;
            }
        }
        if (!S->ix_jobs.empty() && complete_index_jobs(S)) progress = true;
        if (progress) S->done_cv.notify_all();
        if (!S->fatal.empty()) break;
        // queries first (short), then comparisons, each on a free flight
        for (auto &F : S->flights) {
            if (F.busy) continue;
            if (!S->q_queue.empty() && launch_queries(S, F, lk)) {
                progress = true;
                if (F.busy) continue;
            }
            if (!F.busy && !S->cmp_queue.empty() && launch_compares(S, F, lk)) progress = true;
        }
        if (!progress) S->cv.wait_for(lk, std::chrono::microseconds(50));
    }
    // a failure ends every outstanding request
    if (!S->fatal.empty()) {
        for (auto &kv : S->reqs)
            if (kv.second->state != kDone) finish(S, kv.second, SL_ERR_HIP, 0, S->fatal);
        S->done_cv.notify_all();
    }
}

static void server_free(Server *S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> g(S->mu);
        S->stop = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    for (auto &F : S->flights) {
        if (F.stream) (void)hipStreamSynchronize(F.stream);
        if (F.bt) batch_destroy_internal(F.bt);
        if (F.h_sd) (void)hipHostFree(F.h_sd);
        if (F.h_qd) (void)hipHostFree(F.h_qd);
        if (F.h_out) (void)hipHostFree(F.h_out);
        if (F.d_sd) (void)hipFree(F.d_sd);
        if (F.d_qd) (void)hipFree(F.d_qd);
        if (F.d_out) (void)hipFree(F.d_out);
        if (F.done) (void)hipEventDestroy(F.done);
        if (F.stream) (void)hipStreamDestroy(F.stream);
    }
    if (S->ix_stream) (void)hipStreamSynchronize(S->ix_stream);
    for (auto &kv : S->ix_jobs)
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    if (S->ix_stream) (void)hipStreamDestroy(S->ix_stream);
    for (auto &kv : S->reqs) delete kv.second;
    delete S;
}

}  // namespace srv
}  // namespace slhip

using namespace slhip;
using namespace slhip::srv;

extern "C" {

int sl_server_create(sl_server **out, int64_t pool_bytes, int max_inflight, int max_batch, int64_t max_batch_cells) {
    if (!out || pool_bytes < 0 || max_inflight < 0 || max_batch < 0 || max_batch_cells < 0) {
        set_error("sl_server_create: bad argument");
        return kErrArg;
    }
    *out = nullptr;
    int rc = ensure_device();
    if (rc) return rc;
    size_t free_b = 0, total_b = 0;
    SL_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t cap_free = (size_t)((double)free_b * 0.9);
    const size_t cap = pool_bytes > 0 ? std::min<size_t>((size_t)pool_bytes, cap_free) : cap_free;
    Server *S = new Server();
    S->max_inflight = max_inflight > 0 ? max_inflight : 4;
    S->max_batch = max_batch > 0 ? max_batch : 1024;
    S->max_batch_cells = max_batch_cells > 0 ? max_batch_cells : (int64_t)4e10;
    hipError_t e = S->arena.init(cap);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&S->ix_stream, hipStreamNonBlocking);
    S->flights.resize((size_t)S->max_inflight);
    for (auto &F : S->flights) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&F.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&F.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void **)&F.h_sd, sizeof(ScoreDesc) * S->max_batch);
        if (e == hipSuccess) e = hipHostMalloc((void **)&F.h_qd, sizeof(QueryDesc) * S->max_batch);
        if (e == hipSuccess) e = hipHostMalloc((void **)&F.h_out, sizeof(int64_t) * S->max_batch);
        if (e == hipSuccess) e = hipMalloc((void **)&F.d_sd, sizeof(ScoreDesc) * S->max_batch);
        if (e == hipSuccess) e = hipMalloc((void **)&F.d_qd, sizeof(QueryDesc) * S->max_batch);
        if (e == hipSuccess) e = hipMalloc((void **)&F.d_out, sizeof(int64_t) * S->max_batch);
    }
    if (e != hipSuccess) {
        set_error(std::string("sl_server_create: ") + hipGetErrorString(e));
        server_free(S);
        return e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;
    }
    S->worker = std::thread(run_loop, S);
    *out = reinterpret_cast<sl_server *>(S);
    return kOk;
}

int sl_server_submit_compare(sl_server *h, const void *a, int m, const void *b, int n, int sym_bytes, int keep,
                             int64_t *id) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S || !id || m < 0 || n < 0 || (m > 0 && !a) || (n > 0 && !b) ||
        (sym_bytes != 1 && sym_bytes != 2 && sym_bytes != 4)) {
        set_error("sl_server_submit_compare: bad argument");
        return kErrArg;
    }
    Req *r = new Req();
    r->sym_bytes = sym_bytes;
    r->m = m;
    r->n = n;
    r->keep = keep != 0;
    r->a.assign((const uint8_t *)a, (const uint8_t *)a + (size_t)m * sym_bytes);
    r->b.assign((const uint8_t *)b, (const uint8_t *)b + (size_t)n * sym_bytes);
    {
        std::lock_guard<std::mutex> g(S->mu);
        if (!S->fatal.empty()) {
            delete r;
            set_error(S->fatal);
            return kErrHip;
        }
        r->id = S->next_id++;
        r->t_submit = Clock::now();
        if (!S->any_submit) {
            S->any_submit = true;
            S->t_first_submit = r->t_submit;
        }
        S->reqs[r->id] = r;
        S->braids[r->id] = Braid();
        S->cmp_queue.push_back(r);
        *id = r->id;
    }
    S->cv.notify_all();
    return kOk;
}

int sl_server_submit_query(sl_server *h, int64_t braid, int kind, int i, int j, int64_t *id) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S || !id || kind < SL_QUERY_H || kind > SL_QUERY_SUBSTRING_STRING) {
        set_error("sl_server_submit_query: bad argument");
        return kErrArg;
    }
    std::lock_guard<std::mutex> g(S->mu);
    auto bi = S->braids.find(braid);
    if (bi == S->braids.end()) {
        set_error("sl_server_submit_query: no comparison " + std::to_string(braid));
        return kErrArg;
    }
    Braid &b = bi->second;
    if (b.state == kBraidGone || b.released) {
        set_error("sl_server_submit_query: braid " + std::to_string(braid) + " is not resident: " +
                  (b.released ? std::string("released") : b.gone_reason));
        return kErrArg;
    }
    const Req *src = S->reqs[braid];
    const int m = src->m, nn = src->n;
    bool ok;
    if (kind == SL_QUERY_STRING_SUBSTRING) ok = 0 <= i && i <= j && j <= nn;
    else if (kind == SL_QUERY_SUBSTRING_STRING) ok = 0 <= i && i <= j && j <= m;
    else ok = -m <= i && i <= nn && 0 <= j && j <= m + nn;
    if (!ok) {
        set_error("sl_server_submit_query: window out of range");
        return kErrArg;
    }
    Req *r = new Req();
    r->is_query = true;
    r->braid = braid;
    r->kind = kind;
    r->qi = i;
    r->qj = j;
    r->id = S->next_id++;
    r->t_submit = Clock::now();
    if (!S->any_submit) {
        S->any_submit = true;
        S->t_first_submit = r->t_submit;
    }
    ++b.refs;
    S->reqs[r->id] = r;
    S->q_queue.push_back(r);
    *id = r->id;
    S->cv.notify_all();
    return kOk;
}

// Result of a request: SL_OK and *result when done (the request's own status
// if it failed), SL_PENDING while queued or running.  wait != 0 blocks until
// it is done.
int sl_server_result(sl_server *h, int64_t id, int wait, int64_t *result, double *latency_ms) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S) {
        set_error("sl_server_result: null handle");
        return kErrArg;
    }
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) {
        set_error("sl_server_result: unknown request " + std::to_string(id));
        return kErrArg;
    }
    Req *r = it->second;
    if (wait)
        S->done_cv.wait(lk, [&] { return r->state == kDone; });
    if (r->state != kDone) return SL_PENDING;
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    if (r->status != SL_OK) {
        set_error(r->error);
        return r->status;
    }
    if (result) *result = r->result;
    return kOk;
}

int sl_server_download_braid(sl_server *h, int64_t braid, int32_t *row_to_col) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S || !row_to_col) {
        set_error("sl_server_download_braid: bad argument");
        return kErrArg;
    }
    int32_t *d = nullptr;
    int size = 0;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto bi = S->braids.find(braid);
        if (bi == S->braids.end() || bi->second.state != kBraidResident) {
            set_error("sl_server_download_braid: braid " + std::to_string(braid) + " is not resident");
            return kErrArg;
        }
        Braid &b = bi->second;
        ++b.refs;   // pinned while copying
        d = b.d_rc;
        size = b.m + b.n;
    }
    hipError_t e = size ? hipMemcpy(row_to_col, d, sizeof(int32_t) * (size_t)size, hipMemcpyDeviceToHost)
                        : hipSuccess;
    {
        std::lock_guard<std::mutex> g(S->mu);
        Braid &b = S->braids[braid];
        --b.refs;
        b.last_use = ++S->use_clock;
        braid_free_if_unused(S, braid, b);
    }
    if (e != hipSuccess) {
        set_error(std::string("sl_server_download_braid: ") + hipGetErrorString(e));
        return kErrHip;
    }
    return kOk;
}

int sl_server_release(sl_server *h, int64_t braid) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S) {
        set_error("sl_server_release: null handle");
        return kErrArg;
    }
    std::lock_guard<std::mutex> g(S->mu);
    auto bi = S->braids.find(braid);
    if (bi == S->braids.end()) {
        set_error("sl_server_release: no comparison " + std::to_string(braid));
        return kErrArg;
    }
    bi->second.released = true;
    braid_free_if_unused(S, braid, bi->second);
    return kOk;
}

int sl_server_stats(sl_server *h, sl_server_stats_t *st) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S || !st) {
        set_error("sl_server_stats: bad argument");
        return kErrArg;
    }
    std::lock_guard<std::mutex> g(S->mu);
    st->compares_done = S->n_compares;
    st->queries_done = S->n_queries;
    st->cells_done = S->cells_done;
    st->launches = S->launches;
    st->deferrals = S->deferrals;
    st->evictions = S->evictions;
    st->pool_bytes = (int64_t)S->arena.capacity();
    st->pool_in_use = (int64_t)S->arena.in_use();
    st->pool_peak = (int64_t)S->arena.peak();
    int64_t resident = 0;
    for (auto &kv : S->braids) resident += kv.second.state == kBraidResident;
    st->resident_braids = resident;
    st->indexed_braids = S->n_indexed;
    st->index_queries = S->n_index_queries;
    st->peak_inflight = S->peak_inflight;
    st->peak_inflight_compares = S->peak_inflight_compares;
    st->busy_seconds = S->any_submit ? std::chrono::duration<double>(S->t_last_done - S->t_first_submit).count() : 0;
    return kOk;
}

int sl_server_set_index(sl_server *h, int64_t min_size) {
    Server *S = reinterpret_cast<Server *>(h);
    if (!S) {
        set_error("sl_server_set_index: null handle");
        return kErrArg;
    }
    std::lock_guard<std::mutex> g(S->mu);
    S->index_min = min_size;
    return kOk;
}

void sl_server_destroy(sl_server *h) { server_free(reinterpret_cast<Server *>(h)); }

}  // extern "C"
