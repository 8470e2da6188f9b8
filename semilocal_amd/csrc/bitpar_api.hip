// bitpar_api.hip -- host side of kernel (b): resident bit-parallel problems.
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/semilocal_hip.h"
#include "common.h"

namespace slhip {

hipError_t launch_bitpar(const BitDesc &d, unsigned waves, LaunchCtl *ctl, uint32_t spin_limit, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t *p, uint32_t value, uint64_t count, hipStream_t stream);

// The carry form runs the Hyyro kernel on the two full strings (every symbol
// of the packed words, padding included): the longer one bit-packed, which
// the kernel takes up to 2^20 symbols.
constexpr int64_t kCarryMaxColumns = 1 << 20;

struct BitPlan {
    BitDesc d{};
    int64_t a_total = 0;
    uint32_t *d_words = nullptr;   // a_rev (M words) then b (N words), 32-bit view
    uint64_t *d_buf = nullptr;
    uint32_t *d_scratch = nullptr; // [0,2): LaunchCtl, [2]: left strands leaving at the bottom
    int form = SL_BITPAR_COMBING;  // the form the next run uses
    int run_form = SL_BITPAR_COMBING;   // the form of the last run: the one result() reads
    sl_hyyro *carry = nullptr;     // the carry form's handle (null: the form is not offered)
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    uint32_t spin_limit = 1u << 23;
    unsigned waves = 0;
    uint64_t cleared_at = 0;       // run serial when d_buf was last cleared
};

static void bitplan_free(BitPlan *p) {
    if (!p) return;
    (void)hipFree(p->d_words);
    (void)hipFree(p->d_buf);
    (void)hipFree(p->d_scratch);
    if (p->carry) sl_hyyro_destroy(p->carry);
    for (auto &e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

static int bitplan_run(BitPlan *p) {
    if (!p->d_words || !p->d_buf || !p->d_scratch || !p->stream) {
        set_error("bit-parallel plan is missing a device buffer (internal defect); nothing launched");
        return kErrHip;
    }
    p->run_form = p->form;
    if (p->form == SL_BITPAR_CARRY) return sl_hyyro_run(p->carry);
    // unique tags for this run's hand-offs; tag 0 is a cleared buffer
    if (next_run_tag(p->cleared_at, &p->d.epoch))
        SL_HIP_TRY(launch_fill_u32(reinterpret_cast<uint32_t *>(p->d_buf), 0u, 2u * (uint64_t)p->d.N, p->stream));
    SL_HIP_TRY(launch_fill_u32(p->d_scratch, 0u, 4, p->stream));
    SL_HIP_TRY(launch_bitpar(p->d, p->waves, reinterpret_cast<LaunchCtl *>(p->d_scratch), p->spin_limit, p->stream));
    return kOk;
}

static int bitplan_result(BitPlan *p, int *llcs) {
    if (p->run_form == SL_BITPAR_CARRY) {
        // the LLCS of the two full packed strings is the number of left
        // strands leaving at the bottom
        int32_t full = 0;
        const int rc = sl_hyyro_result(p->carry, &full);
        if (rc) return rc;
        if (llcs) *llcs = (int)(p->a_total - 32 * (int64_t)p->d.M + full);
        return kOk;
    }
    SL_HIP_TRY(hipStreamSynchronize(p->stream));
    uint32_t h[4];
    SL_HIP_TRY(hipMemcpy(h, p->d_scratch, sizeof h, hipMemcpyDeviceToHost));
    if (h[1]) {
        set_error("bit-parallel kernel: a bounded wait timed out; result is void");
        return kErrTimeout;
    }
    // LLCS = |a| - 32M + (left strands leaving at the bottom), counted by the last pass
    if (llcs) *llcs = (int)(p->a_total - 32 * (int64_t)p->d.M + (int64_t)h[2]);
    return kOk;
}

}  // namespace slhip

using namespace slhip;

extern "C" {

int sl_bitpar_create(sl_bitpar **out, const uint64_t *a_rev, int a_words, const uint64_t *b, int b_words,
                     int a_total_symbols) {
    if (!out || !a_rev || !b || a_words <= 0 || b_words <= 0 || a_total_symbols < 0 ||
        (int64_t)a_words * 64 > 0x7fffffffLL || (int64_t)b_words * 64 > 0x7fffffffLL) {
        set_error("sl_bitpar_create: bad argument");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    BitPlan *p = new BitPlan();
    p->a_total = a_total_symbols;
    const int64_t M = 2LL * a_words, N = 2LL * b_words;   // 32-bit words
    hipError_t e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&p->ev[k]);
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_words, sizeof(uint32_t) * (size_t)(M + N));
    if (e == hipSuccess) e = hipMemcpy(p->d_words, a_rev, sizeof(uint64_t) * a_words, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(p->d_words + M, b, sizeof(uint64_t) * b_words, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_buf, sizeof(uint64_t) * (size_t)N);
    // cleared on the handle's own stream, ordered before every launch
    if (e == hipSuccess) e = launch_fill_u32(reinterpret_cast<uint32_t *>(p->d_buf), 0u, 2u * (uint64_t)N, p->stream);
    p->cleared_at = run_serial_now();
    if (e == hipSuccess) e = hipMalloc((void **)&p->d_scratch, sizeof(uint32_t) * 4);
    if (e == hipSuccess) e = launch_fill_u32(p->d_scratch, 0u, 4, p->stream);   // no bits left by others
    // the word uploads go through the null stream, which is not ordered with
    // the handle's non-blocking stream: finish them before the first launch
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        set_error(std::string("sl_bitpar_create: ") + hipGetErrorString(e));
        bitplan_free(p);
        return e == hipErrorOutOfMemory ? kErrNoMem : kErrHip;
    }
    p->d.a_rev = p->d_words;
    p->d.b = p->d_words + M;
    p->d.buf = p->d_buf;
    p->d.dis = p->d_scratch + 2;
    p->d.M = (int32_t)M;
    p->d.N = (int32_t)N;
    p->d.npasses = (int32_t)((M + kBitRowLanes - 1) / kBitRowLanes);
    p->d.epoch = 0;
    p->waves = (unsigned)std::min<int64_t>(p->d.npasses, (int64_t)device_cus() * 16);
    // Carry form: the full strings as symbols (a_rev word M-1-g holds a's
    // symbols 32g .. 32g+31, symbol 32g+i at bit 31-i; b word c holds b's
    // symbols 32c .. 32c+31, symbol 32c+j at bit j).  The rows are the
    // shorter string, the bit vector the longer; 32 bits per lane where the
    // rows outnumber the vector's 32-bit words (a lone long chain of steps
    // runs faster on the narrower lanes), the kernel's default otherwise.
    if (std::max(M, N) * 32 <= kCarryMaxColumns) {
        const uint32_t *ar32 = reinterpret_cast<const uint32_t *>(a_rev);
        const uint32_t *b32 = reinterpret_cast<const uint32_t *>(b);
        std::vector<uint8_t> sa((size_t)(32 * M)), sb((size_t)(32 * N));
        for (int64_t g = 0; g < M; ++g)
            for (int i = 0; i < 32; ++i) sa[(size_t)(32 * g + i)] = (uint8_t)(ar32[M - 1 - g] >> (31 - i) & 1u);
        for (int64_t c = 0; c < N; ++c)
            for (int j = 0; j < 32; ++j) sb[(size_t)(32 * c + j)] = (uint8_t)(b32[c] >> j & 1u);
        const bool rows_a = M <= N;
        const std::vector<uint8_t> &rows = rows_a ? sa : sb, &cols = rows_a ? sb : sa;
        const int64_t off = 0;
        const int32_t lr = (int32_t)rows.size(), lc = (int32_t)cols.size();
        int rc2 = sl_hyyro_create(&p->carry, 1, rows.data(), &off, &lr, cols.data(), &off, &lc, 1);
        if (!rc2 && (int64_t)lr >= (int64_t)lc / 32) {
            const int32_t wpl[SL_HYYRO_BUCKETS] = {1, 1, 0, 1, 1};   // bucket 2 offers 2 or 4 only
            rc2 = sl_hyyro_set_words_per_lane(p->carry, wpl);
        }
        if (rc2) {
            bitplan_free(p);
            return rc2;
        }
        p->form = SL_BITPAR_CARRY;
    }
    *out = reinterpret_cast<sl_bitpar *>(p);
    return kOk;
}

int sl_bitpar_set_form(sl_bitpar *h, int form) {
    BitPlan *p = reinterpret_cast<BitPlan *>(h);
    if (!p || (form != SL_BITPAR_CARRY && form != SL_BITPAR_COMBING)) {
        set_error("sl_bitpar_set_form: bad argument");
        return kErrArg;
    }
    if (form == SL_BITPAR_CARRY && !p->carry) {
        set_error("sl_bitpar_set_form: the carry form needs the longer string within 2^20 symbols");
        return kErrArg;
    }
    // each form has its own buffers and stream; result() reads the form of
    // the last run, so a switch between a run and its result is harmless
    p->form = form;
    return kOk;
}

int sl_bitpar_get_form(sl_bitpar *h, int *form) {
    BitPlan *p = reinterpret_cast<BitPlan *>(h);
    if (!p || !form) {
        set_error("sl_bitpar_get_form: null argument");
        return kErrArg;
    }
    *form = p->form;
    return kOk;
}

int sl_bitpar_run(sl_bitpar *h) {
    if (!h) {
        set_error("null plan");
        return kErrArg;
    }
    return bitplan_run(reinterpret_cast<BitPlan *>(h));
}

int sl_bitpar_result(sl_bitpar *h, int *llcs) {
    if (!h || !llcs) {
        set_error("sl_bitpar_result: null argument");
        return kErrArg;
    }
    return bitplan_result(reinterpret_cast<BitPlan *>(h), llcs);
}

int sl_bitpar_time(sl_bitpar *h, int warmup, int iters, double *ms_per_run) {
    BitPlan *p = reinterpret_cast<BitPlan *>(h);
    if (!p || iters <= 0 || warmup < 0 || !ms_per_run) {
        set_error("sl_bitpar_time: bad argument");
        return kErrArg;
    }
    if (p->form == SL_BITPAR_CARRY) {
        p->run_form = SL_BITPAR_CARRY;
        return sl_hyyro_time(p->carry, warmup, iters, ms_per_run);
    }
    int rc = kOk;
    for (int w = 0; w < warmup && !rc; ++w) rc = bitplan_run(p);
    if (!rc) rc = bitplan_result(p, nullptr);
    if (rc) return rc;
    SL_HIP_TRY(hipEventRecord(p->ev[0], p->stream));
    for (int it = 0; it < iters && !rc; ++it) rc = bitplan_run(p);
    if (rc) return rc;
    SL_HIP_TRY(hipEventRecord(p->ev[1], p->stream));
    rc = bitplan_result(p, nullptr);
    if (rc) return rc;
    float ms = 0.f;
    SL_HIP_TRY(hipEventElapsedTime(&ms, p->ev[0], p->ev[1]));
    *ms_per_run = ms / iters;
    return kOk;
}

void sl_bitpar_destroy(sl_bitpar *h) { bitplan_free(reinterpret_cast<BitPlan *>(h)); }

int sl_llcs_bitparallel(const uint64_t *a_rev, int a_words, const uint64_t *b, int b_words, int a_total_symbols,
                        int threads, int formula_one, int *llcs) {
    (void)threads;      // the GPU decides its own parallelism
    (void)formula_one;  // both cell formulas give the same score; formula 2 is used
    sl_bitpar *h = nullptr;
    int rc = sl_bitpar_create(&h, a_rev, a_words, b, b_words, a_total_symbols);
    if (rc) return rc;
    rc = sl_bitpar_run(h);
    if (!rc) rc = sl_bitpar_result(h, llcs);
    sl_bitpar_destroy(h);
    return rc;
}

}  // extern "C"
