// noalloc.hip -- sl_semi_local_comb_noalloc: the plain semi-local LCS of one
// pair (kernel (a)) with device buffers in and out and NO device allocation.
//
// The one-shot entry point sl_semi_local_comb allocates its inputs, outputs,
// pair table, queue, launch control block and inter-pass strand buffer.  This
// variant takes the inputs and outputs from the caller (device pointers) and
// keeps everything else in one file-scope device array, `scratch`, of a fixed
// size: the launch control block, the pair descriptor, the queue of passes
// and the tagged strand buffer (one 64-bit entry per column).  A set-up
// kernel rewrites all of it at the start of every call, so a call never sees
// the previous call's contents.
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address is taken once per device with hipGetSymbolAddress (with that
// device current) and only that address is ever passed to kernels.  Every
// call on a device shares that device's array, so calls are serialised per
// device by a host lock held for the whole (synchronous) call.
#include <mutex>

#include "../../include/semilocal_hip.h"
#include "common.h"

namespace slhip {

hipError_t launch_comb_batch(int sym_bytes, int rows_per_lane, unsigned waves, hipStream_t stream,
                             const PairDesc *pairs, const WorkItem *items, uint32_t nitems, const void *pool,
                             int32_t *rc, int32_t *cr, uint64_t *buf, uint32_t epoch, LaunchCtl *ctl,
                             uint32_t spin_limit, uint32_t table_words, uint32_t *wait_flags);
int model_rows_one_pair(int32_t m, int32_t n, int cus);

namespace {

// Scratch layout (bytes): [0, 256) launch control block; [256, 512) pair
// descriptor; [512, 512 + 8 * kMaxItems) queue; then kMaxCols strand entries.
constexpr int64_t kMaxCols = SL_NOALLOC_MAX_LEN;        // longer string (columns)
constexpr int64_t kMaxItems = SL_NOALLOC_MAX_LEN / 256;  // passes of 64 x 4 rows at most
constexpr size_t kOffCtl = 0;
constexpr size_t kOffPair = 256;
constexpr size_t kOffItems = 512;
constexpr size_t kOffBuf = kOffItems + sizeof(WorkItem) * (size_t)kMaxItems;
constexpr size_t kScratchBytes = kOffBuf + sizeof(uint64_t) * (size_t)kMaxCols;
static_assert(sizeof(LaunchCtl) <= 256 && sizeof(PairDesc) <= 256, "scratch header slots");
static_assert(kOffBuf % 256 == 0, "strand buffer alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu[kMaxDevices];          // one call at a time per device
unsigned char *g_ptr[kMaxDevices];      // device address of `scratch` per device (0: not looked up yet)

// Rewrites the whole per-call state: control block, pair descriptor, queue
// (pass k of pair 0 at position k) and the n strand entries the passes read.
__global__ void noalloc_setup(unsigned char *s, PairDesc d, uint32_t npasses) {
    const uint64_t tid = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    if (tid == 0) {
        *reinterpret_cast<LaunchCtl *>(s + kOffCtl) = LaunchCtl{0u, 0u};
        *reinterpret_cast<PairDesc *>(s + kOffPair) = d;
    }
    WorkItem *items = reinterpret_cast<WorkItem *>(s + kOffItems);
    for (uint64_t k = tid; k < npasses; k += stride) items[k] = WorkItem{0, (int32_t)k};
    uint64_t *buf = reinterpret_cast<uint64_t *>(s + kOffBuf);
    for (uint64_t k = tid; k < (uint64_t)d.n; k += stride) buf[k] = 0;
}

// The kernel of a pair with an empty string is the identity.
__global__ void noalloc_iota(int32_t *rc, int32_t *cr, int32_t size) {
    for (int32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < size; k += gridDim.x * blockDim.x) {
        rc[k] = k;
        if (cr) cr[k] = k;
    }
}

// Device address of this device's `scratch` (looked up on first use, with the
// device current).  Called with g_mu[device] held.
int scratch_for(int device, unsigned char **out) {
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("sl_semi_local_comb_noalloc: hipGetSymbolAddress(scratch): ") +
                      hipGetErrorString(e));
            return kErrHip;
        }
        g_ptr[device] = static_cast<unsigned char *>(p);
    }
    *out = g_ptr[device];
    return kOk;
}

}  // namespace
}  // namespace slhip

using namespace slhip;

extern "C" {

int sl_semi_local_comb_noalloc(const void *d_a, int m, const void *d_b, int n, int sym_bytes, int32_t *d_row_to_col,
                               int32_t *d_col_to_row, void *stream) {
    if (m < 0 || n < 0 || (m > 0 && !d_a) || (n > 0 && !d_b) || (!d_row_to_col && m + n > 0)) {
        set_error("sl_semi_local_comb_noalloc: bad argument");
        return kErrArg;
    }
    if (sym_bytes != 1 && sym_bytes != 2) {
        set_error("sl_semi_local_comb_noalloc: sym_bytes must be 1 or 2 (no host-side remapping here)");
        return kErrArg;
    }
    if (m > SL_NOALLOC_MAX_LEN || n > SL_NOALLOC_MAX_LEN) {
        set_error("sl_semi_local_comb_noalloc: strings longer than " + std::to_string(SL_NOALLOC_MAX_LEN) +
                  " symbols exceed the fixed scratch array; use sl_semi_local_comb");
        return kErrArg;
    }
    if (sym_bytes == 2 && ((reinterpret_cast<uintptr_t>(d_a) | reinterpret_cast<uintptr_t>(d_b)) & 1u)) {
        set_error("sl_semi_local_comb_noalloc: 16-bit symbols must be 2-byte aligned");
        return kErrArg;
    }
    int rc = ensure_device();
    if (rc) return rc;
    int device = 0;
    SL_HIP_TRY(hipGetDevice(&device));
    if (device < 0 || device >= kMaxDevices) {
        set_error("sl_semi_local_comb_noalloc: device index out of range");
        return kErrUnsupported;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> lk(g_mu[device]);
    unsigned char *s = nullptr;
    rc = scratch_for(device, &s);
    if (rc) return rc;
    if (m == 0 || n == 0) {
        if (m + n > 0) {
            hipLaunchKernelGGL(noalloc_iota, dim3((unsigned)std::min(1024, (m + n + 255) / 256)), dim3(256), 0, st,
                               d_row_to_col, d_col_to_row, m + n);
            SL_HIP_TRY(hipGetLastError());
            SL_HIP_TRY(hipStreamSynchronize(st));
        }
        return kOk;
    }
    // orientation as in a batch: rows = the shorter string
    const bool rev = m > n;
    PairDesc d{};
    d.m = rev ? n : m;
    d.n = rev ? m : n;
    d.reverse = rev ? 1 : 0;
    // both strings are addressed from the lower of the two pointers
    const uintptr_t pa = reinterpret_cast<uintptr_t>(d_a), pb = reinterpret_cast<uintptr_t>(d_b);
    const uintptr_t base = pa < pb ? pa : pb;
    const int64_t off_a = (int64_t)((pa - base) / (uintptr_t)sym_bytes);
    const int64_t off_b = (int64_t)((pb - base) / (uintptr_t)sym_bytes);
    d.rows_off = rev ? off_b : off_a;
    d.cols_off = rev ? off_a : off_b;
    d.perm_off = 0;
    d.buf_off = 0;
    d.cell = kCellCode;   // raw symbol codes: no remapping pass over the inputs
    d.sigma = 0;
    d.tag_shift = 0;
    const int R = model_rows_one_pair(d.m, d.n, device_cus());
    d.rows_per_lane = R;
    d.last_rows = R;   // one tile height for every pass (the launch's instantiation)
    d.npasses = (int32_t)((d.m + 64 * R - 1) / (64 * R));
    d.item0 = 0;
    const unsigned waves = (unsigned)std::min<int64_t>(d.npasses, (int64_t)device_cus() * 8);
    LaunchCtl *ctl = reinterpret_cast<LaunchCtl *>(s + kOffCtl);
    const unsigned setup_blocks = (unsigned)std::min<int64_t>(1024, (d.n + 255) / 256);
    hipLaunchKernelGGL(noalloc_setup, dim3(setup_blocks), dim3(256), 0, st, s, d, (uint32_t)d.npasses);
    SL_HIP_TRY(hipGetLastError());
    SL_HIP_TRY(launch_comb_batch(sym_bytes, R, waves, st, reinterpret_cast<const PairDesc *>(s + kOffPair),
                                 reinterpret_cast<const WorkItem *>(s + kOffItems), (uint32_t)d.npasses,
                                 reinterpret_cast<const void *>(base), d_row_to_col, d_col_to_row,
                                 reinterpret_cast<uint64_t *>(s + kOffBuf), 1u, ctl, 1u << 23, 0u, nullptr));
    LaunchCtl h{};
    SL_HIP_TRY(hipMemcpyAsync(&h, ctl, sizeof h, hipMemcpyDeviceToHost, st));
    SL_HIP_TRY(hipStreamSynchronize(st));
    if (h.error & 2u) {
        set_error("combing kernel: strand id out of range (internal defect); results are void");
        return kErrHip;
    }
    if (h.error & 1u) {
        set_error("combing kernel: a bounded wait timed out; results are void");
        return kErrTimeout;
    }
    return kOk;
}

}  // extern "C"
