// common.h -- shared host/device definitions of the semi-local LCS engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace slhip {

// Status codes of the C ABI (mirrored in include/semilocal_hip.h).
enum Status : int {
    kOk = 0,
    kErrArg = 1,        // bad argument (null pointer, negative size, ...)
    kErrHip = 2,        // a HIP runtime call failed
    kErrNoMem = 3,      // device allocation failed / admission refused
    kErrTimeout = 4,    // a kernel's bounded wait ran out (reported by the kernel)
    kErrUnsupported = 5 // input outside what the engine supports (e.g. > 65535 symbols)
};

void set_error(const std::string &msg);
const char *last_error();
int ensure_device(int device = -1);   // selects (and on first use initialises) the HIP device
int device_cus();                     // compute units of the selected device
// Run tags of the tagged hand-off buffers (combing strands, bit-parallel and
// Hyyro carries).  Every run of any handle draws the next value of one
// process-wide 64-bit serial; its 16-bit tag is (serial mod 65535) + 1, never
// 0 (a cleared entry).  Two serials share a tag only 65,535 or more runs
// apart, so a handle keeps `cleared_at`, the serial when its buffer was last
// cleared: every tag in the buffer then belongs to a later serial, and
// next_run_tag() asks for a clear before the tag could repeat one of them.
uint64_t run_serial_now();
bool next_run_tag(uint64_t &cleared_at, uint32_t *tag);   // true: clear the buffer before this run

#define SL_HIP_TRY(expr)                                                              \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::slhip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));    \
            return ::slhip::kErrHip;                                                  \
        }                                                                             \
    } while (0)

// One pair of the ragged batch after orientation: the kernel always combs
// rows = the shorter string (length m) against cols = the longer (length n),
// and `reverse` records whether that swapped the caller's (a, b).
struct PairDesc {
    int64_t rows_off;   // offset of the row string in the symbol pool
    int64_t cols_off;   // offset of the column string in the symbol pool
    int64_t perm_off;   // offset of this pair's (m+n) outputs
    int64_t buf_off;    // offset of this pair's inter-pass strand buffer (n entries)
    int32_t m, n;       // rows, cols (m <= n)
    int32_t reverse;    // 1 if the caller's a was the longer string
    int32_t item0;      // queue index of pass 0 of this pair
    int32_t npasses;    // ceil(m / rows_per_pass)
    int32_t rows_per_lane;  // R: rows held by one lane (rows_per_pass = 64 R)
    int32_t last_rows;  // tile height of the last pass (<= R: the fewest rows that cover the rest)
    int32_t cell;       // how the cell tests symbols: kCellCode, kCellOneHot or kCellTable
    int32_t tag_shift;  // bits of the largest strand id (one-hot tags start here)
    int32_t sigma;      // distinct symbols of the pair (dense codes 0..sigma-1)
    int32_t prio;       // 1: the pair's passes run at raised wave priority (a critical chain)
};

// Symbol test of the combing cell, chosen per pair (and per tile height R).
constexpr int32_t kCellCode = 0;     // codes compared in the cell: 4 VALU ops
constexpr int32_t kCellOneHot = 1;   // one-hot tags above the id bits (bits + sigma <= 32): 3 ops
constexpr int32_t kCellTable = 2;    // per-lane LDS table of row-match masks indexed by the
                                     // column symbol (sigma <= kTableMaxSymbols, bits <= 32 - R):
                                     // 3 ops + one LDS read per step
constexpr int32_t kTableMaxSymbols = 32;

// Queue entry: (pair, pass).  Entries of one pair are consecutive and in pass
// order, so a pass only ever waits on an entry handed out before it.
struct WorkItem {
    int32_t pair;
    int32_t pass;
};

// Source of device memory for a combing batch: hipMalloc by default, or a
// caller's arena (the request loop carves every batch out of one allocation).
struct DeviceArena {
    virtual ~DeviceArena() = default;
    virtual hipError_t alloc(void **p, size_t bytes) = 0;
    virtual void release(void *p) = 0;
};

// One bit-parallel problem (kernel (b)): M 32-bit row-words of the reversed
// a, N column-words of b.  A pass is kBitRowLanes row-words: one wave whose
// lane 0 feeds the column strands and whose lanes 1..63 own a row-word each.
constexpr int kBitRowLanes = 63;
struct BitDesc {
    const uint32_t *a_rev;   // M words: word j = a_rev word j (32-bit view)
    const uint32_t *b;       // N words
    uint64_t *buf;           // N tagged entries: the column words a pass hands to the next
    uint32_t *dis;           // left strands leaving at the bottom (added by the last pass)
    int32_t M, N;
    int32_t npasses;         // ceil(M / kBitRowLanes)
    uint32_t epoch;
};

// Device-side control block of one launch.
struct LaunchCtl {
    uint32_t ticket;    // next queue entry to hand out
    uint32_t error;     // bit 0: a bounded wait ran out; bit 1: a strand id failed the bounds check
};

}  // namespace slhip
