// ltpg_internal.h - context state and helpers shared by the host side of
// libltpg.  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "ltpg.h"

namespace ltpg {

// Thread-local message for ltpg_last_error().
void set_error(const std::string &msg);

#define LTPG_HIP_TRY(expr)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            ::ltpg::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));             \
            return e_ == hipErrorOutOfMemory ? LTPG_ERR_NOMEM : LTPG_ERR_DEVICE;              \
        }                                                                                     \
    } while (0)

#define LTPG_FAIL(code, msg)                                                                 \
    do {                                                                                      \
        ::ltpg::set_error(msg);                                                               \
        return (code);                                                                        \
    } while (0)

// Device-resident counters written by kernels.  Commit / attempt counts are
// striped: kernels add to stripe (block % kCounterStripes) and the host sums
// the stripes, so no single address takes an atomic from every block.
constexpr int kCounterStripes = 64;

struct DeviceCounters {
    uint32_t range_errors;                            // ops outside their table
    uint32_t pad0[63];
    uint32_t commits[kCounterStripes][64];            // [stripe][type]: last epoch
    unsigned long long totals[kCounterStripes][32];   // [stripe][type]: commits,
                                                      // [stripe][3 + type]: attempts
};

// One key-value batch as the kernels see it (device pointers).
struct KvBatchDev {
    uint32_t n_txns = 0;
    uint32_t n_slots = 0;
    const uint64_t *keys = nullptr;
    const uint32_t *meta = nullptr;
    const uint32_t *tids = nullptr;
};

// Register tag of the current epoch: registers hold tag:8 | TID:32 | pos:24
// with tag = 255 - epoch, epochs 1..255.
struct EpochTag {
    uint32_t tag;
    uint8_t epoch;
};

struct KvState {
    uint64_t rows = 0;
    uint32_t fields = 0;
    uint32_t field_words = 0;
    uint32_t field_stride = 0;          // words between field starts (>= field_words)
    uint32_t max_slots = 0;
    uint32_t *table = nullptr;          // rows * fields * field_stride, row-major
    unsigned long long *regs = nullptr; // 2 * rows: [R, W] per row, epoch-tagged
    uint8_t *doomed = nullptr;          // max_txns: epoch number if waw > 0
    uint32_t *read_values = nullptr;    // max_slots * max_txns * field_words
    uint64_t *bitmap = nullptr;         // ceil(max_txns / 64)
    uint32_t *conflicts = nullptr;      // max_txns
    // library-owned copy of host-submitted batches
    uint64_t *keys_buf = nullptr;
    uint32_t *meta_buf = nullptr;
    uint32_t *tids_buf = nullptr;
    KvBatchDev cur;
    bool has_batch = false;
};

}  // namespace ltpg

struct ltpg_ctx {
    ltpg_config cfg;
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    uint32_t epoch = 1;  // 1..255; 0 owns the all-ones fill value of the registers
    ltpg::DeviceCounters *counters = nullptr;
    ltpg::KvState kv;
    // Per-step timing: each timed epoch takes LTPG_TIMING_SLOTS + 1 events
    // (start, after CID, after EXECUTE, after CHECK, after the read-out of
    // WRITEBACK, after WRITEBACK).
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    uint32_t ev_used = 0;
};

namespace ltpg {

// ---- host helpers shared by the API translation units --------------------

template <typename T>
int dev_alloc(T **p, uint64_t count, const char *what)
{
    *p = nullptr;
    if (count == 0) return LTPG_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        set_error(std::string("hipMalloc(") + what + ", " + std::to_string(count * sizeof(T)) +
                  " bytes): " + hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? LTPG_ERR_NOMEM : LTPG_ERR_DEVICE;
    }
    return LTPG_OK;
}

template <typename T>
void dev_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline int check_ctx(ltpg_ctx *ctx)
{
    if (!ctx) LTPG_FAIL(LTPG_ERR_INVALID, "null context");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return LTPG_OK;
}

// Kernel launchers (kv_kernels.hip).  All asynchronous on `s`.
hipError_t launch_reset_epoch_counters(DeviceCounters *ctr, hipStream_t s);
hipError_t launch_copy_u128(void *dst, const void *src, uint64_t n16, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t s);
hipError_t launch_kv_fill_reference(const KvState &kv, hipStream_t s);
hipError_t launch_repack(const uint32_t *src, uint32_t *dst, uint64_t n_fields, uint32_t fw,
                         uint32_t fsw, int to_strided, hipStream_t s);
hipError_t launch_kv_transpose(const uint32_t *src, uint32_t *dst, uint64_t rows, uint32_t cols,
                               int to_row_major, hipStream_t s);
hipError_t launch_kv_execute(const KvState &kv, const KvBatchDev &b, const EpochTag &e,
                             DeviceCounters *ctr, hipStream_t s);
hipError_t launch_kv_check(const KvState &kv, const KvBatchDev &b, const EpochTag &e, bool counts,
                           DeviceCounters *ctr, hipStream_t s);
hipError_t launch_kv_readout(const KvState &kv, const KvBatchDev &b, hipStream_t s);
hipError_t launch_kv_writeback(const KvState &kv, const KvBatchDev &b, uint32_t values,
                               hipStream_t s);

}  // namespace ltpg
