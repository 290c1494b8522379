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

constexpr uint32_t kAttempts = 16;  // attempts of type t at totals[stripe][kAttempts + t]
constexpr uint32_t kCountedTypes = 6;  // LTPG_TYPE_KV .. LTPG_TYPE_DELIVERY

struct DeviceCounters {
    uint32_t range_errors;                            // ops outside their table
    uint32_t pad0[63];
    uint32_t commits[kCounterStripes][64];            // [stripe][type]: last epoch
    unsigned long long totals[kCounterStripes][32];   // [stripe][type]: commits,
                                                      // [stripe][kAttempts + type]: attempts
};

// One key-value batch as the kernels see it (device pointers).
struct KvBatchDev {
    uint32_t n_txns = 0;
    uint32_t n_slots = 0;
    const uint64_t *keys = nullptr;
    const uint32_t *meta = nullptr;
    const uint32_t *tids = nullptr;
    // slot s of transaction t at t * n_slots + s instead of s * n_txns + t
    // (the served form's staging); committed reads stay slot-major
    bool request_major = false;
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
    uint32_t max_scan = 0;              // longest scan accepted (0: none)
    uint32_t read_slot = 1;             // fields of read values per op slot: max(1, max_scan)
    uint32_t *table = nullptr;          // rows * fields * field_stride, row-major
    unsigned long long *regs = nullptr; // 2 * rows: [R, W] per row, epoch-tagged
    uint8_t *doomed = nullptr;          // max_txns: epoch number if waw > 0
    uint8_t *raw_flag = nullptr;        // max_txns: epoch number if a read found a lower W (fast form)
    uint8_t *war_flag = nullptr;        // max_txns: epoch number if a lower reader hit its W (fast form)
    uint32_t *read_values = nullptr;    // max_slots * max_txns * field_words
    uint64_t *bitmap = nullptr;         // ceil(max_txns / 64)
    uint32_t *conflicts = nullptr;      // max_txns
    // library-owned copies of host-submitted batches: two slots, so that
    // the next batch's copy overlaps the current epoch (slot 1 allocated at
    // its first use)
    uint64_t *keys_buf[2] = {nullptr, nullptr};
    uint32_t *meta_buf[2] = {nullptr, nullptr};
    uint32_t *tids_buf[2] = {nullptr, nullptr};
    KvBatchDev cur;
    bool has_batch = false;
    // tuning (ltpg_kv_set_tuning): EXECUTE one lane per op instead of per
    // transaction; R and W registers in two arrays instead of pairs; CHECK
    // and WRITEBACK blocks of 64 transactions instead of 256
    uint32_t exec_per_op = 0;
    uint32_t split_regs = 0;
    uint32_t small_chunks = 0;
    uint32_t lane_copy = 0;
    uint32_t small_epoch = 0;           // epochs of <= LTPG_SMALL_EPOCH_OPS op slots in one kernel
    uint32_t more_items = 0;            // per-op EXECUTE and read marking: twice the ops per lane
    // row-partitioned EXECUTE (fast form): the epoch's register operations
    // binned by row range, each range's W minima kept in LDS; buffers
    // allocated when the shape is first selected (kv_part_prepare)
    uint32_t partitioned = 0;
    struct Part {
        uint32_t n_buckets = 0;
        uint64_t cap_entries = 0;             // register operations an epoch may bin
        uint32_t cap_slices = 0;
        uint32_t *count = nullptr;            // [n_buckets] operations per bucket (zero between epochs)
        uint32_t *start = nullptr;            // [n_buckets + 1] first entry of each bucket
        uint32_t *cursor = nullptr;           // [n_buckets] next free entry while binning
        uint32_t *slices = nullptr;           // [3 * cap_slices + 1] (bucket, begin, end), then their number
        unsigned long long *ent_me = nullptr; // [cap_entries] priority of each binned operation
        uint32_t *ent_key = nullptr;          // [cap_entries] row within its bucket << 1 | is-write
    } part;
    // no kernel has written regs / doomed / raw_flag / war_flag since the
    // last refill ran (the next epoch-tag wrap may skip the refill)
    bool regs_clean = false;
    // hot row (LTPG_NO_HOT_ROW or a row): its ops never register or
    // conflict; committed writes add through hot_acc ([kCounterStripes][fields])
    uint64_t hot_row = LTPG_NO_HOT_ROW;
    uint32_t *hot_acc = nullptr;
};

// ---- order-entry state -------------------------------------------------

constexpr int kTpccTables = LTPG_TPCC_NTABLES;
constexpr uint32_t kNoReads = 51;   // read values per NewOrder
constexpr uint32_t kPayReads = 24;  // read values per Payment
constexpr uint32_t kOsReads = 3;    // read values per Order-Status
constexpr uint32_t kSlReads = 1;    // read values per Stock-Level
constexpr uint32_t kDlReads = 2;    // read values per Delivery

// One table on the device: row-major, `stride` u32 words per row (cols
// rounded up to a multiple of 4 so rows start 16-byte aligned).
struct TpccTable {
    uint32_t *data = nullptr;
    uint64_t rows = 0;
    uint32_t cols = 0;
    uint32_t stride = 0;
};

// A NewOrder / Payment batch as the kernels see it (device pointers; the
// fields of ltpg_neworder_batch / ltpg_payment_batch).
struct NeworderDev {
    uint32_t n = 0;
    uint32_t line_stride = 0;
    const uint32_t *W_ID = nullptr, *D_ID = nullptr, *C_ID = nullptr, *O_ID = nullptr, *N_O_ID = nullptr,
                   *O_OL_CNT = nullptr, *O_OL_ID = nullptr, *OL_I_ID = nullptr, *OL_SUPPLY_W_ID = nullptr,
                   *OL_QUANTITY = nullptr, *TID = nullptr;
};
struct PaymentDev {
    uint32_t n = 0;
    const uint32_t *W_ID = nullptr, *D_ID = nullptr, *C_ID = nullptr, *C_LAST = nullptr, *isName = nullptr,
                   *C_D_ID = nullptr, *C_W_ID = nullptr, *H_AMOUNT = nullptr, *H_ID = nullptr, *TID = nullptr;
};
// Read-only transactions (Order-Status, Stock-Level) of the current batch.
struct ReadOnlyDev {
    uint32_t n_os = 0, n_sl = 0;
    const uint32_t *os_W_ID = nullptr, *os_D_ID = nullptr, *os_C_ID = nullptr, *os_O_ID = nullptr,
                   *os_OL_ID = nullptr, *os_TID = nullptr;
    const uint32_t *sl_W_ID = nullptr, *sl_S_ID = nullptr, *sl_TID = nullptr;
};
// Deliveries of the current batch (the fields of ltpg_delivery_batch).
struct DeliveryDev {
    uint32_t n = 0;
    const uint32_t *O_ID = nullptr, *NO_W_ID = nullptr, *NO_O_ID = nullptr, *TID = nullptr;
};

// Everything the order-entry kernels take, passed by value.
struct TpccDev {
    uint32_t warehouses = 0;
    uint32_t rules = 0;
    TpccTable t[kTpccTables];
    // Registers (tag:8 | TID:32 | pos:24, lower wins), only those a check
    // can observe: CUSTOMER R and W (interleaved, 2 per row), then W of
    // ORDER, STOCK, ORDERLINE, HISTORY and NEWORDER.
    unsigned long long *regs = nullptr;
    uint64_t off_order = 0, off_stock = 0, off_line = 0, off_hist = 0, off_new = 0, n_regs = 0;
    unsigned long long *name_index = nullptr;  // [10W * 3000] sorted C_LAST << 32 | row in district
    // NEWORDER rows: epoch:8 | TID:32 | pos:24 of the last committed inserter
    // (largest); only that inserter writes the row.
    unsigned long long *new_last = nullptr;
    uint8_t *doomed = nullptr;                 // [max_txns] epoch number if waw > 0
    uint64_t *bitmap_no = nullptr, *bitmap_pay = nullptr;
    uint32_t *conf_no = nullptr, *conf_pay = nullptr;
    uint32_t *reads_no = nullptr, *reads_pay = nullptr;  // [51 * n_no], [24 * n_pay] value-major
    uint32_t *cid = nullptr;                             // [n_pay] C_ID each Payment uses
    // read-only transactions: war flags they give NewOrders (epoch number),
    // their bitmaps and read values (value-major)
    uint8_t *war = nullptr;
    uint64_t *bitmap_os = nullptr, *bitmap_sl = nullptr;
    uint32_t *reads_os = nullptr, *reads_sl = nullptr;
    // Deliveries: bitmap and read values (value-major)
    uint64_t *bitmap_dl = nullptr;
    uint32_t *reads_dl = nullptr;
};

struct TpccState {
    TpccDev d;
    // host-submitted batches, two slots as for key-value batches
    uint32_t *no_buf[2] = {nullptr, nullptr};   // packed NEWORDER_QUERY, capacity max_txns
    uint32_t *pay_buf[2] = {nullptr, nullptr};  // packed PAYMENT_QUERY, capacity max_txns
    uint32_t *ro_buf[2] = {nullptr, nullptr};   // read-only batches: 6 + 3 arrays of capacity max_txns
    uint32_t *dl_buf[2] = {nullptr, nullptr};   // Deliveries: 4 arrays of capacity max_txns
    bool host_slot = false;                     // the current epoch's batch went through a host slot
    NeworderDev no;
    PaymentDev pay;
    ReadOnlyDev ro;
    DeliveryDev dl;
    bool has_batch = false;
    bool small_blocks = false;                  // LTPG_TUNE_TPCC_SMALL_BLOCKS
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
    ltpg::TpccState tpcc;
    // Per-step timing: each timed epoch takes LTPG_TIMING_SLOTS + 1 events
    // (start, after CID, after EXECUTE, after CHECK, after the read-out of
    // WRITEBACK, after WRITEBACK).
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    uint32_t ev_used = 0;
    // transport of host-resident batches (two device slots, a copy stream)
    hipStream_t copy_stream = nullptr;
    hipEvent_t xfer_copied[2] = {nullptr, nullptr};
    hipEvent_t xfer_released[2] = {nullptr, nullptr};  // the epochs that used the slot are done
    bool xfer_released_set[2] = {false, false};
    int xfer_slot = 1;  // slot of the current host batch (the next one takes the other)
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
hipError_t launch_grid_barrier(uint32_t blocks, unsigned int *count, unsigned int *gen, uint32_t reps,
                               uint32_t spin_limit, unsigned int *err, hipStream_t s);
hipError_t launch_boundary(uint32_t blocks, unsigned int *p, hipStream_t s);
// Step latencies of the served device loop's epoch chain (ticks): out[0..4].
hipError_t launch_latency_probe(uint64_t *host, uint32_t chain_words, unsigned long long *out, uint32_t reps,
                                uint32_t host_reps, hipStream_t s);
hipError_t launch_gather_calibration(const uint32_t *src, uint64_t slots, uint32_t stride_w, uint32_t fw,
                                     uint32_t *dst, uint64_t n, hipStream_t s);
hipError_t launch_calibration(unsigned long long *a, uint64_t len, uint64_t n, int atomics,
                              unsigned long long *sink, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s);
hipError_t launch_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t s);
hipError_t launch_copy_u64(uint64_t *dst, const uint64_t *src, uint64_t n, hipStream_t s);
// n u32 words; src may be pinned host memory
hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t n, hipStream_t s);
// Up to four bitmaps to copy to pinned host memory (n[k] words from src[k]
// to dst[k]; n[k] = 0 for unused entries).
struct PublishSet {
    const uint64_t *src[5];
    uint64_t *dst[5];
    uint32_t n[5];
};
// copy the bitmaps of p, then store seq to *seq_out (pinned host memory)
// with a system-scope release
hipError_t launch_publish(const PublishSet &p, uint64_t *seq_out, uint64_t seq, hipStream_t s);
hipError_t launch_stamp(unsigned long long *out, hipStream_t s);

// Registers back to "no registrant" (all-ones), doomed flags and NEWORDER
// inserters to 0, on the context's stream (api.cpp).  Epoch-tag wrap and the
// start of a captured batch sequence use it.
int refill_registers(ltpg_ctx *ctx);
int kv_part_prepare(ltpg_ctx *ctx);
hipError_t launch_kv_fill_reference(const KvState &kv, hipStream_t s);
// rows idx[0 .. n-1] (device, each < kv.rows) of every field, packed, to out (device)
hipError_t launch_kv_gather_rows(const KvState &kv, const uint64_t *idx, uint64_t n, uint32_t *out, hipStream_t s);
hipError_t launch_repack(const uint32_t *src, uint32_t *dst, uint64_t n_fields, uint32_t fw,
                         uint32_t fsw, int to_strided, hipStream_t s);
hipError_t launch_kv_transpose(const uint32_t *src, uint32_t *dst, uint64_t rows, uint32_t cols,
                               int to_row_major, hipStream_t s);
// Row-partitioned EXECUTE limits: rows per bucket (1 << kPartShift), buckets,
// and operations per slice (one workgroup).
constexpr uint32_t kPartShift = 13;
constexpr uint32_t kPartMaxBuckets = 16384;
constexpr uint32_t kPartSlice = 32768;
hipError_t launch_kv_execute(const KvState &kv, const KvBatchDev &b, const EpochTag &e, bool counts,
                             DeviceCounters *ctr, hipStream_t s, bool zero_counts = true);
hipError_t launch_kv_check(const KvState &kv, const KvBatchDev &b, const EpochTag &e, bool counts,
                           DeviceCounters *ctr, hipStream_t s);
hipError_t launch_kv_readout(const KvState &kv, const KvBatchDev &b, hipStream_t s);
hipError_t launch_kv_writeback(const KvState &kv, const KvBatchDev &b, uint32_t values,
                               hipStream_t s);
// A whole fast-form epoch of at most LTPG_SMALL_EPOCH_OPS op slots in one
// single-workgroup kernel (a hot row only with at most 1024 fields; rows <
// 2^32 - 1).  With seq_out it
// also publishes: the bitmap to host_bits, then seq with a release.
bool kv_small_epoch_ok(const KvState &kv, const KvBatchDev &b);
hipError_t launch_kv_small_epoch(const KvState &kv, const KvBatchDev &b, uint32_t values, DeviceCounters *ctr,
                                 uint64_t *host_bits, uint64_t *seq_out, uint64_t seq, hipStream_t s);
// The current batch's epoch phase by phase, a stream synchronisation after
// each (the reference's driver loop), then ltpg_sync.
int run_epoch_stepwise(ltpg_ctx *ctx);

// ---- served key-value epochs in one persistent workgroup -----------------
//
// The served form's key-value lane can run as one single-workgroup kernel
// that forms its own epochs: it takes arrivals from a request ring in
// pinned host memory, keeps the aborted requests of each epoch in LDS as the
// head of the next epoch (their priority order is their arrival order), and
// publishes one record per epoch into a done ring in pinned host memory.
// The kernel exits when asked, when idle, when the done ring stays full, or
// after a time slice; the host relaunches it when work is pending, and the
// aborted requests it carried wait in device memory in between.

// Control block in pinned, fine-grained host memory.  The host writes the
// first cache line, the kernel the second.
struct ServeCtl {
    uint64_t tail;       // host: requests published in the ring (release)
    uint64_t stop;       // host: nonzero = exit at the next epoch boundary
    uint64_t done_head;  // host: done-ring words consumed (release)
    uint64_t pad0[5];
    uint64_t head;       // kernel: requests taken from the ring (release)
    uint64_t done_tail;  // kernel: done-ring words published (release)
    uint64_t running;    // host sets 1 before a launch, the kernel clears it last
    uint64_t exit_code;  // why the last launch returned (kServeExit*)
    uint64_t epochs;     // epochs run, all launches
    uint64_t n_carry;    // requests carried to the next epoch
    uint64_t cap;        // the next epoch's transaction cap
    uint64_t pad1;
    // wall-clock ticks inside epochs (all launches)
    uint64_t busy_ticks;
    // profile (ServeLoopArgs.profile): wall-clock ticks spent in each part
    // of the epochs, all launches, and the number of epochs profiled
    uint64_t prof[10];
    uint64_t prof_epochs;
};
// Parts of an epoch timed by the profile: control read, admission, EXECUTE,
// read marking, CHECK, write folding, read-out and write-back, publication,
// requeue, idle or waiting.
constexpr int kServeProfParts = 10;
constexpr uint64_t kServeExitStop = 1;   // stop requested
constexpr uint64_t kServeExitIdle = 2;   // nothing to do for idle_ticks
constexpr uint64_t kServeExitStall = 3;  // the done ring stayed full for stall_ticks
constexpr uint64_t kServeExitSlice = 4;  // ran for slice_ticks

// Done-ring record of one epoch: a header word (bits 0-15 transactions,
// 16-31 commits, 32-63 epoch number), then one word per member in batch
// order: its arrival sequence number, bit 63 set if it committed.
constexpr uint64_t kDoneCommitted = 1ull << 63;

struct ServeLoopArgs {
    // table
    uint32_t *table;
    uint64_t rows;
    uint32_t fields, fw, fsw, max_scan, L, values;
    uint64_t hot;
    // requests: kv_ops slots each, request-major; ring slot of arrival q is
    // q & ring_mask
    uint32_t S;
    const uint64_t *ring_keys;
    const uint32_t *ring_meta;
    uint64_t ring_mask;
    ServeCtl *ctl;
    uint64_t *done;       // done ring (pinned)
    uint64_t done_mask;
    uint32_t *reads;      // or null: committed reads, S * L * fw words per done-ring word
    // carried requests between launches (device memory)
    uint64_t *carry_keys;
    uint32_t *carry_meta;
    uint64_t *carry_seq;
    // epoch sizing: cap in [floor, cap_max], growth x the last epoch's commits
    uint32_t cap_max, floor, growth;
    // bounded waits, in wall-clock ticks
    uint64_t idle_ticks, stall_ticks, slice_ticks;
    DeviceCounters *ctr;
    uint32_t profile;     // nonzero: time the parts of each epoch into ctl->prof
};
bool kv_serve_loop_ok(const KvState &kv, uint32_t S);
hipError_t launch_kv_serve_loop(const ServeLoopArgs &a, hipStream_t s);

// Order-entry launchers (tpcc_kernels.hip).
hipError_t launch_tpcc_populate(const TpccDev &d, uint64_t seed, hipStream_t s);
hipError_t launch_tpcc_name_index(const TpccDev &d, hipStream_t s);
hipError_t launch_tpcc_relayout(const uint32_t *src, uint32_t *dst, uint64_t rows, uint32_t cols,
                                uint32_t stride, uint32_t layout, int to_device, hipStream_t s);
// `small`: 64-thread blocks (LTPG_TUNE_TPCC_SMALL_BLOCKS) instead of 256.
hipError_t launch_tpcc_cid(const TpccDev &d, const PaymentDev &pay, bool small, hipStream_t s);
hipError_t launch_tpcc_execute(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay,
                               const EpochTag &e, bool counts, bool small, DeviceCounters *ctr, hipStream_t s);
hipError_t launch_tpcc_readonly(const TpccDev &d, const ReadOnlyDev &ro, uint32_t base, const EpochTag &e,
                                DeviceCounters *ctr, hipStream_t s);
// Deliveries (base = their first batch position): EXECUTE registrations,
// CHECK (commit, bitmap, read values, counts) and WRITEBACK.
hipError_t launch_tpcc_delivery_execute(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                        DeviceCounters *ctr, hipStream_t s);
hipError_t launch_tpcc_delivery_check(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                      DeviceCounters *ctr, hipStream_t s);
hipError_t launch_tpcc_delivery_writeback(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                          uint32_t values, hipStream_t s);
hipError_t launch_tpcc_check(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e,
                             bool counts, bool ro, bool dl, bool small, DeviceCounters *ctr, hipStream_t s);
hipError_t launch_tpcc_writeback(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay,
                                 const EpochTag &e, uint32_t values, bool small, hipStream_t s);

}  // namespace ltpg
