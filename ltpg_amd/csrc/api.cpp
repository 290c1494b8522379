// api.cpp - the extern "C" entry points of libltpg (include/ltpg.h).
//
// Host-side counterpart of the reference's Database / Query lifecycle
// (LTPG/TPCC-allopt-final/Database.h:37-102, Query.h:330-383,532-631;
// LTPG/YCSB/0/Database.h:27-50, Query.h:145-225) and of the per-epoch kernel
// launches of tp.cu:136-261 / ycsb.cu:78-152.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "ltpg_internal.h"

namespace ltpg {

namespace {
thread_local std::string g_last_error;
constexpr uint32_t kEv = LTPG_TIMING_SLOTS + 1;  // events per timed epoch
}

void set_error(const std::string &msg) { g_last_error = msg; }

namespace {

void free_kv(KvState &kv)
{
    dev_free(kv.table);
    dev_free(kv.regs);
    dev_free(kv.doomed);
    dev_free(kv.raw_flag);
    dev_free(kv.war_flag);
    dev_free(kv.read_values);
    dev_free(kv.bitmap);
    dev_free(kv.conflicts);
    for (int k = 0; k < 2; k++) {
        dev_free(kv.keys_buf[k]);
        dev_free(kv.meta_buf[k]);
        dev_free(kv.tids_buf[k]);
    }
    dev_free(kv.hot_acc);
    dev_free(kv.part.count);
    dev_free(kv.part.start);
    dev_free(kv.part.cursor);
    dev_free(kv.part.slices);
    dev_free(kv.part.ent_me);
    dev_free(kv.part.ent_key);
}

int init_kv(ltpg_ctx *ctx)
{
    const ltpg_config &c = ctx->cfg;
    KvState &kv = ctx->kv;
    if (c.kv_rows == 0 || c.kv_fields == 0) LTPG_FAIL(LTPG_ERR_INVALID, "kv_rows and kv_fields must be > 0");
    if (c.kv_field_bytes == 0 || c.kv_field_bytes % 4 != 0 || c.kv_field_bytes > 1024)
        LTPG_FAIL(LTPG_ERR_INVALID, "kv_field_bytes must be a multiple of 4 in [4, 1024]");
    if (c.kv_fields > 65536) LTPG_FAIL(LTPG_ERR_INVALID, "kv_fields must be <= 65536");
    if (c.kv_max_slots == 0 || c.kv_max_slots > 64)
        LTPG_FAIL(LTPG_ERR_INVALID, "kv_max_slots must be in [1, 64]");
    if (c.kv_hot_row != LTPG_NO_HOT_ROW && c.kv_hot_row >= c.kv_rows)
        LTPG_FAIL(LTPG_ERR_INVALID, "kv_hot_row must be LTPG_NO_HOT_ROW or a row of the table");
    kv.hot_row = c.kv_hot_row;
    kv.rows = c.kv_rows;
    kv.fields = c.kv_fields;
    kv.field_words = c.kv_field_bytes / 4;
    // Fields of 64 bytes or more start on 64-byte boundaries, so a random
    // field read touches the fewest 64-byte memory segments.
    kv.field_stride = kv.field_words >= 16 ? (kv.field_words + 15) / 16 * 16 : kv.field_words;
    kv.max_slots = c.kv_max_slots;
    if (c.kv_max_scan > LTPG_KV_MAX_SCAN) LTPG_FAIL(LTPG_ERR_INVALID, "kv_max_scan must be at most 256");
    kv.max_scan = c.kv_max_scan;
    kv.read_slot = c.kv_max_scan > 1 ? c.kv_max_scan : 1;
    const uint64_t B = c.max_txns, S = c.kv_max_slots;
    int rc;
    if ((rc = dev_alloc(&kv.table, kv.rows * kv.fields * kv.field_stride, "kv table"))) return rc;
    if ((rc = dev_alloc(&kv.regs, kv.rows * 2, "kv registers"))) return rc;
    if ((rc = dev_alloc(&kv.doomed, B, "kv doomed flags"))) return rc;
    if ((rc = dev_alloc(&kv.raw_flag, B, "kv raw flags"))) return rc;
    if ((rc = dev_alloc(&kv.war_flag, B, "kv war flags"))) return rc;
    if ((rc = dev_alloc(&kv.read_values, S * B * kv.read_slot * kv.field_words, "kv read values"))) return rc;
    if ((rc = dev_alloc(&kv.bitmap, (B + 63) / 64, "kv bitmap"))) return rc;
    if ((rc = dev_alloc(&kv.conflicts, B, "kv conflicts"))) return rc;
    if ((rc = dev_alloc(&kv.keys_buf[0], S * B, "kv keys"))) return rc;
    if ((rc = dev_alloc(&kv.meta_buf[0], S * B, "kv meta"))) return rc;
    if ((rc = dev_alloc(&kv.tids_buf[0], B, "kv tids"))) return rc;
    if (kv.hot_row != LTPG_NO_HOT_ROW) {
        if ((rc = dev_alloc(&kv.hot_acc, (uint64_t)kCounterStripes * kv.fields, "hot-row accumulators"))) return rc;
        LTPG_HIP_TRY(hipMemsetAsync(kv.hot_acc, 0, (uint64_t)kCounterStripes * kv.fields * 4, ctx->stream));
    }
    // Registers start as all-ones: tag ~0 belongs to epoch 0, which is never
    // current (epochs start at 1), so every row reads as "no registrant".
    LTPG_HIP_TRY(launch_fill_u64(kv.regs, kv.rows * 2, ~0ull, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(kv.doomed, B, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(kv.raw_flag, B, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(kv.war_flag, B, 0, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(kv.table, 0, kv.rows * kv.fields * kv.field_stride * 4, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(kv.bitmap, 0, (B + 63) / 64 * 8, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(kv.conflicts, 0, B * 4, ctx->stream));
    return LTPG_OK;
}

// ---- order-entry state ------------------------------------------------------

// Rows per table at W warehouses (TPCC-allopt-final/Predefine.h:48-56) and
// columns (Predefine.h:57-65); device rows are padded to a multiple of 4.
const uint64_t kRowsPerW[kTpccTables] = {1, 10, 30000, 30000, 30000, 30000, 450000, 100000, 0};
const uint32_t kTpccCols[kTpccTables] = {9, 11, 21, 5, 3, 8, 10, 17, 5};

uint64_t tpcc_rows(uint32_t W, uint32_t t) { return t == LTPG_TPCC_ITEM ? 100000 : kRowsPerW[t] * W; }

void free_tpcc(TpccState &tp)
{
    for (auto &t : tp.d.t) dev_free(t.data);
    dev_free(tp.d.regs);
    dev_free(tp.d.name_index);
    dev_free(tp.d.new_last);
    dev_free(tp.d.doomed);
    dev_free(tp.d.bitmap_no);
    dev_free(tp.d.bitmap_pay);
    dev_free(tp.d.conf_no);
    dev_free(tp.d.conf_pay);
    dev_free(tp.d.reads_no);
    dev_free(tp.d.reads_pay);
    dev_free(tp.d.cid);
    for (int k = 0; k < 2; k++) {
        dev_free(tp.no_buf[k]);
        dev_free(tp.pay_buf[k]);
        dev_free(tp.ro_buf[k]);
        dev_free(tp.dl_buf[k]);
    }
    dev_free(tp.d.war);
    dev_free(tp.d.bitmap_os);
    dev_free(tp.d.bitmap_sl);
    dev_free(tp.d.reads_os);
    dev_free(tp.d.reads_sl);
    dev_free(tp.d.bitmap_dl);
    dev_free(tp.d.reads_dl);
}

int init_tpcc(ltpg_ctx *ctx)
{
    const ltpg_config &c = ctx->cfg;
    TpccState &tp = ctx->tpcc;
    TpccDev &d = tp.d;
    const uint32_t W = c.tpcc_warehouses;
    if (W == 0 || W > 65536) LTPG_FAIL(LTPG_ERR_INVALID, "tpcc_warehouses must be in [1, 65536]");
    d.warehouses = W;
    d.rules = c.rules;
    int rc;
    for (uint32_t t = 0; t < (uint32_t)kTpccTables; t++) {
        TpccTable &tb = d.t[t];
        tb.rows = tpcc_rows(W, t);
        tb.cols = kTpccCols[t];
        tb.stride = (tb.cols + 3) / 4 * 4;
        if ((rc = dev_alloc(&tb.data, tb.rows * tb.stride, "order-entry table"))) return rc;
        LTPG_HIP_TRY(hipMemsetAsync(tb.data, 0, tb.rows * tb.stride * 4, ctx->stream));
    }
    const uint64_t cust = d.t[LTPG_TPCC_CUSTOMER].rows;
    d.off_order = 2 * cust;
    d.off_stock = d.off_order + d.t[LTPG_TPCC_ORDER].rows;
    d.off_line = d.off_stock + d.t[LTPG_TPCC_STOCK].rows;
    d.off_hist = d.off_line + d.t[LTPG_TPCC_ORDERLINE].rows;
    d.off_new = d.off_hist + d.t[LTPG_TPCC_HISTORY].rows;
    d.n_regs = d.off_new + d.t[LTPG_TPCC_NEWORDER].rows;
    const uint64_t B = c.max_txns;
    if ((rc = dev_alloc(&d.regs, d.n_regs, "order-entry registers"))) return rc;
    if ((rc = dev_alloc(&d.name_index, cust, "by-name index"))) return rc;
    if ((rc = dev_alloc(&d.new_last, d.t[LTPG_TPCC_NEWORDER].rows, "NEWORDER inserters"))) return rc;
    if ((rc = dev_alloc(&d.doomed, B, "doomed flags"))) return rc;
    if ((rc = dev_alloc(&d.bitmap_no, (B + 63) / 64, "NewOrder bitmap"))) return rc;
    if ((rc = dev_alloc(&d.bitmap_pay, (B + 63) / 64, "Payment bitmap"))) return rc;
    if ((rc = dev_alloc(&d.conf_no, B, "NewOrder conflicts"))) return rc;
    if ((rc = dev_alloc(&d.conf_pay, B, "Payment conflicts"))) return rc;
    if ((rc = dev_alloc(&d.reads_no, (uint64_t)kNoReads * B, "NewOrder read values"))) return rc;
    if ((rc = dev_alloc(&d.reads_pay, (uint64_t)kPayReads * B, "Payment read values"))) return rc;
    if ((rc = dev_alloc(&d.cid, B, "Payment C_ID"))) return rc;
    if ((rc = dev_alloc(&tp.no_buf[0], 53 * B, "NewOrder batch"))) return rc;
    if ((rc = dev_alloc(&tp.pay_buf[0], 10 * B, "Payment batch"))) return rc;
    if ((rc = dev_alloc(&tp.ro_buf[0], 9 * B, "read-only batches"))) return rc;
    if ((rc = dev_alloc(&d.war, B, "read-only war flags"))) return rc;
    if ((rc = dev_alloc(&d.bitmap_os, (B + 63) / 64, "Order-Status bitmap"))) return rc;
    if ((rc = dev_alloc(&d.bitmap_sl, (B + 63) / 64, "Stock-Level bitmap"))) return rc;
    if ((rc = dev_alloc(&d.reads_os, (uint64_t)kOsReads * B, "Order-Status read values"))) return rc;
    if ((rc = dev_alloc(&d.reads_sl, (uint64_t)kSlReads * B, "Stock-Level read values"))) return rc;
    if ((rc = dev_alloc(&d.bitmap_dl, (B + 63) / 64, "Delivery bitmap"))) return rc;
    if ((rc = dev_alloc(&d.reads_dl, (uint64_t)kDlReads * B, "Delivery read values"))) return rc;
    LTPG_HIP_TRY(launch_fill_u8(d.war, B, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u64(d.regs, d.n_regs, ~0ull, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(d.doomed, B, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u64(d.new_last, d.t[LTPG_TPCC_NEWORDER].rows, 0, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(d.bitmap_no, 0, (B + 63) / 64 * 8, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(d.bitmap_pay, 0, (B + 63) / 64 * 8, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(d.conf_no, 0, B * 4, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(d.conf_pay, 0, B * 4, ctx->stream));
    LTPG_HIP_TRY(hipMemsetAsync(d.cid, 0, B * 4, ctx->stream));
    LTPG_HIP_TRY(launch_tpcc_name_index(d, ctx->stream));
    return LTPG_OK;
}

}  // namespace

// Buffers of the row-partitioned EXECUTE, allocated on its first selection
// (outside any graph capture: the batched form prepares them before it
// captures).
int kv_part_prepare(ltpg_ctx *ctx)
{
    KvState &kv = ctx->kv;
    if (kv.part.ent_me) return LTPG_OK;
    const uint64_t nb = (kv.rows + (1ull << kPartShift) - 1) >> kPartShift;
    const uint64_t cap = (uint64_t)kv.max_slots * ctx->cfg.max_txns * kv.read_slot;
    if (nb > kPartMaxBuckets || cap >= (1ull << 32))
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "row-partitioned EXECUTE: too many rows or operations per epoch");
    KvState::Part p;
    p.n_buckets = (uint32_t)nb;
    p.cap_entries = cap;
    p.cap_slices = (uint32_t)(nb + cap / kPartSlice + 1);
    int rc;
    if ((rc = dev_alloc(&p.count, nb, "partition counts")) || (rc = dev_alloc(&p.start, nb + 1, "partition starts")) ||
        (rc = dev_alloc(&p.cursor, nb, "partition cursors")) ||
        (rc = dev_alloc(&p.slices, 3ull * p.cap_slices + 1, "partition slices")) ||
        (rc = dev_alloc(&p.ent_me, cap, "partitioned operations")) ||
        (rc = dev_alloc(&p.ent_key, cap, "partitioned operation rows"))) {
        dev_free(p.count);
        dev_free(p.start);
        dev_free(p.cursor);
        dev_free(p.slices);
        dev_free(p.ent_me);
        dev_free(p.ent_key);
        return rc;
    }
    // the bucket sizes start at zero; the binning leaves them at zero
    LTPG_HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t *>(p.count), nb * 4, 0, ctx->stream));
    kv.part = p;
    return LTPG_OK;
}

int refill_registers(ltpg_ctx *ctx)
{
    if (ctx->kv.regs) {
        LTPG_HIP_TRY(launch_fill_u64(ctx->kv.regs, ctx->kv.rows * 2, ~0ull, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u8(ctx->kv.doomed, ctx->cfg.max_txns, 0, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u8(ctx->kv.raw_flag, ctx->cfg.max_txns, 0, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u8(ctx->kv.war_flag, ctx->cfg.max_txns, 0, ctx->stream));
    }
    if (ctx->tpcc.d.regs) {
        LTPG_HIP_TRY(launch_fill_u64(ctx->tpcc.d.regs, ctx->tpcc.d.n_regs, ~0ull, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u8(ctx->tpcc.d.doomed, ctx->cfg.max_txns, 0, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u8(ctx->tpcc.d.war, ctx->cfg.max_txns, 0, ctx->stream));
        LTPG_HIP_TRY(launch_fill_u64(ctx->tpcc.d.new_last, ctx->tpcc.d.t[LTPG_TPCC_NEWORDER].rows, 0, ctx->stream));
    }
    return LTPG_OK;
}

namespace {

// Advance the epoch (1..255).  At wrap-around every register goes back to
// all-ones (the tag of epoch 0, never current) and the doomed flags to 0 --
// unless no key-value kernel has written them since the last refill ran
// (only single-workgroup epochs, which keep their registrations in LDS).
// A refill recorded into a graph capture has not run yet, so it leaves the
// registers marked as written.
int next_epoch(ltpg_ctx *ctx)
{
    ctx->epoch += 1;
    if (ctx->epoch > 255) {
        KvState &kv = ctx->kv;
        const bool skip = kv.regs && !ctx->tpcc.d.regs && kv.regs_clean;
        if (!skip) {
            int rc = refill_registers(ctx);
            if (rc) return rc;
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            LTPG_HIP_TRY(hipStreamIsCapturing(ctx->stream, &cs));
            kv.regs_clean = cs == hipStreamCaptureStatusNone;
        }
        ctx->epoch = 1;
    }
    return LTPG_OK;
}

EpochTag epoch_tag(const ltpg_ctx *ctx)
{
    EpochTag e;
    e.epoch = (uint8_t)ctx->epoch;
    e.tag = 255u - ctx->epoch;
    return e;
}

// Host <-> device table copies between the caller's dense row-major layout
// (field_words apart) and the device's strided one, in chunks through a
// device staging buffer.
int kv_copy_table(ltpg_ctx *ctx, void *host, bool to_device)
{
    KvState &kv = ctx->kv;
    const uint64_t n_fields = kv.rows * kv.fields;
    const uint64_t fw = kv.field_words, fsw = kv.field_stride;
    if (fw == fsw) {
        if (to_device)
            LTPG_HIP_TRY(hipMemcpyAsync(kv.table, host, n_fields * fw * 4, hipMemcpyHostToDevice, ctx->stream));
        else
            LTPG_HIP_TRY(hipMemcpyAsync(host, kv.table, n_fields * fw * 4, hipMemcpyDeviceToHost, ctx->stream));
        LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return LTPG_OK;
    }
    const uint64_t chunk_fields = std::max<uint64_t>(1, (256ull << 20) / (fw * 4));
    uint32_t *tmp = nullptr;
    int rc = dev_alloc(&tmp, std::min(chunk_fields, n_fields) * fw, "repack staging");
    if (rc) return rc;
    hipError_t e = hipSuccess;
    for (uint64_t f0 = 0; f0 < n_fields && e == hipSuccess; f0 += chunk_fields) {
        const uint64_t nf = std::min(chunk_fields, n_fields - f0);
        uint32_t *h = static_cast<uint32_t *>(host) + f0 * fw;
        if (to_device) {
            e = hipMemcpyAsync(tmp, h, nf * fw * 4, hipMemcpyHostToDevice, ctx->stream);
            if (e == hipSuccess) e = launch_repack(tmp, kv.table + f0 * fsw, nf, (uint32_t)fw, (uint32_t)fsw, 1, ctx->stream);
        } else {
            e = launch_repack(kv.table + f0 * fsw, tmp, nf, (uint32_t)fw, (uint32_t)fsw, 0, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(h, tmp, nf * fw * 4, hipMemcpyDeviceToHost, ctx->stream);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    dev_free(tmp);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("table copy: ") + hipGetErrorString(e));
    return LTPG_OK;
}

}  // namespace
}  // namespace ltpg

using namespace ltpg;

extern "C" {

void ltpg_config_init(ltpg_config *cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->size = sizeof(ltpg_config);
    cfg->device = 0;
    cfg->workload = LTPG_WORKLOAD_KV;
    cfg->rules = LTPG_RULES_REFERENCE;
    cfg->values = LTPG_VALUES_RMW;
    cfg->max_txns = 65536;
    cfg->kv_rows = 1000000;
    cfg->kv_fields = 10;
    cfg->kv_field_bytes = 4;
    cfg->kv_max_slots = 16;
    cfg->kv_hot_row = LTPG_NO_HOT_ROW;
    cfg->tpcc_warehouses = 64;
    cfg->hbm_budget = 0;
}

int ltpg_abi_version(void) { return LTPG_ABI_VERSION; }

const char *ltpg_last_error(void) { return g_last_error.c_str(); }

int ltpg_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int ltpg_create(const ltpg_config *cfg, ltpg_ctx **out)
{
    if (!cfg || !out) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->size != sizeof(ltpg_config)) LTPG_FAIL(LTPG_ERR_INVALID, "ltpg_config.size mismatch (call ltpg_config_init)");
    if (cfg->workload != LTPG_WORKLOAD_KV && cfg->workload != LTPG_WORKLOAD_TPCC)
        LTPG_FAIL(LTPG_ERR_INVALID, "unknown workload");
    if (cfg->rules > LTPG_RULES_INTENDED) LTPG_FAIL(LTPG_ERR_INVALID, "unknown rule set");
    if (cfg->values > LTPG_VALUES_RMW) LTPG_FAIL(LTPG_ERR_INVALID, "unknown value mode");
    if (cfg->max_txns == 0 || cfg->max_txns > (1u << 24)) LTPG_FAIL(LTPG_ERR_INVALID, "max_txns must be in [1, 2^24]");
    if (cfg->flags & ~LTPG_FLAG_CONFLICT_COUNTS) LTPG_FAIL(LTPG_ERR_INVALID, "unknown flag");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) LTPG_FAIL(LTPG_ERR_DEVICE, "no HIP device visible");
    if (cfg->device < 0 || cfg->device >= ndev) LTPG_FAIL(LTPG_ERR_INVALID, "device ordinal out of range");

    ltpg_ctx *ctx = new (std::nothrow) ltpg_ctx();
    if (!ctx) LTPG_FAIL(LTPG_ERR_NOMEM, "host allocation of context failed");
    ctx->cfg = *cfg;
    ctx->device = cfg->device;
    int rc = LTPG_OK;
    do {
        hipError_t e = hipSetDevice(ctx->device);
        if (e != hipSuccess) {
            set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
            rc = LTPG_ERR_DEVICE;
            break;
        }
        e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
            rc = LTPG_ERR_DEVICE;
            break;
        }
        ctx->stream = ctx->own_stream;
        if ((rc = dev_alloc(&ctx->counters, 1, "counters"))) break;
        e = hipMemsetAsync(ctx->counters, 0, sizeof(DeviceCounters), ctx->stream);
        if (e != hipSuccess) {
            set_error(std::string("hipMemsetAsync: ") + hipGetErrorString(e));
            rc = LTPG_ERR_DEVICE;
            break;
        }
        if (cfg->workload == LTPG_WORKLOAD_KV) {
            if ((rc = init_kv(ctx))) break;
        } else {
            if ((rc = init_tpcc(ctx))) break;
        }
        e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            set_error(std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
            rc = LTPG_ERR_DEVICE;
            break;
        }
    } while (0);
    if (rc) {
        std::string msg = g_last_error;
        ltpg_destroy(ctx);
        set_error(msg);
        return rc;
    }
    *out = ctx;
    return LTPG_OK;
}

int ltpg_destroy(ltpg_ctx *ctx)
{
    if (!ctx) return LTPG_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->copy_stream) (void)hipStreamSynchronize(ctx->copy_stream);  // copies into the batch slots
    free_kv(ctx->kv);
    free_tpcc(ctx->tpcc);
    dev_free(ctx->counters);
    for (auto &ev : ctx->ev_pool)
        if (ev) (void)hipEventDestroy(ev);
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    for (int k = 0; k < 2; k++) {
        if (ctx->xfer_copied[k]) (void)hipEventDestroy(ctx->xfer_copied[k]);
        if (ctx->xfer_released[k]) (void)hipEventDestroy(ctx->xfer_released[k]);
    }
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return LTPG_OK;
}

int ltpg_set_stream(ltpg_ctx *ctx, void *hip_stream)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own_stream;
    return LTPG_OK;
}

int ltpg_sync(ltpg_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t errors = 0;
    LTPG_HIP_TRY(hipMemcpy(&errors, &ctx->counters->range_errors, sizeof(errors), hipMemcpyDeviceToHost));
    if (errors) {
        LTPG_HIP_TRY(hipMemset(&ctx->counters->range_errors, 0, sizeof(uint32_t)));
        LTPG_FAIL(LTPG_ERR_RANGE, std::to_string(errors) + " op(s) named a row or field outside the table");
    }
    return LTPG_OK;
}

int ltpg_set_timing(ltpg_ctx *ctx, int enable)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    ctx->timing = enable != 0;
    return LTPG_OK;
}

int ltpg_phase_times(ltpg_ctx *ctx, float ms[LTPG_TIMING_SLOTS], uint32_t *n_epochs)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ms) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < LTPG_TIMING_SLOTS; k++) ms[k] = 0.f;
    const uint32_t n = ctx->ev_used / kEv;
    for (uint32_t e = 0; e < n; e++) {
        for (int k = 0; k < LTPG_TIMING_SLOTS; k++) {
            float v = 0.f;
            LTPG_HIP_TRY(hipEventElapsedTime(&v, ctx->ev_pool[kEv * e + k], ctx->ev_pool[kEv * e + k + 1]));
            ms[k] += v;
        }
    }
    if (n_epochs) *n_epochs = n;
    ctx->ev_used = 0;
    return LTPG_OK;
}

int ltpg_commit_totals(ltpg_ctx *ctx, uint32_t type, uint64_t *commits, uint64_t *attempts, int reset)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (type > LTPG_TYPE_DELIVERY) LTPG_FAIL(LTPG_ERR_INVALID, "unknown transaction type");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    static_assert(sizeof(DeviceCounters::totals) == kCounterStripes * 32 * 8, "stripe layout");
    unsigned long long h[kCounterStripes][32];
    LTPG_HIP_TRY(hipMemcpy(h, ctx->counters->totals, sizeof(h), hipMemcpyDeviceToHost));
    uint64_t c = 0, a = 0;
    for (int k = 0; k < kCounterStripes; k++) {
        c += h[k][type];
        a += h[k][kAttempts + type];
    }
    if (commits) *commits = c;
    if (attempts) *attempts = a;
    if (reset) LTPG_HIP_TRY(hipMemset(ctx->counters->totals, 0, sizeof(h)));
    return LTPG_OK;
}

// ---- device memory ---------------------------------------------------------

int ltpg_device_alloc(ltpg_ctx *ctx, uint64_t bytes, void **dptr)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dptr) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    *dptr = nullptr;
    if (bytes == 0) LTPG_FAIL(LTPG_ERR_INVALID, "zero-byte allocation");
    uint8_t *p = nullptr;
    if ((rc = dev_alloc(&p, bytes, "caller buffer"))) return rc;
    *dptr = p;
    return LTPG_OK;
}

int ltpg_device_free(ltpg_ctx *ctx, void *dptr)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (dptr) {
        LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
        LTPG_HIP_TRY(hipFree(dptr));
    }
    return LTPG_OK;
}

int ltpg_copy_to_device(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dst || !src) LTPG_FAIL(LTPG_ERR_INVALID, "null pointer");
    if (bytes) {
        LTPG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return LTPG_OK;
}

int ltpg_copy_to_host(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dst || !src) LTPG_FAIL(LTPG_ERR_INVALID, "null pointer");
    if (bytes) {
        LTPG_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return LTPG_OK;
}

int ltpg_random_access_rates(ltpg_ctx *ctx, uint64_t bytes, uint64_t n_ops, double *load_gops,
                             double *atomic_gops)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    const uint64_t len = bytes / 8;
    if (len == 0 || n_ops == 0) LTPG_FAIL(LTPG_ERR_INVALID, "bytes and n_ops must be non-zero");
    if (n_ops > (1ull << 36)) LTPG_FAIL(LTPG_ERR_INVALID, "n_ops too large");
    unsigned long long *a = nullptr, *sink = nullptr;
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    float ms[2] = {0.f, 0.f};
    if ((rc = dev_alloc(&a, len, "calibration array"))) return rc;
    if ((rc = dev_alloc(&sink, 1, "calibration sink"))) {
        dev_free(a);
        return rc;
    }
    hipError_t h = launch_fill_u64(a, len, ~0ull, ctx->stream);
    for (int k = 0; k < 3 && h == hipSuccess; k++) h = hipEventCreate(&e[k]);
    // warm-up pass of each, then the timed pair (the atomic pass needs fresh
    // all-ones contents so that every probe goes on to its atomic)
    if (h == hipSuccess) h = launch_calibration(a, len, n_ops, 0, sink, ctx->stream);
    if (h == hipSuccess) h = hipEventRecord(e[0], ctx->stream);
    if (h == hipSuccess) h = launch_calibration(a, len, n_ops, 0, sink, ctx->stream);
    if (h == hipSuccess) h = hipEventRecord(e[1], ctx->stream);
    if (h == hipSuccess) h = launch_fill_u64(a, len, ~0ull, ctx->stream);
    if (h == hipSuccess) h = hipEventRecord(e[2], ctx->stream);
    if (h == hipSuccess) h = launch_calibration(a, len, n_ops, 1, sink, ctx->stream);
    hipEvent_t e3 = nullptr;
    if (h == hipSuccess) h = hipEventCreate(&e3);
    if (h == hipSuccess) h = hipEventRecord(e3, ctx->stream);
    if (h == hipSuccess) h = hipEventSynchronize(e3);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms[0], e[0], e[1]);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms[1], e[2], e3);
    for (int k = 0; k < 3; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    if (e3) (void)hipEventDestroy(e3);
    dev_free(a);
    dev_free(sink);
    if (h != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("calibration: ") + hipGetErrorString(h));
    if (load_gops) *load_gops = ms[0] > 0 ? n_ops / (ms[0] * 1e6) : 0.0;
    if (atomic_gops) *atomic_gops = ms[1] > 0 ? n_ops / (ms[1] * 1e6) : 0.0;
    return LTPG_OK;
}

int ltpg_random_gather_rate(ltpg_ctx *ctx, uint64_t bytes, uint32_t field_bytes, uint32_t stride_bytes,
                            uint64_t n_ops, double *gbs)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!gbs) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    if (field_bytes == 0 || field_bytes % 4 || field_bytes > 128 || stride_bytes < field_bytes || stride_bytes % 4)
        LTPG_FAIL(LTPG_ERR_INVALID, "field_bytes a multiple of 4 in [4, 128], stride_bytes >= field_bytes, a multiple of 4");
    const uint64_t slots = bytes / stride_bytes;
    if (slots == 0 || n_ops == 0 || n_ops > (1ull << 32)) LTPG_FAIL(LTPG_ERR_INVALID, "bad bytes or n_ops");
    uint32_t *src = nullptr, *dst = nullptr;
    if ((rc = dev_alloc(&src, slots * (stride_bytes / 4), "gather source"))) return rc;
    if ((rc = dev_alloc(&dst, n_ops * (field_bytes / 4), "gather destination"))) {
        dev_free(src);
        return rc;
    }
    hipEvent_t e[2] = {nullptr, nullptr};
    float ms = 0.f;
    hipError_t h = hipMemsetAsync(src, 0, slots * stride_bytes, ctx->stream);
    for (int k = 0; k < 2 && h == hipSuccess; k++) h = hipEventCreate(&e[k]);
    if (h == hipSuccess)
        h = launch_gather_calibration(src, slots, stride_bytes / 4, field_bytes / 4, dst, n_ops, ctx->stream);
    if (h == hipSuccess) h = hipEventRecord(e[0], ctx->stream);
    if (h == hipSuccess)
        h = launch_gather_calibration(src, slots, stride_bytes / 4, field_bytes / 4, dst, n_ops, ctx->stream);
    if (h == hipSuccess) h = hipEventRecord(e[1], ctx->stream);
    if (h == hipSuccess) h = hipEventSynchronize(e[1]);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms, e[0], e[1]);
    for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    dev_free(src);
    dev_free(dst);
    if (h != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("gather calibration: ") + hipGetErrorString(h));
    *gbs = ms > 0 ? 2.0 * (double)n_ops * field_bytes / (ms * 1e6) : 0.0;
    return LTPG_OK;
}

int ltpg_sync_costs(ltpg_ctx *ctx, uint32_t blocks, uint32_t reps, double out[3])
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out || blocks == 0 || reps == 0 || reps > 100000) LTPG_FAIL(LTPG_ERR_INVALID, "bad blocks, reps or output");
    unsigned int *w = nullptr;  // [0] count, [1] generation, [2] error, then one word per block
    if ((rc = dev_alloc(&w, 3 + (uint64_t)blocks, "sync cost words"))) return rc;
    hipStream_t s = ctx->stream;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e[6] = {};
    float ms[3] = {0.f, 0.f, 0.f};
    unsigned int err = 0;
    hipError_t h = hipMemsetAsync(w, 0, 12, s);
    for (int k = 0; k < 6 && h == hipSuccess; k++) h = hipEventCreate(&e[k]);
    // grid barriers (a warm-up pass first)
    constexpr uint32_t kSpinLimit = 1u << 20;
    if (h == hipSuccess) h = launch_grid_barrier(blocks, w, w + 1, 8, kSpinLimit, w + 2, s);
    if (h == hipSuccess) h = hipEventRecord(e[0], s);
    if (h == hipSuccess) h = launch_grid_barrier(blocks, w, w + 1, reps, kSpinLimit, w + 2, s);
    if (h == hipSuccess) h = hipEventRecord(e[1], s);
    // kernel boundaries on the stream
    if (h == hipSuccess) h = launch_boundary(blocks, w + 3, s);
    if (h == hipSuccess) h = hipEventRecord(e[2], s);
    for (uint32_t r = 0; r < reps && h == hipSuccess; r++) h = launch_boundary(blocks, w + 3, s);
    if (h == hipSuccess) h = hipEventRecord(e[3], s);
    // kernel boundaries inside a graph
    if (h == hipSuccess) h = hipStreamSynchronize(s);
    if (h == hipSuccess) h = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (h == hipSuccess) {
        hipError_t hc = hipSuccess;
        for (uint32_t r = 0; r < reps && hc == hipSuccess; r++) hc = launch_boundary(blocks, w + 3, s);
        h = hipStreamEndCapture(s, &graph);
        if (h == hipSuccess) h = hc;
    }
    if (h == hipSuccess) h = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (h == hipSuccess) h = hipGraphLaunch(exec, s);
    if (h == hipSuccess) h = hipEventRecord(e[4], s);
    if (h == hipSuccess) h = hipGraphLaunch(exec, s);
    if (h == hipSuccess) h = hipEventRecord(e[5], s);
    if (h == hipSuccess) h = hipEventSynchronize(e[5]);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms[0], e[0], e[1]);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms[1], e[2], e[3]);
    if (h == hipSuccess) h = hipEventElapsedTime(&ms[2], e[4], e[5]);
    if (h == hipSuccess) h = hipMemcpy(&err, w + 2, 4, hipMemcpyDeviceToHost);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (int k = 0; k < 6; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    dev_free(w);
    if (h != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("sync costs: ") + hipGetErrorString(h));
    if (err) LTPG_FAIL(LTPG_ERR_DEVICE, "grid barrier did not complete (poll limit reached)");
    out[0] = ms[0] * 1e3 / reps;
    out[1] = ms[2] * 1e3 / reps;
    out[2] = ms[1] * 1e3 / reps;
    return LTPG_OK;
}

int ltpg_latency_costs(ltpg_ctx *ctx, double out_ns[5])
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out_ns) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    constexpr uint32_t kChain = 4096, kReps = 4096, kHostReps = 512;
    uint64_t *host = nullptr;
    unsigned long long *d_out = nullptr;
    if (hipHostMalloc(reinterpret_cast<void **>(&host), (kChain + 1) * sizeof(uint64_t),
                      hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess)
        LTPG_FAIL(LTPG_ERR_NOMEM, "pinned host allocation failed");
    for (uint32_t i = 0; i < kChain; i++) host[i] = (i * 2053u + 1u) & (kChain - 1);
    host[kChain] = 0;
    if ((rc = dev_alloc(&d_out, 8, "latency probe output"))) {
        (void)hipHostFree(host);
        return rc;
    }
    unsigned long long ticks[8] = {};
    int khz = 0;
    hipError_t h = launch_latency_probe(host, kChain, d_out, kReps, kHostReps, ctx->stream);
    if (h == hipSuccess) h = hipStreamSynchronize(ctx->stream);
    if (h == hipSuccess) h = hipMemcpy(ticks, d_out, sizeof(ticks), hipMemcpyDeviceToHost);
    if (h == hipSuccess) h = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device);
    dev_free(d_out);
    (void)hipHostFree(host);
    if (h != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("latency probe: ") + hipGetErrorString(h));
    if (khz <= 0) LTPG_FAIL(LTPG_ERR_DEVICE, "no wall-clock rate");
    const double ns_per_tick = 1e6 / khz;
    const uint32_t reps[5] = {kReps, kReps, kReps, kHostReps, kHostReps};
    for (int k = 0; k < 5; k++) out_ns[k] = ticks[k] * ns_per_tick / reps[k];
    return LTPG_OK;
}

int ltpg_device_copy(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!dst || !src) LTPG_FAIL(LTPG_ERR_INVALID, "null pointer");
    if (bytes % 16 || reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(src) % 16)
        LTPG_FAIL(LTPG_ERR_INVALID, "device copy needs 16-byte multiples and alignment");
    if (bytes) LTPG_HIP_TRY(launch_copy_u128(dst, src, bytes / 16, ctx->stream));
    return LTPG_OK;
}

// ---- key-value tables ---------------------------------------------------

int ltpg_kv_table_load(ltpg_ctx *ctx, const void *host, uint64_t rows, uint32_t layout)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    if (!host) LTPG_FAIL(LTPG_ERR_INVALID, "null host table");
    if (rows != kv.rows) LTPG_FAIL(LTPG_ERR_INVALID, "rows does not match the configured kv_rows");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (layout == LTPG_LAYOUT_ROW_MAJOR) return kv_copy_table(ctx, const_cast<void *>(host), true);
    if (layout != LTPG_LAYOUT_COLUMN_MAJOR) LTPG_FAIL(LTPG_ERR_INVALID, "unknown layout");
    if (kv.field_words != 1) LTPG_FAIL(LTPG_ERR_INVALID, "column-major load needs kv_field_bytes == 4");
    const uint64_t words = kv.rows * kv.fields;
    uint32_t *tmp = nullptr;
    if ((rc = dev_alloc(&tmp, words, "transpose staging"))) return rc;
    hipError_t e = hipMemcpyAsync(tmp, host, words * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_kv_transpose(tmp, kv.table, kv.rows, kv.fields, 1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(tmp);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("column-major load: ") + hipGetErrorString(e));
    return LTPG_OK;
}

int ltpg_kv_table_read(ltpg_ctx *ctx, void *host, uint64_t rows, uint32_t layout)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    if (!host) LTPG_FAIL(LTPG_ERR_INVALID, "null host table");
    if (rows != kv.rows) LTPG_FAIL(LTPG_ERR_INVALID, "rows does not match the configured kv_rows");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (layout == LTPG_LAYOUT_ROW_MAJOR) return kv_copy_table(ctx, host, false);
    if (layout != LTPG_LAYOUT_COLUMN_MAJOR) LTPG_FAIL(LTPG_ERR_INVALID, "unknown layout");
    if (kv.field_words != 1) LTPG_FAIL(LTPG_ERR_INVALID, "column-major read needs kv_field_bytes == 4");
    const uint64_t words = kv.rows * kv.fields;
    uint32_t *tmp = nullptr;
    if ((rc = dev_alloc(&tmp, words, "transpose staging"))) return rc;
    hipError_t e = launch_kv_transpose(kv.table, tmp, kv.rows, kv.fields, 0, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess) e = hipMemcpy(host, tmp, words * 4, hipMemcpyDeviceToHost);
    dev_free(tmp);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("column-major read: ") + hipGetErrorString(e));
    return LTPG_OK;
}

int ltpg_kv_table_read_rows(ltpg_ctx *ctx, const uint64_t *rows, uint64_t n, uint32_t *host)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    if (n && (!rows || !host)) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    for (uint64_t i = 0; i < n; i++)
        if (rows[i] >= kv.rows) LTPG_FAIL(LTPG_ERR_RANGE, "row index outside the table");
    if (n == 0) return LTPG_OK;
    const uint64_t per = (uint64_t)kv.fields * kv.field_words;
    uint64_t *d_idx = nullptr;
    uint32_t *d_out = nullptr;
    if ((rc = dev_alloc(&d_idx, n, "row indices")) || (rc = dev_alloc(&d_out, n * per, "gathered rows"))) {
        dev_free(d_idx);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_idx, rows, n * 8, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_kv_gather_rows(kv, d_idx, n, d_out, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host, d_out, n * per * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(d_idx);
    dev_free(d_out);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("row read: ") + hipGetErrorString(e));
    return LTPG_OK;
}

int ltpg_kv_table_fill_reference(ltpg_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    LTPG_HIP_TRY(launch_kv_fill_reference(kv, ctx->stream));
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return LTPG_OK;
}

int ltpg_kv_set_tuning(ltpg_ctx *ctx, uint32_t flags)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    if (flags & ~(LTPG_TUNE_ALL | LTPG_TUNE_SMALL_EPOCH | LTPG_TUNE_PARTITIONED | LTPG_TUNE_MORE_ITEMS))
        LTPG_FAIL(LTPG_ERR_INVALID, "unknown tuning flag");
    if ((flags & LTPG_TUNE_PARTITIONED) && (rc = kv_part_prepare(ctx))) return rc;
    ctx->kv.partitioned = (flags & LTPG_TUNE_PARTITIONED) ? 1 : 0;
    ctx->kv.exec_per_op = (flags & LTPG_TUNE_EXEC_PER_OP) ? 1 : 0;
    ctx->kv.split_regs = (flags & LTPG_TUNE_SPLIT_REGISTERS) ? 1 : 0;
    ctx->kv.small_chunks = (flags & LTPG_TUNE_SMALL_CHUNKS) ? 1 : 0;
    ctx->kv.lane_copy = (flags & LTPG_TUNE_LANE_COPY) ? 1 : 0;
    ctx->kv.small_epoch = (flags & LTPG_TUNE_SMALL_EPOCH) ? 1 : 0;
    ctx->kv.more_items = (flags & LTPG_TUNE_MORE_ITEMS) ? 1 : 0;
    return LTPG_OK;
}

int ltpg_tpcc_set_tuning(ltpg_ctx *ctx, uint32_t flags)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (ctx->cfg.workload != LTPG_WORKLOAD_TPCC) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no order-entry tables");
    if (flags & ~LTPG_TUNE_TPCC_SMALL_BLOCKS) LTPG_FAIL(LTPG_ERR_INVALID, "unknown tuning flag");
    ctx->tpcc.small_blocks = (flags & LTPG_TUNE_TPCC_SMALL_BLOCKS) != 0;
    return LTPG_OK;
}

// ---- key-value batches --------------------------------------------------

}  // extern "C"

namespace {

// Transport of host-resident batches.  The reference keeps two device query
// slots and copies the batch of epoch e+1 into one while epoch e runs on the
// other (TPCC-allopt-final/Query.h:532-605, Predefine.h:18-19).  Here a host
// submit takes the slot the epoch before last used, copies into it on the
// context's copy stream once that epoch is done with it, and makes its own
// epoch's kernels (on the context stream) wait for the copy: the copy of a
// batch overlaps the epoch still running on the other slot.  The submit
// returns once its copy is done, so the caller may reuse its buffers at once
// (pageable memory is staged by the runtime; pinned memory goes at full
// PCIe speed), without waiting for the device to go idle.
int xfer_begin(ltpg_ctx *ctx, int *slot)
{
    if (!ctx->copy_stream) {
        LTPG_HIP_TRY(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
        for (int k = 0; k < 2; k++) {
            LTPG_HIP_TRY(hipEventCreateWithFlags(&ctx->xfer_copied[k], hipEventDisableTiming));
            LTPG_HIP_TRY(hipEventCreateWithFlags(&ctx->xfer_released[k], hipEventDisableTiming));
        }
    }
    const int k = ctx->xfer_slot ^ 1;
    // everything queued so far -- the epoch on the other slot and all before
    // it -- is what must finish before the other slot is free again
    LTPG_HIP_TRY(hipEventRecord(ctx->xfer_released[k ^ 1], ctx->stream));
    ctx->xfer_released_set[k ^ 1] = true;
    if (ctx->xfer_released_set[k]) LTPG_HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->xfer_released[k], 0));
    ctx->xfer_slot = k;
    *slot = k;
    return LTPG_OK;
}

// The slot's copies are queued on the copy stream: the epoch waits for them,
// and the caller's buffers are free once they are done.
int xfer_end(ltpg_ctx *ctx)
{
    const int k = ctx->xfer_slot;
    LTPG_HIP_TRY(hipEventRecord(ctx->xfer_copied[k], ctx->copy_stream));
    LTPG_HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->xfer_copied[k], 0));
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    return LTPG_OK;
}

template <typename T>
int slot_buffer(T **buf, uint64_t count, const char *what)
{
    return *buf ? LTPG_OK : dev_alloc(buf, count, what);
}

}  // namespace

extern "C" {

int ltpg_kv_submit(ltpg_ctx *ctx, const ltpg_kv_batch *b, uint32_t mem)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    if (!b) LTPG_FAIL(LTPG_ERR_INVALID, "null batch");
    if (b->n_txns > ctx->cfg.max_txns) LTPG_FAIL(LTPG_ERR_INVALID, "n_txns exceeds max_txns");
    if (b->n_slots > kv.max_slots) LTPG_FAIL(LTPG_ERR_INVALID, "n_slots exceeds kv_max_slots");
    if (b->n_txns > 0 && b->n_slots == 0) LTPG_FAIL(LTPG_ERR_INVALID, "n_slots must be > 0 for a non-empty batch");
    if (b->n_txns > 0 && (!b->keys || !b->meta || !b->tids)) LTPG_FAIL(LTPG_ERR_INVALID, "null batch array");
    const bool request_major = (mem & LTPG_BATCH_REQUEST_MAJOR) != 0;
    mem &= ~LTPG_BATCH_REQUEST_MAJOR;
    if (mem != LTPG_MEM_HOST && mem != LTPG_MEM_DEVICE) LTPG_FAIL(LTPG_ERR_INVALID, "unknown memory kind");
    if ((rc = next_epoch(ctx))) return rc;
    KvBatchDev d;
    d.n_txns = b->n_txns;
    d.n_slots = b->n_slots;
    d.request_major = request_major;
    if (mem == LTPG_MEM_DEVICE) {
        d.keys = b->keys;
        d.meta = b->meta;
        d.tids = b->tids;
    } else {
        int k = 0;
        const uint64_t cap = (uint64_t)ctx->cfg.max_txns, n = (uint64_t)b->n_txns * b->n_slots;
        if ((rc = xfer_begin(ctx, &k))) return rc;
        if ((rc = slot_buffer(&kv.keys_buf[k], kv.max_slots * cap, "kv keys")) ||
            (rc = slot_buffer(&kv.meta_buf[k], kv.max_slots * cap, "kv meta")) ||
            (rc = slot_buffer(&kv.tids_buf[k], cap, "kv tids")))
            return rc;
        hipStream_t cs = ctx->copy_stream;
        if (n) {
            LTPG_HIP_TRY(hipMemcpyAsync(kv.keys_buf[k], b->keys, n * 8, hipMemcpyHostToDevice, cs));
            LTPG_HIP_TRY(hipMemcpyAsync(kv.meta_buf[k], b->meta, n * 4, hipMemcpyHostToDevice, cs));
        }
        if (b->n_txns)
            LTPG_HIP_TRY(hipMemcpyAsync(kv.tids_buf[k], b->tids, (uint64_t)b->n_txns * 4, hipMemcpyHostToDevice, cs));
        if ((rc = xfer_end(ctx))) return rc;
        d.keys = kv.keys_buf[k];
        d.meta = kv.meta_buf[k];
        d.tids = kv.tids_buf[k];
    }
    kv.cur = d;
    kv.has_batch = true;
    return LTPG_OK;
}

// ---- epoch phases --------------------------------------------------------

namespace {
// Take the next LTPG_TIMING_SLOTS + 1 timing events of this epoch.
static int take_events(ltpg_ctx *ctx, hipEvent_t **out)
{
    if (ctx->ev_used + kEv > ctx->ev_pool.size()) {
        for (int k = 0; k < 8 * kEv; k++) {
            hipEvent_t e = nullptr;
            LTPG_HIP_TRY(hipEventCreate(&e));
            ctx->ev_pool.push_back(e);
        }
    }
    *out = &ctx->ev_pool[ctx->ev_used];
    ctx->ev_used += kEv;
    return LTPG_OK;
}

// CID, EXECUTE, CHECK (with the committed reads), WRITEBACK of the current
// order-entry batch; the read-out timing slot stays empty.
static int tpcc_epoch_run(ltpg_ctx *ctx, uint32_t phase_mask)
{
    TpccState &tp = ctx->tpcc;
    if (!tp.has_batch) LTPG_FAIL(LTPG_ERR_STATE, "no batch submitted");
    const EpochTag e = epoch_tag(ctx);
    const bool counts = (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) != 0;
    hipStream_t s = ctx->stream;
    hipEvent_t *ev = nullptr;
    int rc;
    if (ctx->timing) {
        if ((rc = take_events(ctx, &ev))) return rc;
        LTPG_HIP_TRY(hipEventRecord(ev[0], s));
    }
    if (phase_mask & LTPG_PHASE_CID) LTPG_HIP_TRY(launch_tpcc_cid(tp.d, tp.pay, tp.small_blocks, s));
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[1], s));
    const bool ro = tp.ro.n_os + tp.ro.n_sl > 0, dl = tp.dl.n > 0;
    const uint32_t dl_base = tp.no.n + tp.pay.n + tp.ro.n_os + tp.ro.n_sl;
    if (phase_mask & LTPG_PHASE_EXECUTE) {
        LTPG_HIP_TRY(launch_tpcc_execute(tp.d, tp.no, tp.pay, e, counts, tp.small_blocks, ctx->counters, s));
        // Deliveries register after NewOrders and Payments, then read-only
        // transactions mark war on the registered writers
        if (dl) LTPG_HIP_TRY(launch_tpcc_delivery_execute(tp.d, tp.dl, dl_base, e, ctx->counters, s));
        if (ro) LTPG_HIP_TRY(launch_tpcc_readonly(tp.d, tp.ro, tp.no.n + tp.pay.n, e, ctx->counters, s));
    }
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[2], s));
    if (phase_mask & LTPG_PHASE_CHECK) {
        if (!(phase_mask & LTPG_PHASE_EXECUTE)) LTPG_HIP_TRY(launch_reset_epoch_counters(ctx->counters, s));
        LTPG_HIP_TRY(launch_tpcc_check(tp.d, tp.no, tp.pay, e, counts, ro, dl, tp.small_blocks, ctx->counters, s));
        if (dl) LTPG_HIP_TRY(launch_tpcc_delivery_check(tp.d, tp.dl, dl_base, e, ctx->counters, s));
    }
    if (ev) {
        LTPG_HIP_TRY(hipEventRecord(ev[3], s));
        LTPG_HIP_TRY(hipEventRecord(ev[4], s));
    }
    if (phase_mask & LTPG_PHASE_WRITEBACK) {
        LTPG_HIP_TRY(launch_tpcc_writeback(tp.d, tp.no, tp.pay, e, ctx->cfg.values, tp.small_blocks, s));
        if (dl) LTPG_HIP_TRY(launch_tpcc_delivery_writeback(tp.d, tp.dl, dl_base, e, ctx->cfg.values, s));
    }
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[5], s));
    return LTPG_OK;
}
}  // namespace

int ltpg_epoch_run(ltpg_ctx *ctx, uint32_t phase_mask)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (phase_mask & ~LTPG_PHASE_ALL) LTPG_FAIL(LTPG_ERR_INVALID, "unknown phase bit");
    if (ctx->cfg.workload == LTPG_WORKLOAD_TPCC) return tpcc_epoch_run(ctx, phase_mask);
    KvState &kv = ctx->kv;
    if (!kv.has_batch) LTPG_FAIL(LTPG_ERR_STATE, "no batch submitted");
    const EpochTag e = epoch_tag(ctx);
    hipStream_t s = ctx->stream;
    if (kv.small_epoch && phase_mask == LTPG_PHASE_ALL && !ctx->timing &&
        !(ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) && kv_small_epoch_ok(kv, kv.cur)) {
        LTPG_HIP_TRY(launch_kv_small_epoch(kv, kv.cur, ctx->cfg.values, ctx->counters, nullptr, nullptr, 0, s));
        return LTPG_OK;
    }
    kv.regs_clean = false;  // the phase kernels below write registers and flags
    hipEvent_t *ev = nullptr;
    if (ctx->timing) {
        if ((rc = take_events(ctx, &ev))) return rc;
        LTPG_HIP_TRY(hipEventRecord(ev[0], s));
    }
    // CID has no key-value work: its bracket is empty.
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[1], s));
    if (phase_mask & LTPG_PHASE_EXECUTE)
        LTPG_HIP_TRY(launch_kv_execute(kv, kv.cur, e, (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) != 0,
                                       ctx->counters, s));
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[2], s));
    if (phase_mask & LTPG_PHASE_CHECK) {
        if (!(phase_mask & LTPG_PHASE_EXECUTE)) LTPG_HIP_TRY(launch_reset_epoch_counters(ctx->counters, s));
        LTPG_HIP_TRY(launch_kv_check(kv, kv.cur, e, (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) != 0,
                                     ctx->counters, s));
    }
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[3], s));
    // Committed reads see the snapshot, so they are delivered before any
    // write.  The fast CHECK already gathered them; the counting CHECK
    // leaves them to a read-out pass here.
    if ((phase_mask & LTPG_PHASE_WRITEBACK) && (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS))
        LTPG_HIP_TRY(launch_kv_readout(kv, kv.cur, s));
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[4], s));
    if (phase_mask & LTPG_PHASE_WRITEBACK) LTPG_HIP_TRY(launch_kv_writeback(kv, kv.cur, ctx->cfg.values, s));
    if (ev) LTPG_HIP_TRY(hipEventRecord(ev[5], s));
    return LTPG_OK;
}

}  // extern "C"

namespace ltpg {

// The reference's launch pattern for one epoch (the driver loops of
// TPF/tp.cu:136-261 and Y0/ycsb.cu:78-152): one phase kernel at a time, each
// followed by a stream synchronisation, then the error check.  The plain
// per-batch form that the batched and served forms are compared with.
int run_epoch_stepwise(ltpg_ctx *ctx)
{
    static const uint32_t kPhases[4] = {LTPG_PHASE_CID, LTPG_PHASE_EXECUTE, LTPG_PHASE_CHECK, LTPG_PHASE_WRITEBACK};
    for (uint32_t p : kPhases) {
        if (p == LTPG_PHASE_CID && ctx->cfg.workload != LTPG_WORKLOAD_TPCC) continue;  // no key-value CID
        int rc = ltpg_epoch_run(ctx, p);
        if (rc) return rc;
        LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    return ltpg_sync(ctx);
}

}  // namespace ltpg

extern "C" {

int ltpg_epoch_reset(ltpg_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    return next_epoch(ctx);
}

// ---- results --------------------------------------------------------------

int ltpg_commit_bitmap(ltpg_ctx *ctx, uint32_t type, uint64_t *out, uint64_t n_words)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    const uint64_t *src = nullptr;
    uint64_t n = 0;
    if (type == LTPG_TYPE_KV && ctx->kv.table) {
        src = ctx->kv.bitmap;
        n = ctx->kv.cur.n_txns;
    } else if (type == LTPG_TYPE_NEWORDER && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.bitmap_no;
        n = ctx->tpcc.no.n;
    } else if (type == LTPG_TYPE_PAYMENT && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.bitmap_pay;
        n = ctx->tpcc.pay.n;
    } else if (type == LTPG_TYPE_ORDERSTATUS && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.bitmap_os;
        n = ctx->tpcc.ro.n_os;
    } else if (type == LTPG_TYPE_STOCKLEVEL && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.bitmap_sl;
        n = ctx->tpcc.ro.n_sl;
    } else if (type == LTPG_TYPE_DELIVERY && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.bitmap_dl;
        n = ctx->tpcc.dl.n;
    } else {
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "transaction type not in this context's workload");
    }
    const uint64_t need = (n + 63) / 64;
    if (n_words < need) LTPG_FAIL(LTPG_ERR_INVALID, "bitmap buffer too small");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (need) LTPG_HIP_TRY(hipMemcpy(out, src, need * 8, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

int ltpg_commit_count(ltpg_ctx *ctx, uint32_t type, uint32_t *out)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    if (type > LTPG_TYPE_DELIVERY) LTPG_FAIL(LTPG_ERR_INVALID, "unknown transaction type");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t h[kCounterStripes][64];
    LTPG_HIP_TRY(hipMemcpy(h, ctx->counters->commits, sizeof(h), hipMemcpyDeviceToHost));
    uint32_t sum = 0;
    for (int k = 0; k < kCounterStripes; k++) sum += h[k][type];
    *out = sum;
    return LTPG_OK;
}

int ltpg_conflict_counters(ltpg_ctx *ctx, uint32_t type, uint32_t *out, uint32_t n)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    const uint32_t *src = nullptr;
    uint32_t have = 0;
    if (type == LTPG_TYPE_KV && ctx->kv.table) {
        src = ctx->kv.conflicts;
        have = ctx->kv.cur.n_txns;
    } else if (type == LTPG_TYPE_NEWORDER && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.conf_no;
        have = ctx->tpcc.no.n;
    } else if (type == LTPG_TYPE_PAYMENT && ctx->tpcc.d.regs) {
        src = ctx->tpcc.d.conf_pay;
        have = ctx->tpcc.pay.n;
    } else {
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "transaction type not in this context's workload");
    }
    if (!(ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS))
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "conflict counters need LTPG_FLAG_CONFLICT_COUNTS");
    if (n > have) LTPG_FAIL(LTPG_ERR_INVALID, "n exceeds the batch size");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) LTPG_HIP_TRY(hipMemcpy(out, src, (uint64_t)n * 4, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

int ltpg_kv_read_values(ltpg_ctx *ctx, uint32_t *out, uint64_t n_words)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    KvState &kv = ctx->kv;
    if (!kv.table) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no key-value table");
    const uint64_t have = (uint64_t)kv.cur.n_slots * kv.cur.n_txns * kv.read_slot * kv.field_words;
    if (n_words > have) LTPG_FAIL(LTPG_ERR_INVALID, "n_words exceeds the staged read values");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_words) LTPG_HIP_TRY(hipMemcpy(out, kv.read_values, n_words * 4, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

// ---- order-entry tables, batches, results -----------------------------------

namespace {
static int tpcc_ctx(ltpg_ctx *ctx)
{
    int rc = check_ctx(ctx);
    if (rc) return rc;
    if (!ctx->tpcc.d.regs) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "context has no order-entry tables");
    return LTPG_OK;
}

static int tpcc_copy_table(ltpg_ctx *ctx, uint32_t table_id, uint32_t *host, uint64_t rows, uint32_t layout,
                    bool to_device)
{
    TpccState &tp = ctx->tpcc;
    if (table_id >= (uint32_t)kTpccTables) LTPG_FAIL(LTPG_ERR_INVALID, "unknown table id");
    if (!host) LTPG_FAIL(LTPG_ERR_INVALID, "null host table");
    if (layout != LTPG_LAYOUT_ROW_MAJOR && layout != LTPG_LAYOUT_COLUMN_MAJOR) LTPG_FAIL(LTPG_ERR_INVALID, "unknown layout");
    const TpccTable &tb = tp.d.t[table_id];
    if (rows != tb.rows) LTPG_FAIL(LTPG_ERR_INVALID, "rows does not match the table's row count");
    const uint64_t words = tb.rows * tb.cols;
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t *tmp = nullptr;
    int rc = dev_alloc(&tmp, words, "table staging");
    if (rc) return rc;
    hipError_t e = hipSuccess;
    if (to_device) {
        e = hipMemcpyAsync(tmp, host, words * 4, hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess) e = launch_tpcc_relayout(tmp, tb.data, tb.rows, tb.cols, tb.stride, layout, 1, ctx->stream);
        if (e == hipSuccess && table_id == LTPG_TPCC_CUSTOMER) e = launch_tpcc_name_index(tp.d, ctx->stream);
    } else {
        e = launch_tpcc_relayout(tb.data, tmp, tb.rows, tb.cols, tb.stride, layout, 0, ctx->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, tmp, words * 4, hipMemcpyDeviceToHost, ctx->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(tmp);
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("order-entry table copy: ") + hipGetErrorString(e));
    return LTPG_OK;
}
}  // namespace

int ltpg_tpcc_table_shape(ltpg_ctx *ctx, uint32_t table_id, uint64_t *rows, uint32_t *cols)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    if (table_id >= (uint32_t)kTpccTables) LTPG_FAIL(LTPG_ERR_INVALID, "unknown table id");
    if (rows) *rows = ctx->tpcc.d.t[table_id].rows;
    if (cols) *cols = ctx->tpcc.d.t[table_id].cols;
    return LTPG_OK;
}

int ltpg_tpcc_table_load(ltpg_ctx *ctx, uint32_t table_id, const uint32_t *host, uint64_t rows, uint32_t layout)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    return tpcc_copy_table(ctx, table_id, const_cast<uint32_t *>(host), rows, layout, true);
}

int ltpg_tpcc_table_read(ltpg_ctx *ctx, uint32_t table_id, uint32_t *host, uint64_t rows, uint32_t layout)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    return tpcc_copy_table(ctx, table_id, host, rows, layout, false);
}

int ltpg_tpcc_populate(ltpg_ctx *ctx, uint64_t seed)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    LTPG_HIP_TRY(launch_tpcc_populate(ctx->tpcc.d, seed, ctx->stream));
    LTPG_HIP_TRY(launch_tpcc_name_index(ctx->tpcc.d, ctx->stream));
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return LTPG_OK;
}

int ltpg_tpcc_submit(ltpg_ctx *ctx, const ltpg_neworder_batch *nb, const ltpg_payment_batch *pb, uint32_t mem)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    TpccState &tp = ctx->tpcc;
    if (mem != LTPG_MEM_HOST && mem != LTPG_MEM_DEVICE) LTPG_FAIL(LTPG_ERR_INVALID, "unknown memory kind");
    const uint32_t n_no = nb ? nb->n : 0, n_pay = pb ? pb->n : 0;
    if ((uint64_t)n_no + n_pay > ctx->cfg.max_txns) LTPG_FAIL(LTPG_ERR_INVALID, "batch exceeds max_txns");
    if (n_no) {
        if (nb->line_stride < n_no) LTPG_FAIL(LTPG_ERR_INVALID, "line_stride must be >= n");
        if (!nb->W_ID || !nb->D_ID || !nb->C_ID || !nb->O_ID || !nb->N_O_ID || !nb->O_OL_CNT || !nb->O_OL_ID ||
            !nb->OL_I_ID || !nb->OL_SUPPLY_W_ID || !nb->OL_QUANTITY || !nb->TID)
            LTPG_FAIL(LTPG_ERR_INVALID, "null NewOrder array");
    }
    if (n_pay) {
        if (!pb->W_ID || !pb->D_ID || !pb->C_ID || !pb->C_LAST || !pb->isName || !pb->C_D_ID || !pb->C_W_ID ||
            !pb->H_AMOUNT || !pb->H_ID || !pb->TID)
            LTPG_FAIL(LTPG_ERR_INVALID, "null Payment array");
    }
    if ((rc = next_epoch(ctx))) return rc;
    NeworderDev no;
    PaymentDev pay;
    no.n = n_no;
    pay.n = n_pay;
    if (mem == LTPG_MEM_DEVICE) {
        if (n_no) {
            no.line_stride = nb->line_stride;
            no.W_ID = nb->W_ID; no.D_ID = nb->D_ID; no.C_ID = nb->C_ID; no.O_ID = nb->O_ID; no.N_O_ID = nb->N_O_ID;
            no.O_OL_CNT = nb->O_OL_CNT; no.O_OL_ID = nb->O_OL_ID; no.OL_I_ID = nb->OL_I_ID;
            no.OL_SUPPLY_W_ID = nb->OL_SUPPLY_W_ID; no.OL_QUANTITY = nb->OL_QUANTITY; no.TID = nb->TID;
        }
        if (n_pay) {
            pay.W_ID = pb->W_ID; pay.D_ID = pb->D_ID; pay.C_ID = pb->C_ID; pay.C_LAST = pb->C_LAST;
            pay.isName = pb->isName; pay.C_D_ID = pb->C_D_ID; pay.C_W_ID = pb->C_W_ID; pay.H_AMOUNT = pb->H_AMOUNT;
            pay.H_ID = pb->H_ID; pay.TID = pb->TID;
        }
    } else {
        // pack into the library's NEWORDER_QUERY / PAYMENT_QUERY blocks of
        // the next host slot
        const uint64_t C = ctx->cfg.max_txns;
        int slot = 0;
        if ((rc = xfer_begin(ctx, &slot))) return rc;
        if ((rc = slot_buffer(&tp.no_buf[slot], 53 * C, "NewOrder batch")) ||
            (rc = slot_buffer(&tp.pay_buf[slot], 10 * C, "Payment batch")) ||
            (rc = slot_buffer(&tp.ro_buf[slot], 9 * C, "read-only batches")))
            return rc;
        hipStream_t s = ctx->copy_stream;
        if (n_no) {
            uint32_t *b = tp.no_buf[slot];
            const uint32_t *heads[7] = {nb->W_ID, nb->D_ID, nb->C_ID, nb->O_ID, nb->N_O_ID, nb->O_OL_CNT, nb->O_OL_ID};
            for (int k = 0; k < 7; k++)
                LTPG_HIP_TRY(hipMemcpyAsync(b + k * C, heads[k], (uint64_t)n_no * 4, hipMemcpyHostToDevice, s));
            const uint32_t *lines[3] = {nb->OL_I_ID, nb->OL_SUPPLY_W_ID, nb->OL_QUANTITY};
            for (int k = 0; k < 3; k++)
                LTPG_HIP_TRY(hipMemcpy2DAsync(b + (7 + 15 * k) * C, C * 4, lines[k], (uint64_t)nb->line_stride * 4,
                                              (uint64_t)n_no * 4, LTPG_TPCC_MAX_LINES, hipMemcpyHostToDevice, s));
            LTPG_HIP_TRY(hipMemcpyAsync(b + 52 * C, nb->TID, (uint64_t)n_no * 4, hipMemcpyHostToDevice, s));
            no.line_stride = (uint32_t)C;
            no.W_ID = b; no.D_ID = b + C; no.C_ID = b + 2 * C; no.O_ID = b + 3 * C; no.N_O_ID = b + 4 * C;
            no.O_OL_CNT = b + 5 * C; no.O_OL_ID = b + 6 * C; no.OL_I_ID = b + 7 * C; no.OL_SUPPLY_W_ID = b + 22 * C;
            no.OL_QUANTITY = b + 37 * C; no.TID = b + 52 * C;
        }
        if (n_pay) {
            uint32_t *b = tp.pay_buf[slot];
            const uint32_t *cols[10] = {pb->W_ID, pb->D_ID, pb->C_ID, pb->C_LAST, pb->isName,
                                        pb->C_D_ID, pb->C_W_ID, pb->H_AMOUNT, pb->H_ID, pb->TID};
            for (int k = 0; k < 10; k++)
                LTPG_HIP_TRY(hipMemcpyAsync(b + k * C, cols[k], (uint64_t)n_pay * 4, hipMemcpyHostToDevice, s));
            pay.W_ID = b; pay.D_ID = b + C; pay.C_ID = b + 2 * C; pay.C_LAST = b + 3 * C; pay.isName = b + 4 * C;
            pay.C_D_ID = b + 5 * C; pay.C_W_ID = b + 6 * C; pay.H_AMOUNT = b + 7 * C; pay.H_ID = b + 8 * C;
            pay.TID = b + 9 * C;
        }
        if ((rc = xfer_end(ctx))) return rc;
    }
    tp.host_slot = mem == LTPG_MEM_HOST;
    // Until a CID phase runs, each Payment uses its query's C_ID (a copy
    // kernel rather than a memcpy: a batched run captures this submit into a
    // graph, which then holds kernel nodes only)
    if (n_pay) LTPG_HIP_TRY(launch_copy_u32(tp.d.cid, pay.C_ID, n_pay, ctx->stream));
    tp.no = no;
    tp.pay = pay;
    tp.ro = ReadOnlyDev();
    tp.dl = DeliveryDev();
    tp.has_batch = true;
    return LTPG_OK;
}

int ltpg_tpcc_submit_readonly(ltpg_ctx *ctx, const ltpg_orderstatus_batch *ob, const ltpg_stocklevel_batch *sb,
                              uint32_t mem)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    TpccState &tp = ctx->tpcc;
    if (!tp.has_batch) LTPG_FAIL(LTPG_ERR_STATE, "ltpg_tpcc_submit first");
    if (mem != LTPG_MEM_HOST && mem != LTPG_MEM_DEVICE) LTPG_FAIL(LTPG_ERR_INVALID, "unknown memory kind");
    if (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS)
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "read-only transactions need the fast check (no LTPG_FLAG_CONFLICT_COUNTS)");
    const uint32_t n_os = ob ? ob->n : 0, n_sl = sb ? sb->n : 0;
    if ((uint64_t)tp.no.n + tp.pay.n + n_os + n_sl + tp.dl.n > ctx->cfg.max_txns)
        LTPG_FAIL(LTPG_ERR_INVALID, "batch exceeds max_txns");
    if (n_os && (!ob->W_ID || !ob->D_ID || !ob->C_ID || !ob->O_ID || !ob->OL_ID || !ob->TID))
        LTPG_FAIL(LTPG_ERR_INVALID, "null Order-Status array");
    if (n_sl && (!sb->W_ID || !sb->S_ID || !sb->TID)) LTPG_FAIL(LTPG_ERR_INVALID, "null Stock-Level array");
    ReadOnlyDev ro;
    ro.n_os = n_os;
    ro.n_sl = n_sl;
    const uint32_t *os_src[6] = {n_os ? ob->W_ID : nullptr, n_os ? ob->D_ID : nullptr, n_os ? ob->C_ID : nullptr,
                                 n_os ? ob->O_ID : nullptr, n_os ? ob->OL_ID : nullptr, n_os ? ob->TID : nullptr};
    const uint32_t *sl_src[3] = {n_sl ? sb->W_ID : nullptr, n_sl ? sb->S_ID : nullptr, n_sl ? sb->TID : nullptr};
    const uint32_t *os_dev[6], *sl_dev[3];
    if (mem == LTPG_MEM_DEVICE) {
        for (int k = 0; k < 6; k++) os_dev[k] = os_src[k];
        for (int k = 0; k < 3; k++) sl_dev[k] = sl_src[k];
    } else {
        // the epoch's host slot (its NewOrders and Payments took it), or, if
        // they were device-resident, the next one
        const uint64_t C = ctx->cfg.max_txns;
        int slot = ctx->xfer_slot;
        if (!tp.host_slot && (rc = xfer_begin(ctx, &slot))) return rc;
        tp.host_slot = true;
        if ((rc = slot_buffer(&tp.ro_buf[slot], 9 * C, "read-only batches"))) return rc;
        uint32_t *rb = tp.ro_buf[slot];
        for (int k = 0; k < 6; k++) {
            os_dev[k] = rb + k * C;
            if (n_os)
                LTPG_HIP_TRY(hipMemcpyAsync(rb + k * C, os_src[k], (uint64_t)n_os * 4, hipMemcpyHostToDevice,
                                            ctx->copy_stream));
        }
        for (int k = 0; k < 3; k++) {
            sl_dev[k] = rb + (6 + k) * C;
            if (n_sl)
                LTPG_HIP_TRY(hipMemcpyAsync(rb + (6 + k) * C, sl_src[k], (uint64_t)n_sl * 4, hipMemcpyHostToDevice,
                                            ctx->copy_stream));
        }
        if ((rc = xfer_end(ctx))) return rc;
    }
    ro.os_W_ID = os_dev[0]; ro.os_D_ID = os_dev[1]; ro.os_C_ID = os_dev[2];
    ro.os_O_ID = os_dev[3]; ro.os_OL_ID = os_dev[4]; ro.os_TID = os_dev[5];
    ro.sl_W_ID = sl_dev[0]; ro.sl_S_ID = sl_dev[1]; ro.sl_TID = sl_dev[2];
    tp.ro = ro;
    return LTPG_OK;
}

int ltpg_tpcc_submit_delivery(ltpg_ctx *ctx, const ltpg_delivery_batch *db, uint32_t mem)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    TpccState &tp = ctx->tpcc;
    if (!tp.has_batch) LTPG_FAIL(LTPG_ERR_STATE, "ltpg_tpcc_submit first");
    if (mem != LTPG_MEM_HOST && mem != LTPG_MEM_DEVICE) LTPG_FAIL(LTPG_ERR_INVALID, "unknown memory kind");
    if (ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS)
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "Deliveries need the fast check (no LTPG_FLAG_CONFLICT_COUNTS)");
    const uint32_t n = db ? db->n : 0;
    if ((uint64_t)tp.no.n + tp.pay.n + tp.ro.n_os + tp.ro.n_sl + n > ctx->cfg.max_txns)
        LTPG_FAIL(LTPG_ERR_INVALID, "batch exceeds max_txns");
    if (n && (!db->O_ID || !db->NO_W_ID || !db->NO_O_ID || !db->TID)) LTPG_FAIL(LTPG_ERR_INVALID, "null Delivery array");
    const uint32_t *src[4] = {n ? db->O_ID : nullptr, n ? db->NO_W_ID : nullptr, n ? db->NO_O_ID : nullptr,
                              n ? db->TID : nullptr};
    const uint32_t *dev[4];
    if (mem == LTPG_MEM_DEVICE || n == 0) {
        for (int k = 0; k < 4; k++) dev[k] = src[k];
    } else {
        // the epoch's host slot, or, if its batch was device-resident, the next one
        const uint64_t C = ctx->cfg.max_txns;
        int slot = ctx->xfer_slot;
        if (!tp.host_slot && (rc = xfer_begin(ctx, &slot))) return rc;
        tp.host_slot = true;
        if ((rc = slot_buffer(&tp.dl_buf[slot], 4 * C, "Delivery batches"))) return rc;
        uint32_t *b = tp.dl_buf[slot];
        for (int k = 0; k < 4; k++) {
            dev[k] = b + k * C;
            LTPG_HIP_TRY(hipMemcpyAsync(b + k * C, src[k], (uint64_t)n * 4, hipMemcpyHostToDevice, ctx->copy_stream));
        }
        if ((rc = xfer_end(ctx))) return rc;
    }
    DeliveryDev dl;
    dl.n = n;
    dl.O_ID = dev[0];
    dl.NO_W_ID = dev[1];
    dl.NO_O_ID = dev[2];
    dl.TID = dev[3];
    tp.dl = dl;
    return LTPG_OK;
}

int ltpg_tpcc_payment_cid(ltpg_ctx *ctx, uint32_t *out, uint32_t n)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    if (n > ctx->tpcc.pay.n) LTPG_FAIL(LTPG_ERR_INVALID, "n exceeds the Payment batch size");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) LTPG_HIP_TRY(hipMemcpy(out, ctx->tpcc.d.cid, (uint64_t)n * 4, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

int ltpg_tpcc_read_values(ltpg_ctx *ctx, uint32_t type, uint32_t *out, uint64_t n_words)
{
    int rc = tpcc_ctx(ctx);
    if (rc) return rc;
    if (!out) LTPG_FAIL(LTPG_ERR_INVALID, "null output");
    const uint32_t *src = nullptr;
    uint64_t have = 0;
    if (type == LTPG_TYPE_NEWORDER) {
        src = ctx->tpcc.d.reads_no;
        have = (uint64_t)kNoReads * ctx->tpcc.no.n;
    } else if (type == LTPG_TYPE_PAYMENT) {
        src = ctx->tpcc.d.reads_pay;
        have = (uint64_t)kPayReads * ctx->tpcc.pay.n;
    } else if (type == LTPG_TYPE_ORDERSTATUS) {
        src = ctx->tpcc.d.reads_os;
        have = (uint64_t)kOsReads * ctx->tpcc.ro.n_os;
    } else if (type == LTPG_TYPE_STOCKLEVEL) {
        src = ctx->tpcc.d.reads_sl;
        have = (uint64_t)kSlReads * ctx->tpcc.ro.n_sl;
    } else if (type == LTPG_TYPE_DELIVERY) {
        src = ctx->tpcc.d.reads_dl;
        have = (uint64_t)kDlReads * ctx->tpcc.dl.n;
    } else {
        LTPG_FAIL(LTPG_ERR_INVALID, "type must be an order-entry type");
    }
    if (n_words > have) LTPG_FAIL(LTPG_ERR_INVALID, "n_words exceeds the batch's read values");
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_words) LTPG_HIP_TRY(hipMemcpy(out, src, n_words * 4, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

}  // extern "C"
