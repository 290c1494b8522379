// batched.cpp - the batched form: many batches of uneven size and mixed
// types in one submission.
//
// The reference drives one epoch at a time from the host: launch a phase,
// synchronise, launch the next, copy COMMIT back, prefetch the next batch on
// a helper thread (LTPG/TPCC-allopt-final/tp.cu:136-261, Query.h:640-725).
// Here the batches are resident in device memory and each workload's whole
// sequence of epochs is captured once into a HIP graph: a register refill
// (so that every replay starts from the same register state and uses the
// same epoch tags), then per batch its submit, the phase kernels and a copy
// of its commit bitmap(s) into a device archive.  The key-value and the
// order-entry graphs replay concurrently on their contexts' streams (the two
// workloads share no table), with one host synchronisation for the lot.
//
// Key-value epochs are pipelined inside their graph.  EXECUTE (write
// registration + read marking) touches only the registers, the flags and the
// batch, never the table; CHECK and WRITEBACK touch the table.  With two
// register sets (and two sets of doomed / raw / war flags) used by alternate
// batches, EXECUTE of batch j+1 runs on a side stream while CHECK and
// WRITEBACK of batch j run on the main stream; EXECUTE of batch j+2 waits
// for CHECK of batch j, the last reader of that set.  Each batch keeps its
// own epoch tag, so the result is the plain loop's, bit for bit.
#include <hip/hip_runtime.h>

#include <chrono>
#include <new>
#include <string>
#include <vector>

#include "ltpg_internal.h"

using namespace ltpg;

struct ltpg_batched {
    ltpg_ctx *ctx[2] = {nullptr, nullptr};  // [0] key-value, [1] order-entry
    std::vector<ltpg_batch_desc> d;
    std::vector<uint64_t> off;              // archive offset (u64 words) of batch i's first bitmap
    uint64_t *archive = nullptr;
    uint64_t archive_words = 0;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint32_t final_epoch[2] = {0, 0};       // host epoch counter after a replay
    std::vector<uint8_t> tune;              // batch i: its LTPG_TUNE_* flags (key-value or order-entry)
    unsigned long long *stamps = nullptr;   // graph runs: device clock after batch i; [n + k] start of graph k
    std::vector<double> plain_ms;           // plain runs: host ms from the run's start to batch i's readback
    // key-value pipelining: side stream for EXECUTE, second register / flag set
    hipStream_t side = nullptr;
    unsigned long long *regs2 = nullptr;
    uint8_t *doomed2 = nullptr, *raw2 = nullptr, *war2 = nullptr;
    std::vector<hipEvent_t> events;         // capture dependencies (kept until destroy)
    uint32_t last_mode = LTPG_RUN_PLAIN;
    bool ran = false;
};

namespace {

uint64_t words(uint64_t n) { return (n + 63) / 64; }

int slot_of(uint32_t workload) { return workload == LTPG_WORKLOAD_KV ? 0 : 1; }

// One epoch of batch i on its context: submit (device-resident), all
// phases, archive copy of the commit bitmap(s), and (captured runs) a device
// clock stamp.  No synchronisation, so it can be captured -- unless
// `stepwise`: then the phases run one at a time with a synchronisation after
// each, as the reference's driver loop does (the plain loop).
int enqueue_batch(ltpg_batched *b, uint32_t i, bool stamp, bool stepwise = false)
{
    const ltpg_batch_desc &d = b->d[i];
    ltpg_ctx *ctx = b->ctx[slot_of(d.workload)];
    int rc;
    if (d.workload == LTPG_WORKLOAD_KV) {
        if ((rc = ltpg_kv_set_tuning(ctx, b->tune[i]))) return rc;
        if ((rc = ltpg_kv_submit(ctx, &d.kv, LTPG_MEM_DEVICE))) return rc;
    } else {
        if ((rc = ltpg_tpcc_set_tuning(ctx, b->tune[i]))) return rc;
        if ((rc = ltpg_tpcc_submit(ctx, &d.neworder, &d.payment, LTPG_MEM_DEVICE))) return rc;
        if ((d.orderstatus.n || d.stocklevel.n) &&
            (rc = ltpg_tpcc_submit_readonly(ctx, &d.orderstatus, &d.stocklevel, LTPG_MEM_DEVICE)))
            return rc;
        if (d.delivery.n && (rc = ltpg_tpcc_submit_delivery(ctx, &d.delivery, LTPG_MEM_DEVICE))) return rc;
    }
    if ((rc = stepwise ? run_epoch_stepwise(ctx) : ltpg_epoch_run(ctx, LTPG_PHASE_ALL))) return rc;
    if (d.workload == LTPG_WORKLOAD_KV) {
        LTPG_HIP_TRY(launch_copy_u64(b->archive + b->off[i], ctx->kv.bitmap, words(d.kv.n_txns), ctx->stream));
    } else {
        LTPG_HIP_TRY(launch_copy_u64(b->archive + b->off[i], ctx->tpcc.d.bitmap_no, words(d.neworder.n),
                                     ctx->stream));
        uint64_t at = b->off[i] + words(d.neworder.n);
        LTPG_HIP_TRY(launch_copy_u64(b->archive + at, ctx->tpcc.d.bitmap_pay, words(d.payment.n), ctx->stream));
        at += words(d.payment.n);
        LTPG_HIP_TRY(launch_copy_u64(b->archive + at, ctx->tpcc.d.bitmap_os, words(d.orderstatus.n), ctx->stream));
        at += words(d.orderstatus.n);
        LTPG_HIP_TRY(launch_copy_u64(b->archive + at, ctx->tpcc.d.bitmap_sl, words(d.stocklevel.n), ctx->stream));
        at += words(d.stocklevel.n);
        LTPG_HIP_TRY(launch_copy_u64(b->archive + at, ctx->tpcc.d.bitmap_dl, words(d.delivery.n), ctx->stream));
    }
    if (stamp) LTPG_HIP_TRY(launch_stamp(b->stamps + i, ctx->stream));
    return LTPG_OK;
}

hipError_t new_event(ltpg_batched *b, hipEvent_t *e)
{
    const hipError_t r = hipEventCreateWithFlags(e, hipEventDisableTiming);
    if (r == hipSuccess) b->events.push_back(*e);
    return r;
}

// the second register set and its flags back to their initial state
int refill_second_set(ltpg_batched *b, ltpg_ctx *ctx)
{
    LTPG_HIP_TRY(launch_fill_u64(b->regs2, ctx->kv.rows * 2, ~0ull, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(b->doomed2, ctx->cfg.max_txns, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(b->raw2, ctx->cfg.max_txns, 0, ctx->stream));
    LTPG_HIP_TRY(launch_fill_u8(b->war2, ctx->cfg.max_txns, 0, ctx->stream));
    return LTPG_OK;
}

// Capture topology: side joins the capture by a fork from main (an event
// recorded on the capturing main stream, which side then waits for), and
// rejoins main by a join (an event recorded on side, which main waits for).
// Every event is recorded inside the capture before it is waited for, and
// the capture ends on main after a join, with nothing left on side.
int fork_side(ltpg_batched *b, hipStream_t main, hipStream_t side)
{
    hipEvent_t e;
    LTPG_HIP_TRY(new_event(b, &e));
    LTPG_HIP_TRY(hipEventRecord(e, main));
    LTPG_HIP_TRY(hipStreamWaitEvent(side, e, 0));
    return LTPG_OK;
}

int join_side(ltpg_batched *b, hipStream_t main, hipStream_t side)
{
    hipEvent_t e;
    LTPG_HIP_TRY(new_event(b, &e));
    LTPG_HIP_TRY(hipEventRecord(e, side));
    LTPG_HIP_TRY(hipStreamWaitEvent(main, e, 0));
    return LTPG_OK;
}

// The key-value epochs of the runner, pipelined over two streams (captured).
int enqueue_kv_pipelined(ltpg_batched *b)
{
    ltpg_ctx *ctx = b->ctx[0];
    hipStream_t main = ctx->stream, side = b->side;
    int rc;
    if ((rc = refill_second_set(b, ctx))) return rc;
    if ((rc = fork_side(b, main, side))) return rc;  // side joins the capture after the refills
    std::vector<hipEvent_t> checked;                 // per key-value batch: its CHECK is done
    uint32_t j = 0;
    for (uint32_t i = 0; i < b->d.size(); i++) {
        const ltpg_batch_desc &d = b->d[i];
        if (d.workload != LTPG_WORKLOAD_KV) continue;
        if ((rc = ltpg_kv_set_tuning(ctx, b->tune[i]))) return rc;
        if (ctx->epoch >= 255) {
            // the submit below wraps the epoch and refills the first set on
            // main: nothing may be in flight on side, and the second set
            // starts over too
            if ((rc = join_side(b, main, side))) return rc;
            if ((rc = refill_second_set(b, ctx))) return rc;
        }
        const uint32_t before = ctx->epoch;
        if ((rc = ltpg_kv_submit(ctx, &d.kv, LTPG_MEM_DEVICE))) return rc;
        if (ctx->epoch < before) {
            if ((rc = fork_side(b, main, side))) return rc;  // side waits for the refills
            checked.assign(checked.size(), nullptr);        // older checks precede the refill
        }
        EpochTag e;
        e.epoch = (uint8_t)ctx->epoch;
        e.tag = 255u - ctx->epoch;
        KvState st = ctx->kv;
        if (j & 1) {
            st.regs = b->regs2;
            st.doomed = b->doomed2;
            st.raw_flag = b->raw2;
            st.war_flag = b->war2;
        }
        if (j >= 2 && checked[j - 2]) LTPG_HIP_TRY(hipStreamWaitEvent(side, checked[j - 2], 0));
        LTPG_HIP_TRY(launch_kv_execute(st, st.cur, e, false, ctx->counters, side, false));
        hipEvent_t executed, done;
        LTPG_HIP_TRY(new_event(b, &executed));
        LTPG_HIP_TRY(hipEventRecord(executed, side));
        LTPG_HIP_TRY(hipStreamWaitEvent(main, executed, 0));
        LTPG_HIP_TRY(launch_reset_epoch_counters(ctx->counters, main));
        LTPG_HIP_TRY(launch_kv_check(st, st.cur, e, false, ctx->counters, main));
        LTPG_HIP_TRY(new_event(b, &done));
        LTPG_HIP_TRY(hipEventRecord(done, main));
        checked.push_back(done);
        LTPG_HIP_TRY(launch_kv_writeback(st, st.cur, ctx->cfg.values, main));
        LTPG_HIP_TRY(launch_copy_u64(b->archive + b->off[i], ctx->kv.bitmap, words(d.kv.n_txns), main));
        LTPG_HIP_TRY(launch_stamp(b->stamps + i, main));
        j++;
    }
    return join_side(b, main, side);  // side rejoins main before the capture ends
}

// Capture the epochs of workload slot k into a graph.
int capture(ltpg_batched *b, int k)
{
    ltpg_ctx *ctx = b->ctx[k];
    ctx->kv.regs_clean = false;  // replays write registers outside ltpg_epoch_run
    const bool timing = ctx->timing;
    const uint32_t epoch = ctx->epoch;  // capture runs nothing: the device state keeps this epoch
    ctx->timing = false;
    LTPG_HIP_TRY(hipStreamSynchronize(ctx->stream));
    LTPG_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    // every replay starts from all-ones registers and epoch 1
    ctx->epoch = 0;
    int rc = refill_registers(ctx);
    if (!rc && launch_stamp(b->stamps + b->d.size() + k, ctx->stream) != hipSuccess) rc = LTPG_ERR_DEVICE;
    if (k == 0 && b->side) {
        if (!rc) rc = enqueue_kv_pipelined(b);
    } else {
        for (uint32_t i = 0; i < b->d.size() && !rc; i++)
            if (slot_of(b->d[i].workload) == k) rc = enqueue_batch(b, i, true);
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    b->final_epoch[k] = ctx->epoch;
    ctx->epoch = epoch;
    ctx->timing = timing;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) LTPG_FAIL(LTPG_ERR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    b->graph[k] = g;
    LTPG_HIP_TRY(hipGraphInstantiate(&b->exec[k], g, nullptr, nullptr, 0));
    return LTPG_OK;
}

}  // namespace

extern "C" {

int ltpg_batched_destroy(ltpg_batched *b)
{
    if (!b) return LTPG_OK;
    for (int k = 0; k < 2; k++) {
        if (b->ctx[k]) (void)hipStreamSynchronize(b->ctx[k]->stream);
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
    }
    if (b->side) (void)hipStreamSynchronize(b->side);
    for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
    if (b->side) (void)hipStreamDestroy(b->side);
    dev_free(b->regs2);
    dev_free(b->doomed2);
    dev_free(b->raw2);
    dev_free(b->war2);
    dev_free(b->archive);
    dev_free(b->stamps);
    delete b;
    return LTPG_OK;
}

int ltpg_batched_create(ltpg_ctx *kv_ctx, ltpg_ctx *tpcc_ctx, const ltpg_batch_desc *descs, uint32_t n,
                        ltpg_batched **out)
{
    if (!out || (!descs && n)) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (kv_ctx && !kv_ctx->kv.table) LTPG_FAIL(LTPG_ERR_INVALID, "kv_ctx is not a key-value context");
    if (tpcc_ctx && !tpcc_ctx->tpcc.d.regs) LTPG_FAIL(LTPG_ERR_INVALID, "tpcc_ctx is not an order-entry context");
    if (kv_ctx && tpcc_ctx && kv_ctx->device != tpcc_ctx->device)
        LTPG_FAIL(LTPG_ERR_INVALID, "both contexts must be on one device");
    ltpg_batched *b = new (std::nothrow) ltpg_batched();
    if (!b) LTPG_FAIL(LTPG_ERR_NOMEM, "host allocation of batched runner failed");
    b->ctx[0] = kv_ctx;
    b->ctx[1] = tpcc_ctx;
    b->d.assign(descs, descs + n);
    b->off.resize(n);
    b->tune.assign(n, 0);
    uint64_t w = 0;
    for (uint32_t i = 0; i < n; i++) {
        const ltpg_batch_desc &d = b->d[i];
        b->off[i] = w;
        if (d.workload == LTPG_WORKLOAD_KV) {
            if (!kv_ctx) { ltpg_batched_destroy(b); LTPG_FAIL(LTPG_ERR_INVALID, "key-value batch without kv_ctx"); }
            w += words(d.kv.n_txns);
        } else if (d.workload == LTPG_WORKLOAD_TPCC) {
            if (!tpcc_ctx) { ltpg_batched_destroy(b); LTPG_FAIL(LTPG_ERR_INVALID, "order-entry batch without tpcc_ctx"); }
            w += words(d.neworder.n) + words(d.payment.n) + words(d.orderstatus.n) + words(d.stocklevel.n) +
                 words(d.delivery.n);
        } else {
            ltpg_batched_destroy(b);
            LTPG_FAIL(LTPG_ERR_INVALID, "unknown workload in batch descriptor");
        }
    }
    b->archive_words = w;
    int rc = check_ctx(kv_ctx ? kv_ctx : tpcc_ctx);
    if (!rc) rc = dev_alloc(&b->archive, w ? w : 1, "bitmap archive");
    if (!rc) rc = dev_alloc(&b->stamps, (uint64_t)n + 2, "batch clock stamps");
    b->plain_ms.assign(n, 0.0);
    bool any_kv = false;
    for (const auto &d : b->d) any_kv |= d.workload == LTPG_WORKLOAD_KV;
    if (!rc && any_kv && !(kv_ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS)) {
        // the counting form has no read marking to pipeline; it runs unpipelined
        const uint64_t B = kv_ctx->cfg.max_txns;
        if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess) {
            b->side = nullptr;
            set_error("hipStreamCreateWithFlags failed");
            rc = LTPG_ERR_DEVICE;
        }
        if (!rc) rc = dev_alloc(&b->regs2, kv_ctx->kv.rows * 2, "second register set");
        if (!rc) rc = dev_alloc(&b->doomed2, B, "second doomed flags");
        if (!rc) rc = dev_alloc(&b->raw2, B, "second raw flags");
        if (!rc) rc = dev_alloc(&b->war2, B, "second war flags");
    }
    for (int k = 0; k < 2 && !rc; k++) {
        bool any = false;
        for (const auto &d : b->d) any |= slot_of(d.workload) == k;
        if (any) rc = capture(b, k);
    }
    if (rc) {
        const std::string msg = ltpg_last_error();
        ltpg_batched_destroy(b);
        set_error(msg);
        return rc;
    }
    *out = b;
    return LTPG_OK;
}

int ltpg_batched_run(ltpg_batched *b, uint32_t mode)
{
    if (!b) LTPG_FAIL(LTPG_ERR_INVALID, "null runner");
    int rc = check_ctx(b->ctx[0] ? b->ctx[0] : b->ctx[1]);
    if (rc) return rc;
    if (mode == LTPG_RUN_PLAIN) {
        // the reference's loop: per batch its phases, each launch followed
        // by a synchronisation, then the commit count back on the host
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t i = 0; i < b->d.size(); i++) {
            ltpg_ctx *ctx = b->ctx[slot_of(b->d[i].workload)];
            if ((rc = enqueue_batch(b, i, false, true))) return rc;
            uint32_t c = 0;
            if ((rc = ltpg_commit_count(ctx, b->d[i].workload == LTPG_WORKLOAD_KV ? LTPG_TYPE_KV
                                                                                  : LTPG_TYPE_PAYMENT, &c)))
                return rc;
            b->plain_ms[i] =
                std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
        b->last_mode = LTPG_RUN_PLAIN;
        b->ran = true;
        return LTPG_OK;
    }
    if (mode != LTPG_RUN_GRAPH) LTPG_FAIL(LTPG_ERR_INVALID, "unknown run mode");
    for (int k = 0; k < 2; k++) {
        if (!b->exec[k]) continue;
        LTPG_HIP_TRY(hipGraphLaunch(b->exec[k], b->ctx[k]->stream));
        // the registers now carry the captured sequence's epoch tags
        b->ctx[k]->epoch = b->final_epoch[k];
        b->ctx[k]->kv.regs_clean = false;
    }
    int first = LTPG_OK;
    for (int k = 0; k < 2; k++) {
        if (!b->exec[k]) continue;
        rc = ltpg_sync(b->ctx[k]);
        if (rc && !first) first = rc;
    }
    b->last_mode = LTPG_RUN_GRAPH;
    b->ran = true;
    return first;
}

int ltpg_batched_autotune(ltpg_batched *b, uint32_t reps, ltpg_tuning_entry *out, uint32_t max_out,
                          uint32_t *n_out)
{
    if (!b) LTPG_FAIL(LTPG_ERR_INVALID, "null runner");
    if (n_out) *n_out = 0;
    if (reps == 0) reps = 1;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    LTPG_HIP_TRY(hipEventCreate(&e0));
    LTPG_HIP_TRY(hipEventCreate(&e1));
    int rc = LTPG_OK;
    uint32_t n = 0;
    for (int k = 0; k < 2 && !rc; k++) {
        ltpg_ctx *ctx = b->ctx[k];
        if (!ctx) continue;
        if ((rc = check_ctx(ctx))) break;
        const uint32_t workload = k == 0 ? LTPG_WORKLOAD_KV : LTPG_WORKLOAD_TPCC;
        // size classes and their first batch
        std::vector<int> rep(33, -1);
        for (uint32_t i = 0; i < b->d.size(); i++) {
            const ltpg_batch_desc &d = b->d[i];
            if (d.workload != workload) continue;
            const uint32_t txns = k == 0 ? d.kv.n_txns
                                         : d.neworder.n + d.payment.n + d.orderstatus.n + d.stocklevel.n +
                                               d.delivery.n;
            if (txns == 0) continue;
            const int c = 31 - __builtin_clz(txns);
            if (rep[c] < 0) rep[c] = (int)i;
        }
        if (k == 0) ctx->kv.regs_clean = false;
        const uint32_t values = ctx->cfg.values;
        ctx->cfg.values = LTPG_VALUES_COPYBACK;  // timing runs leave the tables as they are
        const bool timing = ctx->timing;
        ctx->timing = false;
        // the distinct launch shapes of the workload
        std::vector<uint32_t> shapes;
        if (k == 0) {
            const bool part_ok = kv_part_prepare(ctx) == LTPG_OK;
            for (uint32_t idx = 0; idx < 64; idx++) {
                const bool part = idx & 16, more = idx & 32;
                // EXEC_PER_OP and MORE_ITEMS shape only the register path
                if (part && (!part_ok || (idx & LTPG_TUNE_EXEC_PER_OP) || more)) continue;
                shapes.push_back((idx & LTPG_TUNE_ALL) | (part ? LTPG_TUNE_PARTITIONED : 0u) |
                                 (more ? LTPG_TUNE_MORE_ITEMS : 0u));
            }
        } else {
            shapes = {0u, LTPG_TUNE_TPCC_SMALL_BLOCKS};
        }
        for (int c = 0; c < 33 && !rc; c++) {
            if (rep[c] < 0) continue;
            const uint32_t i = (uint32_t)rep[c];
            float ms[64] = {};
            uint32_t best = shapes[0];
            float best_ms = -1.f;
            for (uint32_t f : shapes) {
                for (uint32_t r = 0; r <= reps && !rc; r++) {
                    if (r == 1) rc = hipEventRecord(e0, ctx->stream) == hipSuccess ? LTPG_OK : LTPG_ERR_DEVICE;
                    if (rc) break;
                    // submit + phases of the batch under shape f (no archive copy)
                    const uint8_t keep = b->tune[i];
                    b->tune[i] = (uint8_t)f;
                    rc = enqueue_batch(b, i, false);
                    b->tune[i] = keep;
                }
                if (!rc && hipEventRecord(e1, ctx->stream) != hipSuccess) rc = LTPG_ERR_DEVICE;
                if (!rc) rc = ltpg_sync(ctx);
                float t = 0;
                if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = LTPG_ERR_DEVICE;
                if (rc) break;
                t /= (float)reps;
                ms[k == 0 ? LTPG_TUNE_INDEX(f) : (f ? 1u : 0u)] = t;
                if (best_ms < 0.f || t < best_ms) {
                    best_ms = t;
                    best = f;
                }
            }
            if (rc) break;
            for (uint32_t j = 0; j < b->d.size(); j++) {
                const ltpg_batch_desc &d = b->d[j];
                if (d.workload != workload) continue;
                const uint32_t txns = k == 0 ? d.kv.n_txns
                                             : d.neworder.n + d.payment.n + d.orderstatus.n + d.stocklevel.n +
                                                   d.delivery.n;
                if (txns && 31 - __builtin_clz(txns) == c) b->tune[j] = (uint8_t)best;
            }
            if (out && n < max_out) {
                ltpg_tuning_entry &o = out[n];
                o.workload = workload;
                o.log2_txns = (uint32_t)c;
                o.flags = best;
                o.reserved = 0;
                for (int m = 0; m < 64; m++) o.ms[m] = ms[m];
            }
            n++;
        }
        ctx->cfg.values = values;
        ctx->timing = timing;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (n_out) *n_out = n < max_out ? n : max_out;
    // capture both sequences again with the chosen shapes
    for (int k = 0; k < 2; k++) {
        if (!b->exec[k] && !b->graph[k]) continue;
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
        if ((rc = capture(b, k))) return rc;
    }
    return LTPG_OK;
}

int ltpg_batched_set_tuning(ltpg_batched *b, uint32_t flags)
{
    if (!b) LTPG_FAIL(LTPG_ERR_INVALID, "null runner");
    if (flags & ~(LTPG_TUNE_ALL | LTPG_TUNE_PARTITIONED | LTPG_TUNE_MORE_ITEMS | LTPG_TUNE_TPCC_SMALL_BLOCKS))
        LTPG_FAIL(LTPG_ERR_INVALID, "unknown tuning flag");
    const uint32_t kv_flags = flags & ~LTPG_TUNE_TPCC_SMALL_BLOCKS, tp_flags = flags & LTPG_TUNE_TPCC_SMALL_BLOCKS;
    int rc;
    // validates the flags and prepares their buffers outside any capture
    if (b->ctx[0] && ((rc = check_ctx(b->ctx[0])) || (rc = ltpg_kv_set_tuning(b->ctx[0], kv_flags)))) return rc;
    if (b->ctx[1] && ((rc = check_ctx(b->ctx[1])) || (rc = ltpg_tpcc_set_tuning(b->ctx[1], tp_flags)))) return rc;
    for (uint32_t i = 0; i < b->d.size(); i++)
        b->tune[i] = (uint8_t)(b->d[i].workload == LTPG_WORKLOAD_KV ? kv_flags : tp_flags);
    for (int k = 0; k < 2; k++) {
        if (!b->exec[k] && !b->graph[k]) continue;
        if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
        if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
        b->exec[k] = nullptr;
        b->graph[k] = nullptr;
        if ((rc = capture(b, k))) return rc;
    }
    return LTPG_OK;
}

int ltpg_batched_times(ltpg_batched *b, double *ms, uint32_t n)
{
    if (!b || (!ms && n)) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    if (n > b->d.size()) LTPG_FAIL(LTPG_ERR_INVALID, "n exceeds the number of batches");
    if (!b->ran) LTPG_FAIL(LTPG_ERR_STATE, "no run yet");
    if (b->last_mode == LTPG_RUN_PLAIN) {
        for (uint32_t i = 0; i < n; i++) ms[i] = b->plain_ms[i];
        return LTPG_OK;
    }
    ltpg_ctx *ctx = b->ctx[0] ? b->ctx[0] : b->ctx[1];
    int rc = check_ctx(ctx);
    if (rc) return rc;
    int khz = 0;
    LTPG_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device));
    if (khz <= 0) LTPG_FAIL(LTPG_ERR_DEVICE, "device reports no wall-clock rate");
    std::vector<unsigned long long> st(b->d.size() + 2);
    LTPG_HIP_TRY(hipMemcpy(st.data(), b->stamps, st.size() * 8, hipMemcpyDeviceToHost));
    unsigned long long start = ~0ull;
    for (int k = 0; k < 2; k++)
        if (b->exec[k] && st[b->d.size() + k] < start) start = st[b->d.size() + k];
    for (uint32_t i = 0; i < n; i++) ms[i] = (double)(st[i] - start) / (double)khz;
    return LTPG_OK;
}

int ltpg_batched_bitmap(ltpg_batched *b, uint32_t batch, uint32_t type, uint64_t *out, uint64_t n_words)
{
    if (!b || !out) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    if (batch >= b->d.size()) LTPG_FAIL(LTPG_ERR_INVALID, "batch index out of range");
    const ltpg_batch_desc &d = b->d[batch];
    uint64_t off = b->off[batch], n = 0;
    if (d.workload == LTPG_WORKLOAD_KV && type == LTPG_TYPE_KV) {
        n = words(d.kv.n_txns);
    } else if (d.workload == LTPG_WORKLOAD_TPCC && type == LTPG_TYPE_NEWORDER) {
        n = words(d.neworder.n);
    } else if (d.workload == LTPG_WORKLOAD_TPCC && type == LTPG_TYPE_PAYMENT) {
        off += words(d.neworder.n);
        n = words(d.payment.n);
    } else if (d.workload == LTPG_WORKLOAD_TPCC && type == LTPG_TYPE_ORDERSTATUS) {
        off += words(d.neworder.n) + words(d.payment.n);
        n = words(d.orderstatus.n);
    } else if (d.workload == LTPG_WORKLOAD_TPCC && type == LTPG_TYPE_STOCKLEVEL) {
        off += words(d.neworder.n) + words(d.payment.n) + words(d.orderstatus.n);
        n = words(d.stocklevel.n);
    } else if (d.workload == LTPG_WORKLOAD_TPCC && type == LTPG_TYPE_DELIVERY) {
        off += words(d.neworder.n) + words(d.payment.n) + words(d.orderstatus.n) + words(d.stocklevel.n);
        n = words(d.delivery.n);
    } else {
        LTPG_FAIL(LTPG_ERR_INVALID, "type does not match the batch's workload");
    }
    if (n_words < n) LTPG_FAIL(LTPG_ERR_INVALID, "bitmap buffer too small");
    int rc = check_ctx(b->ctx[slot_of(d.workload)]);
    if (rc) return rc;
    for (int k = 0; k < 2; k++)
        if (b->ctx[k]) LTPG_HIP_TRY(hipStreamSynchronize(b->ctx[k]->stream));
    if (n) LTPG_HIP_TRY(hipMemcpy(out, b->archive + off, n * 8, hipMemcpyDeviceToHost));
    return LTPG_OK;
}

}  // extern "C"
