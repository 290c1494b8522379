// serve.cpp - the served form: a request loop over resident tables.
//
// The reference feeds its epoch loop from generator threads that fill
// pre-allocated query batches, one epoch at a time (LTPG/TPCC-allopt-final/
// Query.h:640-725, tp.cu:136-261).  Here requests arrive individually
// (submission = arrival) and a serving thread forms epochs continuously:
//
//   admission  pending requests of a workload -- re-queued aborts first, in
//              priority (TID) order, then new arrivals in arrival order -- fill
//              the next epoch up to the per-epoch transaction budget, derived
//              from the HBM budget of the access-set buffers (batch slot and
//              read values per transaction);
//   overlap    each workload has `slots` epochs in flight: while one runs on
//              the context's stream, the next is packed into pinned staging
//              and queued behind it; key-value and order-entry epochs run on
//              their own contexts' streams at the same time;
//   results    the epoch's last kernel writes its commit bitmaps and then a
//              sequence number into pinned host memory, which the loop polls
//              (no event or copy call per epoch);
//   requeue    an aborted transaction keeps the TID given at its arrival and
//              joins the next epoch formed after its result is read, ahead of
//              every later arrival, so it commits after finitely many epochs;
//   latency    commit time (result read on the host) minus arrival time.
//
// Tables and registers stay resident in the contexts; nothing is reset
// between epochs (epoch-tagged registers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <queue>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ltpg_internal.h"

using namespace ltpg;

namespace {

uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

uint64_t words(uint64_t n) { return (n + 63) / 64; }

constexpr uint32_t kNoFields = 52;  // NewOrder request: 7 head fields + 3 x 15 line fields
constexpr uint32_t kPayFields = 9;
constexpr uint32_t kOsFields = 5;    // Order-Status: W_ID, D_ID, C_ID, O_ID, OL_ID
constexpr uint32_t kSlFields = 2;    // Stock-Level: W_ID, S_ID
constexpr uint64_t kAdaptiveFloor = 32;  // default smallest epoch the commit-based cap allows

struct Req {
    uint64_t id;
    uint32_t type;      // LTPG_TYPE_*
    uint32_t tid;
    uint32_t attempts;
    uint64_t arrival;
    uint64_t data;      // index into the workload's request store
};

struct ByTid {
    const std::vector<Req> *reqs;
    bool operator()(uint64_t a, uint64_t b) const { return (*reqs)[a].tid > (*reqs)[b].tid; }
};

// Pinned host memory; `coherent` (fine-grained) for buffers that kernels read
// or write directly while the host polls them.
template <typename T>
int host_alloc(T **p, uint64_t count, bool coherent = false)
{
    *p = nullptr;
    const unsigned flags = coherent ? (hipHostMallocCoherent | hipHostMallocMapped) : 0u;
    if (hipHostMalloc(reinterpret_cast<void **>(p), std::max<uint64_t>(count, 1) * sizeof(T), flags) != hipSuccess) {
        *p = nullptr;
        LTPG_FAIL(LTPG_ERR_NOMEM, "pinned host allocation failed");
    }
    return LTPG_OK;
}

template <typename T>
void host_free(T *&p)
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

// One epoch in flight (or a free slot).
struct Flight {
    bool busy = false;
    uint64_t epoch = 0;
    std::vector<uint64_t> members;  // request indices, batch order (NewOrders, then Payments)
    uint32_t n_no = 0, n_pay = 0, n_os = 0, n_sl = 0;  // order-entry members by type, in this order
    hipEvent_t copied = nullptr;     // order-entry staging copied (copy stream)
    uint64_t *h_seq = nullptr;       // pinned: sequence number of the last published epoch
    uint64_t seq = 0;                // sequence number of the epoch in flight
    uint64_t launched_ns = 0;
    // key-value staging (pinned) and device slot, one block each: an epoch
    // of B transactions packs its keys (8 S B bytes), metas (4 S B) and TIDs
    // (4 B) back to back, request-major, and goes over in one copy
    uint8_t *h_stage = nullptr, *d_stage = nullptr;
    // order-entry staging (pinned, packed query blocks) and device slot
    uint32_t *h_no = nullptr, *d_no = nullptr, *h_pay = nullptr, *d_pay = nullptr;
    uint32_t *h_os = nullptr, *d_os = nullptr, *h_sl = nullptr, *d_sl = nullptr;  // read-only, value-major + TID
    // results (pinned)
    uint64_t *h_bits = nullptr, *h_bits2 = nullptr, *h_bits3 = nullptr, *h_bits4 = nullptr;
    uint32_t *h_reads = nullptr, *h_reads2 = nullptr, *h_reads3 = nullptr, *h_reads4 = nullptr;
};

struct Lane {                        // one workload
    ltpg_ctx *ctx = nullptr;
    hipStream_t copy = nullptr;
    uint32_t cap = 0;                // transactions per epoch (budget)
    uint64_t last_commits = ~0ull;   // commits of the last completed epoch (none yet: ~0)
    std::deque<uint64_t> fresh;      // arrival order
    std::priority_queue<uint64_t, std::vector<uint64_t>, ByTid> requeue;
    std::vector<Flight> flights;
    uint64_t epochs = 0;
    uint32_t next_tid = 0;
    uint64_t max_epoch = 0;
    std::vector<std::vector<uint64_t>> log;  // record_epochs: request ids per epoch
    explicit Lane(const std::vector<Req> *r) : requeue(ByTid{r}) {}
};

}  // namespace

struct ltpg_server {
    ltpg_serve_config cfg;
    std::vector<Req> reqs;
    std::vector<uint64_t> kv_keys;   // kv_ops per request
    std::vector<uint32_t> kv_meta;
    std::vector<uint32_t> no_data;   // kNoFields per NewOrder
    std::vector<uint32_t> pay_data;  // kPayFields per Payment
    std::vector<uint32_t> os_data;   // kOsFields per Order-Status
    std::vector<uint32_t> sl_data;   // kSlFields per Stock-Level
    Lane lane[2];                    // [0] key-value, [1] order-entry
    std::mutex mu;
    std::atomic<int> waiting{0};     // API calls waiting for mu (the loop yields to them)
    std::condition_variable cv;
    std::deque<ltpg_serve_done> done;
    std::unordered_map<uint64_t, std::vector<uint32_t>> reads;
    uint64_t committed[kCountedTypes] = {}, attempts[kCountedTypes] = {};
    uint64_t outstanding = 0;        // submitted, not yet committed
    uint64_t dropped = 0;            // dropped by ltpg_serve_cancel
    bool stop = false;
    bool cancel = false;
    int error = LTPG_OK;
    std::string error_msg;
    std::thread th;
    ltpg_server() : lane{Lane(&reqs), Lane(&reqs)} {}
};

namespace {

// The serving thread holds mu while it forms, launches and completes
// epochs, and re-takes it at once; a plain lock from an API call could wait
// behind thousands of epochs.  API calls announce themselves in `waiting`
// and the loop steps aside until they have the lock.
struct ClientLock {
    std::unique_lock<std::mutex> lk;
    explicit ClientLock(ltpg_server *sv) : lk(lock_for(sv)) {}
    static std::unique_lock<std::mutex> lock_for(ltpg_server *sv)
    {
        sv->waiting.fetch_add(1);
        std::unique_lock<std::mutex> l(sv->mu);
        sv->waiting.fetch_sub(1);
        return l;
    }
};

void free_flight(Flight &f)
{
    if (f.copied) (void)hipEventDestroy(f.copied);
    host_free(f.h_seq);
    host_free(f.h_stage);
    host_free(f.h_no);
    host_free(f.h_pay);
    host_free(f.h_os);
    host_free(f.h_sl);
    host_free(f.h_bits3);
    host_free(f.h_bits4);
    host_free(f.h_reads3);
    host_free(f.h_reads4);
    dev_free(f.d_os);
    dev_free(f.d_sl);
    host_free(f.h_bits);
    host_free(f.h_bits2);
    host_free(f.h_reads);
    host_free(f.h_reads2);
    dev_free(f.d_stage);
    dev_free(f.d_no);
    dev_free(f.d_pay);
}

int alloc_flight(ltpg_server *sv, int w, Flight &f)
{
    const uint64_t C = sv->lane[w].cap;
    int rc;
    LTPG_HIP_TRY(hipEventCreateWithFlags(&f.copied, hipEventDisableTiming));
    if ((rc = host_alloc(&f.h_seq, 1, true))) return rc;
    *f.h_seq = 0;
    if (w == 0) {
        const uint64_t S = sv->cfg.kv_ops;
        const uint64_t fw = (uint64_t)sv->lane[0].ctx->kv.field_words * sv->lane[0].ctx->kv.read_slot;
        const uint64_t stage = (12 * S + 4) * C;
        if ((rc = host_alloc(&f.h_stage, stage, true)) || (rc = host_alloc(&f.h_bits, words(C), true))) return rc;
        if (sv->cfg.keep_reads && (rc = host_alloc(&f.h_reads, S * C * fw))) return rc;
        if ((rc = dev_alloc(&f.d_stage, stage, "serving batch slot"))) return rc;
    } else {
        if ((rc = host_alloc(&f.h_no, 53 * C)) || (rc = host_alloc(&f.h_pay, 10 * C)) ||
            (rc = host_alloc(&f.h_os, (kOsFields + 1) * C)) || (rc = host_alloc(&f.h_sl, (kSlFields + 1) * C)) ||
            (rc = host_alloc(&f.h_bits, words(C), true)) || (rc = host_alloc(&f.h_bits2, words(C), true)) ||
            (rc = host_alloc(&f.h_bits3, words(C), true)) || (rc = host_alloc(&f.h_bits4, words(C), true)))
            return rc;
        if (sv->cfg.keep_reads &&
            ((rc = host_alloc(&f.h_reads, (uint64_t)kNoReads * C)) || (rc = host_alloc(&f.h_reads2, (uint64_t)kPayReads * C)) ||
             (rc = host_alloc(&f.h_reads3, (uint64_t)kOsReads * C)) || (rc = host_alloc(&f.h_reads4, (uint64_t)kSlReads * C))))
            return rc;
        if ((rc = dev_alloc(&f.d_no, 53 * C, "serving NewOrder slot")) ||
            (rc = dev_alloc(&f.d_pay, 10 * C, "serving Payment slot")) ||
            (rc = dev_alloc(&f.d_os, (kOsFields + 1) * C, "serving Order-Status slot")) ||
            (rc = dev_alloc(&f.d_sl, (kSlFields + 1) * C, "serving Stock-Level slot")))
            return rc;
    }
    return LTPG_OK;
}

// Position of an order-entry type in an epoch's batch.
uint32_t type_rank(uint32_t type)
{
    return type == LTPG_TYPE_NEWORDER ? 0 : type == LTPG_TYPE_PAYMENT ? 1 : type == LTPG_TYPE_ORDERSTATUS ? 2 : 3;
}

// Pack an order-entry epoch's members into the flight's pinned blocks
// (lock held): value-major columns of stride C, TIDs in the last column.
void pack_tpcc(ltpg_server *sv, Flight &f, uint64_t C)
{
    uint32_t at = 0;
    for (uint32_t q = 0; q < f.n_no; q++, at++) {
        const Req &r = sv->reqs[f.members[at]];
        const uint32_t *src = &sv->no_data[r.data * kNoFields];
        for (uint32_t k = 0; k < kNoFields; k++) f.h_no[(uint64_t)k * C + q] = src[k];
        f.h_no[52 * C + q] = r.tid;
    }
    for (uint32_t p = 0; p < f.n_pay; p++, at++) {
        const Req &r = sv->reqs[f.members[at]];
        const uint32_t *src = &sv->pay_data[r.data * kPayFields];
        for (uint32_t k = 0; k < kPayFields; k++) f.h_pay[(uint64_t)k * C + p] = src[k];
        f.h_pay[9 * C + p] = r.tid;
    }
    for (uint32_t o = 0; o < f.n_os; o++, at++) {
        const Req &r = sv->reqs[f.members[at]];
        const uint32_t *src = &sv->os_data[r.data * kOsFields];
        for (uint32_t k = 0; k < kOsFields; k++) f.h_os[(uint64_t)k * C + o] = src[k];
        f.h_os[kOsFields * C + o] = r.tid;
    }
    for (uint32_t l = 0; l < f.n_sl; l++, at++) {
        const Req &r = sv->reqs[f.members[at]];
        const uint32_t *src = &sv->sl_data[r.data * kSlFields];
        for (uint32_t k = 0; k < kSlFields; k++) f.h_sl[(uint64_t)k * C + l] = src[k];
        f.h_sl[kSlFields * C + l] = r.tid;
    }
}

// Read-only batches over value-major columns of stride C (host or device).
void readonly_views(uint32_t *os, uint32_t *sl, uint64_t C, uint32_t n_os, uint32_t n_sl,
                    ltpg_orderstatus_batch *ob, ltpg_stocklevel_batch *sb)
{
    *ob = ltpg_orderstatus_batch{n_os, 0, os, os + C, os + 2 * C, os + 3 * C, os + 4 * C, os + kOsFields * C};
    *sb = ltpg_stocklevel_batch{n_sl, 0, sl, sl + C, sl + kSlFields * C};
}

// Take the next epoch's members (lock held): re-queued first, by TID, then
// new arrivals.
void take_members(ltpg_server *sv, int w, Flight &f)
{
    Lane &L = sv->lane[w];
    f.members.clear();
    // Under contention an epoch commits a short chain of the lowest
    // priorities whatever its size, so a large epoch mostly re-runs aborts:
    // cap it at epoch_growth times the last epoch's commits (not below
    // epoch_floor), so that epochs turn over faster.
    uint64_t cap = L.cap;
    const uint64_t min_epoch = sv->cfg.epoch_floor ? sv->cfg.epoch_floor : kAdaptiveFloor;
    if (sv->cfg.epoch_growth && L.last_commits != ~0ull)
        cap = std::min<uint64_t>(cap, std::max<uint64_t>(min_epoch, sv->cfg.epoch_growth * L.last_commits));
    while (f.members.size() < cap && !L.requeue.empty()) {
        f.members.push_back(L.requeue.top());
        L.requeue.pop();
    }
    while (f.members.size() < cap && !L.fresh.empty()) {
        f.members.push_back(L.fresh.front());
        L.fresh.pop_front();
    }
    if (w == 1) {  // NewOrders, Payments, Order-Status, Stock-Level, each in member order
        std::stable_sort(f.members.begin(), f.members.end(),
                         [&](uint64_t a, uint64_t b) { return type_rank(sv->reqs[a].type) < type_rank(sv->reqs[b].type); });
        f.n_no = f.n_pay = f.n_os = f.n_sl = 0;
        for (uint64_t r : f.members) {
            const uint32_t ty = sv->reqs[r].type;
            f.n_no += ty == LTPG_TYPE_NEWORDER;
            f.n_pay += ty == LTPG_TYPE_PAYMENT;
            f.n_os += ty == LTPG_TYPE_ORDERSTATUS;
            f.n_sl += ty == LTPG_TYPE_STOCKLEVEL;
        }
    }
}

// Pack, copy and launch one epoch (lock held: the request store may grow
// under concurrent submissions).
int launch(ltpg_server *sv, int w, Flight &f)
{
    Lane &L = sv->lane[w];
    ltpg_ctx *ctx = L.ctx;
    const uint64_t C = L.cap;
    int rc;
    // a flight is restaged only after its last epoch was published, so
    // nothing on the device still reads its buffers
    if (w == 0) {
        // requests are stored request-major and the epoch runs on that
        // layout: one contiguous copy per request into the staging, which a
        // kernel on the epoch's stream reads over the bus; no transpose
        const uint32_t S = sv->cfg.kv_ops, B = (uint32_t)f.members.size();
        const uint64_t keys_at = 0, meta_at = 8ull * S * B, tids_at = 12ull * S * B;
        uint64_t *h_keys = reinterpret_cast<uint64_t *>(f.h_stage + keys_at);
        uint32_t *h_meta = reinterpret_cast<uint32_t *>(f.h_stage + meta_at);
        uint32_t *h_tids = reinterpret_cast<uint32_t *>(f.h_stage + tids_at);
        for (uint32_t t = 0; t < B; t++) {
            const Req &r = sv->reqs[f.members[t]];
            h_tids[t] = r.tid;
            std::memcpy(h_keys + (uint64_t)t * S, &sv->kv_keys[r.data * S], S * sizeof(uint64_t));
            std::memcpy(h_meta + (uint64_t)t * S, &sv->kv_meta[r.data * S], S * sizeof(uint32_t));
        }
        KvBatchDev shape;
        shape.n_txns = B;
        shape.n_slots = S;
        const bool small = !(ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) && kv_small_epoch_ok(ctx->kv, shape);
        bool published = false;
        if (small) {
            // one single-workgroup kernel reads the staging in place and,
            // unless read values follow, publishes the epoch itself
            ltpg_kv_batch b = {B, S, h_keys, h_meta, h_tids};
            if ((rc = ltpg_kv_submit(ctx, &b, LTPG_MEM_DEVICE | LTPG_BATCH_REQUEST_MAJOR))) return rc;
            published = !sv->cfg.keep_reads;
            const uint64_t seq = published ? ++f.seq : 0;
            LTPG_HIP_TRY(launch_kv_small_epoch(ctx->kv, ctx->kv.cur, ctx->cfg.values, ctx->counters,
                                               published ? f.h_bits : nullptr, published ? f.h_seq : nullptr, seq,
                                               ctx->stream));
        } else {
            LTPG_HIP_TRY(launch_copy_u32(reinterpret_cast<uint32_t *>(f.d_stage),
                                         reinterpret_cast<const uint32_t *>(f.h_stage), (tids_at + 4ull * B) / 4,
                                         ctx->stream));
            ltpg_kv_batch b = {B, S, reinterpret_cast<const uint64_t *>(f.d_stage + keys_at),
                               reinterpret_cast<const uint32_t *>(f.d_stage + meta_at),
                               reinterpret_cast<const uint32_t *>(f.d_stage + tids_at)};
            if ((rc = ltpg_kv_submit(ctx, &b, LTPG_MEM_DEVICE | LTPG_BATCH_REQUEST_MAJOR))) return rc;
            if ((rc = ltpg_epoch_run(ctx, LTPG_PHASE_ALL))) return rc;
        }
        if (sv->cfg.keep_reads)
            LTPG_HIP_TRY(hipMemcpyAsync(f.h_reads, ctx->kv.read_values,
                                        (uint64_t)S * B * ctx->kv.read_slot * ctx->kv.field_words * 4,
                                        hipMemcpyDeviceToHost, ctx->stream));
        if (!published) {
            PublishSet ps{{ctx->kv.bitmap, nullptr, nullptr, nullptr}, {f.h_bits, nullptr, nullptr, nullptr},
                          {(uint32_t)words(B), 0, 0, 0}};
            LTPG_HIP_TRY(launch_publish(ps, f.h_seq, ++f.seq, ctx->stream));
        }
    } else {
        const uint32_t n_no = f.n_no, n_pay = f.n_pay, n_os = f.n_os, n_sl = f.n_sl;
        pack_tpcc(sv, f, C);
        if (n_no)
            LTPG_HIP_TRY(hipMemcpy2DAsync(f.d_no, C * 4, f.h_no, C * 4, (uint64_t)n_no * 4, 53,
                                          hipMemcpyHostToDevice, L.copy));
        if (n_pay)
            LTPG_HIP_TRY(hipMemcpy2DAsync(f.d_pay, C * 4, f.h_pay, C * 4, (uint64_t)n_pay * 4, 10,
                                          hipMemcpyHostToDevice, L.copy));
        if (n_os)
            LTPG_HIP_TRY(hipMemcpy2DAsync(f.d_os, C * 4, f.h_os, C * 4, (uint64_t)n_os * 4, kOsFields + 1,
                                          hipMemcpyHostToDevice, L.copy));
        if (n_sl)
            LTPG_HIP_TRY(hipMemcpy2DAsync(f.d_sl, C * 4, f.h_sl, C * 4, (uint64_t)n_sl * 4, kSlFields + 1,
                                          hipMemcpyHostToDevice, L.copy));
        LTPG_HIP_TRY(hipEventRecord(f.copied, L.copy));
        LTPG_HIP_TRY(hipStreamWaitEvent(ctx->stream, f.copied, 0));
        ltpg_neworder_batch nb;
        ltpg_payment_batch pb;
        ltpg_neworder_view(f.d_no, (uint32_t)C, n_no, &nb);
        ltpg_payment_view(f.d_pay, (uint32_t)C, n_pay, &pb);
        if ((rc = ltpg_tpcc_submit(ctx, &nb, &pb, LTPG_MEM_DEVICE))) return rc;
        if (n_os || n_sl) {
            ltpg_orderstatus_batch ob;
            ltpg_stocklevel_batch sb;
            readonly_views(f.d_os, f.d_sl, C, n_os, n_sl, &ob, &sb);
            if ((rc = ltpg_tpcc_submit_readonly(ctx, &ob, &sb, LTPG_MEM_DEVICE))) return rc;
        }
        if ((rc = ltpg_epoch_run(ctx, LTPG_PHASE_ALL))) return rc;
        if (sv->cfg.keep_reads) {
            const std::pair<uint32_t *, const uint32_t *> reads[4] = {
                {f.h_reads, ctx->tpcc.d.reads_no}, {f.h_reads2, ctx->tpcc.d.reads_pay},
                {f.h_reads3, ctx->tpcc.d.reads_os}, {f.h_reads4, ctx->tpcc.d.reads_sl}};
            const uint64_t n_words[4] = {(uint64_t)kNoReads * n_no, (uint64_t)kPayReads * n_pay,
                                         (uint64_t)kOsReads * n_os, (uint64_t)kSlReads * n_sl};
            for (int k = 0; k < 4; k++)
                if (n_words[k])
                    LTPG_HIP_TRY(hipMemcpyAsync(reads[k].first, reads[k].second, n_words[k] * 4,
                                                hipMemcpyDeviceToHost, ctx->stream));
        }
        PublishSet ps{{ctx->tpcc.d.bitmap_no, ctx->tpcc.d.bitmap_pay, ctx->tpcc.d.bitmap_os, ctx->tpcc.d.bitmap_sl},
                      {f.h_bits, f.h_bits2, f.h_bits3, f.h_bits4},
                      {(uint32_t)words(n_no), (uint32_t)words(n_pay), (uint32_t)words(n_os), (uint32_t)words(n_sl)}};
        LTPG_HIP_TRY(launch_publish(ps, f.h_seq, ++f.seq, ctx->stream));
    }
    f.launched_ns = now_ns();
    f.busy = true;
    return LTPG_OK;
}

// The plain per-batch form of one epoch, for comparison (ltpg_serve_config.
// plain): the batch is packed on the host in the reference's layout
// (slot-major key-value ops, value-major order-entry blocks), copied by the
// submit, run phase by phase with a synchronisation after each, and its
// bitmaps (and reads) copied back, one step after the other.  Packing happens under the lock (the request store
// may grow); the device work runs without it.  On return the flight holds
// a published epoch, as the served path leaves it.
int launch_plain(ltpg_server *sv, int w, Flight &f, std::unique_lock<std::mutex> &lk)
{
    Lane &L = sv->lane[w];
    ltpg_ctx *ctx = L.ctx;
    const uint64_t C = L.cap;
    const uint32_t B = (uint32_t)f.members.size();
    int rc = LTPG_OK;
    if (w == 0) {
        const uint32_t S = sv->cfg.kv_ops;
        uint64_t *h_keys = reinterpret_cast<uint64_t *>(f.h_stage);
        uint32_t *h_meta = reinterpret_cast<uint32_t *>(f.h_stage + 8ull * S * B);
        uint32_t *h_tids = reinterpret_cast<uint32_t *>(f.h_stage + 12ull * S * B);
        for (uint32_t t = 0; t < B; t++) {
            const Req &r = sv->reqs[f.members[t]];
            h_tids[t] = r.tid;
            for (uint32_t s = 0; s < S; s++) {
                h_keys[(uint64_t)s * B + t] = sv->kv_keys[r.data * S + s];
                h_meta[(uint64_t)s * B + t] = sv->kv_meta[r.data * S + s];
            }
        }
        lk.unlock();
        const ltpg_kv_batch b = {B, S, h_keys, h_meta, h_tids};
        rc = ltpg_kv_submit(ctx, &b, LTPG_MEM_HOST);
        if (!rc) rc = run_epoch_stepwise(ctx);
        if (!rc) rc = ltpg_commit_bitmap(ctx, LTPG_TYPE_KV, f.h_bits, words(B));
        if (!rc && sv->cfg.keep_reads)
            rc = ltpg_kv_read_values(ctx, f.h_reads, (uint64_t)S * B * ctx->kv.read_slot * ctx->kv.field_words);
    } else {
        const uint32_t n_no = f.n_no, n_pay = f.n_pay, n_os = f.n_os, n_sl = f.n_sl;
        pack_tpcc(sv, f, C);
        lk.unlock();
        ltpg_neworder_batch nb;
        ltpg_payment_batch pb;
        ltpg_neworder_view(f.h_no, (uint32_t)C, n_no, &nb);
        ltpg_payment_view(f.h_pay, (uint32_t)C, n_pay, &pb);
        rc = ltpg_tpcc_submit(ctx, &nb, &pb, LTPG_MEM_HOST);
        if (!rc && (n_os || n_sl)) {
            ltpg_orderstatus_batch ob;
            ltpg_stocklevel_batch sb;
            readonly_views(f.h_os, f.h_sl, C, n_os, n_sl, &ob, &sb);
            rc = ltpg_tpcc_submit_readonly(ctx, &ob, &sb, LTPG_MEM_HOST);
        }
        if (!rc) rc = run_epoch_stepwise(ctx);
        const uint32_t types[4] = {LTPG_TYPE_NEWORDER, LTPG_TYPE_PAYMENT, LTPG_TYPE_ORDERSTATUS, LTPG_TYPE_STOCKLEVEL};
        const uint32_t counts[4] = {n_no, n_pay, n_os, n_sl};
        const uint32_t per[4] = {kNoReads, kPayReads, kOsReads, kSlReads};
        uint64_t *bits[4] = {f.h_bits, f.h_bits2, f.h_bits3, f.h_bits4};
        uint32_t *reads[4] = {f.h_reads, f.h_reads2, f.h_reads3, f.h_reads4};
        for (int k = 0; k < 4 && !rc; k++) {
            if (!counts[k]) continue;
            rc = ltpg_commit_bitmap(ctx, types[k], bits[k], words(counts[k]));
            if (!rc && sv->cfg.keep_reads)
                rc = ltpg_tpcc_read_values(ctx, types[k], reads[k], (uint64_t)per[k] * counts[k]);
        }
    }
    lk.lock();
    if (rc) return rc;
    *f.h_seq = ++f.seq;
    f.launched_ns = now_ns();
    f.busy = true;
    return LTPG_OK;
}

bool bit(const uint64_t *b, uint32_t i) { return (b[i / 64] >> (i % 64)) & 1ull; }

// Read a finished epoch's results (lock held).
void complete(ltpg_server *sv, int w, Flight &f)
{
    Lane &L = sv->lane[w];
    const uint64_t t_now = now_ns();
    const uint32_t B = (uint32_t)f.members.size();
    const uint32_t S = sv->cfg.kv_ops;
    uint64_t commits = 0;
    for (uint32_t t = 0; t < B; t++) {
        const uint64_t ri = f.members[t];
        Req &r = sv->reqs[ri];
        r.attempts++;
        sv->attempts[r.type]++;
        bool ok;
        // order-entry members: NewOrders, Payments, Order-Status, Stock-Level
        const uint32_t c1 = f.n_no, c2 = c1 + f.n_pay, c3 = c2 + f.n_os;
        if (w == 0) ok = bit(f.h_bits, t);
        else if (t < c1) ok = bit(f.h_bits, t);
        else if (t < c2) ok = bit(f.h_bits2, t - c1);
        else if (t < c3) ok = bit(f.h_bits3, t - c2);
        else ok = bit(f.h_bits4, t - c3);
        if (!ok) {
            L.requeue.push(ri);
            continue;
        }
        sv->committed[r.type]++;
        sv->outstanding--;
        commits++;
        ltpg_serve_done d;
        d.id = r.id;
        d.type = r.type;
        d.tid = r.tid;
        d.attempts = r.attempts;
        d.reserved = 0;
        d.epoch = f.epoch;
        d.arrival_ns = r.arrival;
        d.commit_ns = t_now;
        sv->done.push_back(d);
        if (sv->cfg.keep_reads) {
            std::vector<uint32_t> v;
            if (w == 0) {  // each read's field, each scan's rows inside the table, in slot order
                const KvState &kv = L.ctx->kv;
                const uint32_t fw = kv.field_words;
                for (uint32_t s = 0; s < S; s++) {
                    const uint32_t m = sv->kv_meta[r.data * S + s];
                    const uint64_t row = sv->kv_keys[r.data * S + s];
                    const uint32_t kind = LTPG_KV_KIND(m);
                    if (row >= kv.rows || LTPG_KV_FIELD(m) >= kv.fields) continue;
                    uint64_t n = 0;
                    if (kind == LTPG_KV_READ)
                        n = 1;
                    else if (kind == LTPG_KV_SCAN && LTPG_KV_SCAN_LEN(m) <= kv.max_scan)
                        n = std::min<uint64_t>(LTPG_KV_SCAN_LEN(m), kv.rows - row);
                    const uint32_t *src = f.h_reads + (((uint64_t)s * B) + t) * kv.read_slot * fw;
                    v.insert(v.end(), src, src + n * fw);
                }
            } else if (t < c1) {
                for (uint32_t k = 0; k < kNoReads; k++) v.push_back(f.h_reads[(uint64_t)k * f.n_no + t]);
            } else if (t < c2) {
                const uint32_t p = t - c1;
                for (uint32_t k = 0; k < kPayReads; k++) v.push_back(f.h_reads2[(uint64_t)k * f.n_pay + p]);
            } else if (t < c3) {
                const uint32_t o = t - c2;
                for (uint32_t k = 0; k < kOsReads; k++) v.push_back(f.h_reads3[(uint64_t)k * f.n_os + o]);
            } else {
                const uint32_t l = t - c3;
                for (uint32_t k = 0; k < kSlReads; k++) v.push_back(f.h_reads4[(uint64_t)k * f.n_sl + l]);
            }
            sv->reads[r.id] = std::move(v);
        }
    }
    L.last_commits = commits;
    f.busy = false;
}

// hipErrorNotReady unless the stream has failed, or has been idle for
// 100 ms since the launch without the epoch's results becoming visible
// (reported as an error rather than waited for forever).
hipError_t stream_error(hipStream_t s, uint64_t launched_ns)
{
    const hipError_t q = hipStreamQuery(s);
    if (q != hipSuccess) return q;
    return now_ns() - launched_ns > 100000000ull ? hipErrorUnknown : hipErrorNotReady;
}

// The flight's epoch has been published (its results are in host memory).
bool published(const Flight &f) { return __atomic_load_n(f.h_seq, __ATOMIC_ACQUIRE) == f.seq; }

bool any_flight_done(ltpg_server *sv)
{
    for (auto &L : sv->lane) {
        if (!L.ctx) continue;
        for (auto &f : L.flights)
            if (f.busy && published(f)) return true;
    }
    return false;
}

void serve_loop(ltpg_server *sv)
{
    (void)hipSetDevice((sv->lane[0].ctx ? sv->lane[0].ctx : sv->lane[1].ctx)->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    while (true) {
        bool progress = false, inflight = false;
        if (sv->cancel) {
            for (auto &L : sv->lane) {
                sv->dropped += L.fresh.size() + L.requeue.size();
                sv->outstanding -= L.fresh.size() + L.requeue.size();
                L.fresh.clear();
                while (!L.requeue.empty()) L.requeue.pop();
            }
        }
        for (int w = 0; w < 2; w++) {
            Lane &L = sv->lane[w];
            if (!L.ctx) continue;
            // results first: re-queued aborts join the next epoch formed
            for (auto &f : L.flights) {
                if (!f.busy) continue;
                // a device error never publishes: after 2 ms ask the stream
                const hipError_t q = published(f)                              ? hipSuccess
                                     : now_ns() - f.launched_ns < 2000000ull ? hipErrorNotReady
                                                                             : stream_error(L.ctx->stream, f.launched_ns);
                if (q == hipSuccess) {
                    complete(sv, w, f);
                    if (sv->cancel) {  // aborts of a cancelled run are dropped too
                        sv->dropped += L.requeue.size();
                        sv->outstanding -= L.requeue.size();
                        while (!L.requeue.empty()) L.requeue.pop();
                    }
                    progress = true;
                } else if (q != hipErrorNotReady) {
                    sv->error = LTPG_ERR_DEVICE;
                    sv->error_msg = std::string("serving epoch: ") + hipGetErrorString(q);
                    f.busy = false;
                }
            }
            // start epochs on free slots, in slot order so epochs stay ordered
            for (auto &f : L.flights) {
                if (f.busy || (L.fresh.empty() && L.requeue.empty()) || sv->error) continue;
                take_members(sv, w, f);
                f.epoch = L.epochs++;
                L.max_epoch = std::max<uint64_t>(L.max_epoch, f.members.size());
                if (sv->cfg.record_epochs) {
                    std::vector<uint64_t> ids;
                    for (uint64_t r : f.members) ids.push_back(sv->reqs[r].id);
                    L.log.push_back(std::move(ids));
                }
                const int rc = sv->cfg.plain ? launch_plain(sv, w, f, lk) : launch(sv, w, f);
                if (rc) {
                    sv->error = rc;
                    sv->error_msg = ltpg_last_error();
                }
                progress = true;
            }
            for (auto &f : L.flights) inflight |= f.busy;
        }
        if (progress) sv->cv.notify_all();
        const bool pending = !sv->lane[0].fresh.empty() || !sv->lane[0].requeue.empty() ||
                             !sv->lane[1].fresh.empty() || !sv->lane[1].requeue.empty();
        if (sv->stop && !inflight && (!pending || sv->error)) break;
        if (!progress) {
            lk.unlock();
            if (inflight) {
                // Poll the epochs in flight rather than sleep: a sleep lasts
                // at least a scheduler tick (~50 us), as long as a whole
                // small epoch on the device.  Only this thread touches the
                // flights, so their events are read without the lock; a
                // client asking for the lock ends the poll.
                const uint64_t until = now_ns() + 200000;
                while (sv->waiting.load() == 0 && !any_flight_done(sv) && now_ns() < until)
                    std::this_thread::yield();
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(50));
            }
            lk.lock();
        } else if (sv->waiting.load() > 0) {
            lk.unlock();
            while (sv->waiting.load() > 0) std::this_thread::yield();
            lk.lock();
        }
    }
    sv->cv.notify_all();
}

}  // namespace

extern "C" {

void ltpg_serve_config_init(ltpg_serve_config *c)
{
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->size = sizeof(*c);
    c->slots = 2;
    c->kv_ops = 16;
    c->epoch_growth = 8;
    c->epoch_floor = (uint32_t)kAdaptiveFloor;
}

int ltpg_serve_cancel(ltpg_server *sv, uint64_t *dropped)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    {
        ClientLock g(sv);
        sv->cancel = true;  // the loop drops what is pending, finishes its
        sv->stop = true;    // epochs in flight and returns
    }
    if (sv->th.joinable()) sv->th.join();
    if (dropped) *dropped = sv->dropped;
    return ltpg_serve_stop(sv);
}

int ltpg_serve_stop(ltpg_server *sv)
{
    if (!sv) return LTPG_OK;
    {
        ClientLock g(sv);
        sv->stop = true;
    }
    if (sv->th.joinable()) sv->th.join();
    const int err = sv->error;
    const std::string msg = sv->error_msg;
    for (int w = 0; w < 2; w++) {
        Lane &L = sv->lane[w];
        if (!L.ctx) continue;
        (void)hipSetDevice(L.ctx->device);
        (void)hipStreamSynchronize(L.ctx->stream);
        if (L.copy) (void)hipStreamSynchronize(L.copy);
        for (auto &f : L.flights) free_flight(f);
        if (L.copy) (void)hipStreamDestroy(L.copy);
    }
    delete sv;
    if (err) LTPG_FAIL(err, msg);
    return LTPG_OK;
}

int ltpg_serve_start(ltpg_ctx *kv_ctx, ltpg_ctx *tpcc_ctx, const ltpg_serve_config *cfg, ltpg_server **out)
{
    if (!out || !cfg) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    *out = nullptr;
    if (cfg->size != sizeof(ltpg_serve_config)) LTPG_FAIL(LTPG_ERR_INVALID, "ltpg_serve_config.size mismatch");
    if (!kv_ctx && !tpcc_ctx) LTPG_FAIL(LTPG_ERR_INVALID, "no context");
    if (kv_ctx && !kv_ctx->kv.table) LTPG_FAIL(LTPG_ERR_INVALID, "kv_ctx is not a key-value context");
    if (tpcc_ctx && !tpcc_ctx->tpcc.d.regs) LTPG_FAIL(LTPG_ERR_INVALID, "tpcc_ctx is not an order-entry context");
    if (cfg->slots < 1 || cfg->slots > 4) LTPG_FAIL(LTPG_ERR_INVALID, "slots must be in [1, 4]");
    if (kv_ctx && (cfg->kv_ops == 0 || cfg->kv_ops > kv_ctx->kv.max_slots))
        LTPG_FAIL(LTPG_ERR_INVALID, "kv_ops must be in [1, kv_max_slots]");
    ltpg_server *sv = new (std::nothrow) ltpg_server();
    if (!sv) LTPG_FAIL(LTPG_ERR_NOMEM, "host allocation of server failed");
    sv->cfg = *cfg;
    sv->lane[0].ctx = kv_ctx;
    sv->lane[1].ctx = tpcc_ctx;
    int rc = LTPG_OK;
    for (int w = 0; w < 2 && !rc; w++) {
        Lane &L = sv->lane[w];
        if (!L.ctx) continue;
        if ((rc = check_ctx(L.ctx))) break;
        // per-transaction bytes of the access-set buffers an epoch holds:
        // its batch slot and its read values
        const uint64_t per = w == 0 ? (uint64_t)cfg->kv_ops * (12 + 4ull * L.ctx->kv.field_words * L.ctx->kv.read_slot) + 4
                                    : 4ull * (53 + kNoReads);
        uint64_t cap = L.ctx->cfg.max_txns;
        if (cfg->max_epoch_txns) cap = std::min<uint64_t>(cap, cfg->max_epoch_txns);
        if (cfg->hbm_budget) cap = std::min<uint64_t>(cap, cfg->hbm_budget / per);
        if (cap == 0) {
            set_error("the HBM budget admits no transaction");
            rc = LTPG_ERR_BUDGET;
            break;
        }
        L.cap = (uint32_t)cap;
        if (hipStreamCreateWithFlags(&L.copy, hipStreamNonBlocking) != hipSuccess) {
            set_error("hipStreamCreate failed");
            rc = LTPG_ERR_DEVICE;
            break;
        }
        L.flights.resize(cfg->plain ? 1 : cfg->slots);  // the plain form: one epoch at a time
        for (auto &f : L.flights)
            if (!rc) rc = alloc_flight(sv, w, f);
    }
    if (rc) {
        const std::string msg = ltpg_last_error();
        ltpg_serve_stop(sv);
        set_error(msg);
        return rc;
    }
    sv->th = std::thread(serve_loop, sv);
    *out = sv;
    return LTPG_OK;
}

int ltpg_serve_submit_kv(ltpg_server *sv, uint32_t n, const uint64_t *keys, const uint32_t *meta, uint64_t *first_id)
{
    if (!sv || (n && (!keys || !meta))) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    if (!sv->lane[0].ctx) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server has no key-value context");
    const uint64_t t = now_ns();
    const uint32_t S = sv->cfg.kv_ops;
    ClientLock g(sv);
    if (first_id) *first_id = sv->reqs.size();
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t data = sv->kv_keys.size() / S;
        sv->kv_keys.insert(sv->kv_keys.end(), keys + (uint64_t)i * S, keys + (uint64_t)(i + 1) * S);
        sv->kv_meta.insert(sv->kv_meta.end(), meta + (uint64_t)i * S, meta + (uint64_t)(i + 1) * S);
        sv->lane[0].fresh.push_back(sv->reqs.size());
        sv->reqs.push_back(Req{sv->reqs.size(), LTPG_TYPE_KV, sv->lane[0].next_tid++, 0, t, data});
    }
    sv->outstanding += n;
    return LTPG_OK;
}

int ltpg_serve_submit_tpcc(ltpg_server *sv, const ltpg_neworder_batch *no, const ltpg_payment_batch *pay,
                           uint64_t *first_id)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    if (!sv->lane[1].ctx) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server has no order-entry context");
    const uint32_t n_no = no ? no->n : 0, n_pay = pay ? pay->n : 0;
    if (n_no && (no->line_stride < n_no || !no->W_ID || !no->D_ID || !no->C_ID || !no->O_ID || !no->N_O_ID ||
                 !no->O_OL_CNT || !no->O_OL_ID || !no->OL_I_ID || !no->OL_SUPPLY_W_ID || !no->OL_QUANTITY))
        LTPG_FAIL(LTPG_ERR_INVALID, "bad NewOrder arrays");
    if (n_pay && (!pay->W_ID || !pay->D_ID || !pay->C_ID || !pay->C_LAST || !pay->isName || !pay->C_D_ID ||
                  !pay->C_W_ID || !pay->H_AMOUNT || !pay->H_ID))
        LTPG_FAIL(LTPG_ERR_INVALID, "bad Payment arrays");
    const uint64_t t = now_ns();
    ClientLock g(sv);
    if (first_id) *first_id = sv->reqs.size();
    Lane &L = sv->lane[1];
    for (uint32_t q = 0; q < n_no; q++) {
        const uint64_t data = sv->no_data.size() / kNoFields;
        const uint32_t *heads[7] = {no->W_ID, no->D_ID, no->C_ID, no->O_ID, no->N_O_ID, no->O_OL_CNT, no->O_OL_ID};
        for (auto h : heads) sv->no_data.push_back(h[q]);
        for (const uint32_t *a : {no->OL_I_ID, no->OL_SUPPLY_W_ID, no->OL_QUANTITY})
            for (uint32_t l = 0; l < LTPG_TPCC_MAX_LINES; l++) sv->no_data.push_back(a[(uint64_t)l * no->line_stride + q]);
        L.fresh.push_back(sv->reqs.size());
        sv->reqs.push_back(Req{sv->reqs.size(), LTPG_TYPE_NEWORDER, L.next_tid++, 0, t, data});
    }
    for (uint32_t p = 0; p < n_pay; p++) {
        const uint64_t data = sv->pay_data.size() / kPayFields;
        const uint32_t *cols[9] = {pay->W_ID, pay->D_ID, pay->C_ID, pay->C_LAST, pay->isName,
                                   pay->C_D_ID, pay->C_W_ID, pay->H_AMOUNT, pay->H_ID};
        for (auto c : cols) sv->pay_data.push_back(c[p]);
        L.fresh.push_back(sv->reqs.size());
        sv->reqs.push_back(Req{sv->reqs.size(), LTPG_TYPE_PAYMENT, L.next_tid++, 0, t, data});
    }
    sv->outstanding += n_no + n_pay;
    return LTPG_OK;
}

int ltpg_serve_submit_readonly(ltpg_server *sv, const ltpg_orderstatus_batch *os, const ltpg_stocklevel_batch *sl,
                               uint64_t *first_id)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    if (!sv->lane[1].ctx) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server has no order-entry context");
    if (sv->lane[1].ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS)
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "read-only transactions need the fast check (no conflict counts)");
    const uint32_t n_os = os ? os->n : 0, n_sl = sl ? sl->n : 0;
    if (n_os && (!os->W_ID || !os->D_ID || !os->C_ID || !os->O_ID || !os->OL_ID))
        LTPG_FAIL(LTPG_ERR_INVALID, "bad Order-Status arrays");
    if (n_sl && (!sl->W_ID || !sl->S_ID)) LTPG_FAIL(LTPG_ERR_INVALID, "bad Stock-Level arrays");
    const uint64_t t = now_ns();
    ClientLock g(sv);
    if (first_id) *first_id = sv->reqs.size();
    Lane &L = sv->lane[1];
    for (uint32_t o = 0; o < n_os; o++) {
        const uint64_t data = sv->os_data.size() / kOsFields;
        for (const uint32_t *c : {os->W_ID, os->D_ID, os->C_ID, os->O_ID, os->OL_ID}) sv->os_data.push_back(c[o]);
        L.fresh.push_back(sv->reqs.size());
        sv->reqs.push_back(Req{sv->reqs.size(), LTPG_TYPE_ORDERSTATUS, L.next_tid++, 0, t, data});
    }
    for (uint32_t l = 0; l < n_sl; l++) {
        const uint64_t data = sv->sl_data.size() / kSlFields;
        for (const uint32_t *c : {sl->W_ID, sl->S_ID}) sv->sl_data.push_back(c[l]);
        L.fresh.push_back(sv->reqs.size());
        sv->reqs.push_back(Req{sv->reqs.size(), LTPG_TYPE_STOCKLEVEL, L.next_tid++, 0, t, data});
    }
    sv->outstanding += n_os + n_sl;
    return LTPG_OK;
}

int ltpg_serve_poll(ltpg_server *sv, ltpg_serve_done *out, uint32_t max, uint32_t *n_out)
{
    if (!sv || (!out && max)) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    ClientLock g(sv);
    uint32_t n = 0;
    while (n < max && !sv->done.empty()) {
        out[n++] = sv->done.front();
        sv->done.pop_front();
    }
    if (n_out) *n_out = n;
    if (sv->error) LTPG_FAIL(sv->error, sv->error_msg);
    return LTPG_OK;
}

int ltpg_serve_read_values(ltpg_server *sv, uint64_t id, uint32_t *out, uint32_t max_words, uint32_t *n_words)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    ClientLock g(sv);
    if (!sv->cfg.keep_reads) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server does not keep read values");
    auto it = sv->reads.find(id);
    if (it == sv->reads.end()) LTPG_FAIL(LTPG_ERR_STATE, "request has not committed");
    const auto &v = it->second;
    if (n_words) *n_words = (uint32_t)v.size();
    if (out) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), max_words), out);
    return LTPG_OK;
}

int ltpg_serve_epoch(ltpg_server *sv, uint32_t workload, uint64_t e, uint64_t *ids, uint32_t max, uint32_t *n)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    if (workload != LTPG_WORKLOAD_KV && workload != LTPG_WORKLOAD_TPCC) LTPG_FAIL(LTPG_ERR_INVALID, "unknown workload");
    ClientLock g(sv);
    if (!sv->cfg.record_epochs) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server does not record epochs");
    const Lane &L = sv->lane[workload == LTPG_WORKLOAD_KV ? 0 : 1];
    if (e >= L.log.size()) LTPG_FAIL(LTPG_ERR_INVALID, "no such epoch");
    const auto &v = L.log[e];
    if (n) *n = (uint32_t)v.size();
    if (ids) std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), max), ids);
    return LTPG_OK;
}

int ltpg_serve_drain(ltpg_server *sv)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    ClientLock g(sv);
    sv->cv.wait(g.lk, [&] { return sv->outstanding == 0 || sv->error; });
    if (sv->error) LTPG_FAIL(sv->error, sv->error_msg);
    return LTPG_OK;
}

int ltpg_serve_stats_get(ltpg_server *sv, ltpg_serve_stats *st)
{
    if (!sv || !st) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    ClientLock g(sv);
    for (uint32_t k = 0; k < kCountedTypes; k++) {
        st->committed[k] = sv->committed[k];
        st->attempts[k] = sv->attempts[k];
    }
    for (int w = 0; w < 2; w++) {
        st->epochs[w] = sv->lane[w].epochs;
        st->max_epoch_txns[w] = sv->lane[w].max_epoch;
    }
    return LTPG_OK;
}

}  // extern "C"
