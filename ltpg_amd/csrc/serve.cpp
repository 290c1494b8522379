// serve.cpp - the served form: a request loop over resident tables.
//
// The reference feeds its epoch loop from generator threads that fill
// pre-allocated query batches, one epoch at a time (LTPG/TPCC-allopt-final/
// Query.h:640-725, tp.cu:136-261).  Here requests arrive individually
// (submission = arrival) and epochs are formed continuously:
//
//   priority   each workload numbers its arrivals with a 64-bit sequence
//              number; a batch TID is its low 32 bits (the key-value RMW
//              delta and the Delivery's values use it), and an epoch never
//              spans a 2^32 boundary of the low bits, so within an epoch the
//              TIDs are in arrival order;
//   admission  pending requests of a workload -- re-queued aborts first, in
//              sequence order, then new arrivals -- fill the next epoch up to
//              the per-epoch transaction budget, derived from the HBM budget
//              of the access-set buffers (batch slot and read values per
//              transaction), and a cap that follows the last epoch's commits;
//   key-value  epochs form and run on the device (kv_loop): one persistent
//              single-workgroup kernel takes arrivals from a ring in pinned
//              host memory, keeps each epoch's aborts in LDS as the head of
//              the next epoch and publishes one record per epoch into a done
//              ring, which this thread drains; or, host-formed, as below;
//   order-     each workload has `slots` host-formed epochs in flight: while
//   entry      one runs on the context's stream, the next is packed into
//              pinned staging and queued behind it; the epoch's last kernel
//              writes its commit bitmaps and a sequence number into pinned
//              host memory, which the loop polls;
//   requeue    an aborted transaction keeps its sequence number and joins
//              the next epoch formed after its result is known, ahead of
//              every later arrival, so it commits after finitely many epochs
//              (the minimum-priority member of an epoch always commits);
//   latency    commit time (result read on the host) minus arrival time;
//   memory     a request's record is freed once it has committed and every
//              older request has too (a ring indexed by sequence number).
//
// Tables and registers stay resident in the contexts; nothing is reset
// between epochs (epoch-tagged registers).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
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

uint64_t pow2_at_least(uint64_t n)
{
    uint64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

constexpr uint32_t kNoFields = 52;  // NewOrder request: 7 head fields + 3 x 15 line fields
constexpr uint32_t kPayFields = 9;
constexpr uint32_t kOsFields = 5;    // Order-Status: W_ID, D_ID, C_ID, O_ID, OL_ID
constexpr uint32_t kSlFields = 2;    // Stock-Level: W_ID, S_ID
constexpr uint32_t kDlFields = 3;    // Delivery: O_ID, NO_W_ID, NO_O_ID
constexpr uint64_t kAdaptiveFloor = 32;  // default smallest epoch the commit-based cap allows
constexpr uint64_t kDefaultRing = 1ull << 16;
constexpr uint64_t kDoneWords = 1ull << 20;          // done ring of the device loop
constexpr uint64_t kReadsRingBytes = 256ull << 20;   // cap on its read-value ring
// The device loop returns after 2 ms idle, 20 ms with its done ring full, or
// 20 ms in all: a device-wide synchronisation elsewhere in the process (a
// hipFree, say) waits at most one slice for it.
constexpr uint64_t kLoopIdleUs = 2000, kLoopStallUs = 20000, kLoopSliceUs = 20000;

enum : uint8_t { kPending = 0, kCommitted = 1, kDropped = 2 };

struct Rec {
    uint64_t id = 0;
    uint64_t arrival = 0;
    uint32_t type = 0;
    uint32_t attempts = 0;
    uint8_t state = kPending;
};

// The requests of one workload, indexed by sequence number: a ring that
// doubles when full; records older than the oldest pending request are
// freed (retire).  Each record has `pay` payload bytes: the request's
// fields where the host needs them after submission.
class SeqStore {
  public:
    void init(uint64_t first_seq, size_t pay)
    {
        base_ = first_seq;
        pay_ = pay;
        grow(1024);
    }
    uint64_t base() const { return base_; }
    uint64_t end() const { return base_ + count_; }
    bool has(uint64_t seq) const { return seq >= base_ && seq < base_ + count_; }
    Rec &at(uint64_t seq) { return recs_[seq & mask_]; }
    uint8_t *data(uint64_t seq) { return pay_ ? &payload_[(seq & mask_) * pay_] : nullptr; }
    // the next sequence number's record (must equal end())
    Rec &push()
    {
        if (count_ == recs_.size()) grow(recs_.size() * 2);
        Rec &r = recs_[(base_ + count_) & mask_];
        r = Rec();
        ++count_;
        return r;
    }
    void retire()
    {
        while (count_ && recs_[base_ & mask_].state != kPending) {
            ++base_;
            --count_;
        }
    }

  private:
    void grow(size_t cap)
    {
        std::vector<Rec> r(cap);
        std::vector<uint8_t> p(cap * pay_);
        for (uint64_t s = base_; s < base_ + count_; ++s) {
            r[s & (cap - 1)] = recs_[s & mask_];
            if (pay_) std::memcpy(&p[(s & (cap - 1)) * pay_], &payload_[(s & mask_) * pay_], pay_);
        }
        recs_.swap(r);
        payload_.swap(p);
        mask_ = cap - 1;
    }
    std::vector<Rec> recs_;
    std::vector<uint8_t> payload_;
    uint64_t base_ = 0, count_ = 0, mask_ = 0;
    size_t pay_ = 0;
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

// One host-formed epoch in flight (or a free slot).
struct Flight {
    bool busy = false;
    uint64_t epoch = 0;
    std::vector<uint64_t> members;  // sequence numbers, batch order (order-entry: NewOrders, Payments, ...)
    uint32_t n_no = 0, n_pay = 0, n_os = 0, n_sl = 0, n_dl = 0;  // order-entry members by type, in this order
    hipEvent_t copied = nullptr;     // order-entry staging copied (copy stream)
    uint64_t *h_seq = nullptr;       // pinned: sequence number of the last published epoch
    uint64_t seq = 0;                // sequence number of the epoch in flight
    uint64_t launched_ns = 0;
    uint64_t idle_since_ns = 0;      // stream first seen idle with the epoch unpublished
    // key-value staging (pinned) and device slot, one block each: an epoch
    // of B transactions packs its keys (8 S B bytes), metas (4 S B) and TIDs
    // (4 B) back to back, request-major, and goes over in one copy
    uint8_t *h_stage = nullptr, *d_stage = nullptr;
    // order-entry staging (pinned, packed query blocks) and device slot
    uint32_t *h_no = nullptr, *d_no = nullptr, *h_pay = nullptr, *d_pay = nullptr;
    uint32_t *h_os = nullptr, *d_os = nullptr, *h_sl = nullptr, *d_sl = nullptr;  // read-only, value-major + TID
    uint32_t *h_dl = nullptr, *d_dl = nullptr;                                    // Deliveries, likewise
    // results (pinned)
    uint64_t *h_bits = nullptr, *h_bits2 = nullptr, *h_bits3 = nullptr, *h_bits4 = nullptr, *h_bits5 = nullptr;
    uint32_t *h_reads = nullptr, *h_reads2 = nullptr, *h_reads3 = nullptr, *h_reads4 = nullptr, *h_reads5 = nullptr;
};

// The device loop of the key-value lane (kv_loop).
struct Loop {
    ServeCtl *ctl = nullptr;           // pinned, fine-grained
    uint64_t *ring_keys = nullptr;     // pinned, ring * S
    uint32_t *ring_meta = nullptr;
    uint64_t ring = 0;                 // capacity in requests (power of two)
    uint64_t *done = nullptr;          // pinned done ring
    uint64_t done_words = 0;
    uint64_t done_head = 0;            // host copy of ctl->done_head
    uint32_t *reads = nullptr;         // pinned read values (keep_reads)
    uint64_t *carry_keys = nullptr;    // device
    uint32_t *carry_meta = nullptr;
    uint64_t *carry_seq = nullptr;
    ServeLoopArgs args{};
    bool launched = false;
    uint64_t launched_ns = 0;
    uint64_t launches = 0;
    uint64_t exits[5] = {};
    uint64_t slice_ns = 0;
    double ticks_per_us = 100.0;
};

struct Lane {                        // one workload
    ltpg_ctx *ctx = nullptr;
    hipStream_t copy = nullptr;
    uint32_t cap = 0;                // transactions per epoch (budget)
    uint64_t last_commits = ~0ull;   // commits of the last completed epoch (none yet: ~0)
    std::deque<uint64_t> fresh;      // arrival order
    std::priority_queue<uint64_t, std::vector<uint64_t>, std::greater<uint64_t>> requeue;  // by sequence
    std::vector<Flight> flights;
    SeqStore store;
    uint64_t next_seq = 0;           // sequence number of the next arrival
    uint64_t epochs = 0;
    uint64_t max_epoch = 0;
    std::vector<std::vector<uint64_t>> log;  // record_epochs: request ids per epoch
    Loop *loop = nullptr;            // key-value lane run by the device loop
};

}  // namespace

struct ltpg_server {
    ltpg_serve_config cfg;
    Lane lane[2];                    // [0] key-value, [1] order-entry
    std::mutex mu;
    std::atomic<int> waiting{0};     // API calls waiting for mu (the loop yields to them)
    std::condition_variable cv;
    std::deque<ltpg_serve_done> done;
    std::unordered_map<uint64_t, std::vector<uint32_t>> reads;
    uint64_t committed[kCountedTypes] = {}, attempts[kCountedTypes] = {};
    uint64_t next_id = 0;
    uint64_t outstanding = 0;        // submitted, not yet committed
    uint64_t dropped = 0;            // dropped by ltpg_serve_cancel
    bool stop = false;
    bool cancel = false;
    // test hook (environment LTPG_TEST_FAIL_EPOCH = n): the n-th epoch to
    // complete, counting both workloads, is reported as a failed stream
    // although its kernels ran normally, so the error path can be driven
    // from the host
    uint64_t fail_epoch = ~0ull;
    uint64_t completed_epochs = 0;
    bool trace = false;  // environment LTPG_SERVE_TRACE: device-loop events on stderr
    int error = LTPG_OK;
    std::string error_msg;
    std::thread th;
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

// A failure of the serving thread: recorded once, and every waiter woken
// (ltpg_serve_drain returns it instead of waiting for commits that will not
// come).
void fail(ltpg_server *sv, int code, const std::string &msg)
{
    if (!sv->error) {
        sv->error = code;
        sv->error_msg = msg;
    }
    sv->cv.notify_all();
}

uint64_t ld_acquire(const uint64_t *p) { return __atomic_load_n(p, __ATOMIC_ACQUIRE); }
void st_release(uint64_t *p, uint64_t v) { __atomic_store_n(p, v, __ATOMIC_RELEASE); }

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
    host_free(f.h_dl);
    host_free(f.h_bits5);
    host_free(f.h_reads5);
    dev_free(f.d_os);
    dev_free(f.d_sl);
    dev_free(f.d_dl);
    host_free(f.h_bits);
    host_free(f.h_bits2);
    host_free(f.h_reads);
    host_free(f.h_reads2);
    dev_free(f.d_stage);
    dev_free(f.d_no);
    dev_free(f.d_pay);
}

void free_loop(Loop *&lp)
{
    if (!lp) return;
    host_free(lp->ctl);
    host_free(lp->ring_keys);
    host_free(lp->ring_meta);
    host_free(lp->done);
    host_free(lp->reads);
    dev_free(lp->carry_keys);
    dev_free(lp->carry_meta);
    dev_free(lp->carry_seq);
    delete lp;
    lp = nullptr;
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
            (rc = host_alloc(&f.h_bits3, words(C), true)) || (rc = host_alloc(&f.h_bits4, words(C), true)) ||
            (rc = host_alloc(&f.h_dl, (kDlFields + 1) * C)) || (rc = host_alloc(&f.h_bits5, words(C), true)))
            return rc;
        if (sv->cfg.keep_reads &&
            ((rc = host_alloc(&f.h_reads, (uint64_t)kNoReads * C)) || (rc = host_alloc(&f.h_reads2, (uint64_t)kPayReads * C)) ||
             (rc = host_alloc(&f.h_reads3, (uint64_t)kOsReads * C)) || (rc = host_alloc(&f.h_reads4, (uint64_t)kSlReads * C)) ||
             (rc = host_alloc(&f.h_reads5, (uint64_t)kDlReads * C))))
            return rc;
        if ((rc = dev_alloc(&f.d_no, 53 * C, "serving NewOrder slot")) ||
            (rc = dev_alloc(&f.d_pay, 10 * C, "serving Payment slot")) ||
            (rc = dev_alloc(&f.d_os, (kOsFields + 1) * C, "serving Order-Status slot")) ||
            (rc = dev_alloc(&f.d_sl, (kSlFields + 1) * C, "serving Stock-Level slot")) ||
            (rc = dev_alloc(&f.d_dl, (kDlFields + 1) * C, "serving Delivery slot")))
            return rc;
    }
    return LTPG_OK;
}

// The device loop's buffers and launch arguments.
int alloc_loop(ltpg_server *sv)
{
    Lane &L = sv->lane[0];
    const KvState &kv = L.ctx->kv;
    const uint32_t S = sv->cfg.kv_ops;
    Loop *lp = new (std::nothrow) Loop();
    if (!lp) LTPG_FAIL(LTPG_ERR_NOMEM, "host allocation of the device loop failed");
    L.loop = lp;
    lp->ring = pow2_at_least(sv->cfg.ring_requests ? sv->cfg.ring_requests : kDefaultRing);
    const uint64_t cap_max = std::min<uint64_t>(L.cap, LTPG_SMALL_EPOCH_OPS / S);
    const uint64_t floor = std::min<uint64_t>(sv->cfg.epoch_floor ? sv->cfg.epoch_floor : kAdaptiveFloor, cap_max);
    // done ring: at least two full epochs; with kept reads, small enough
    // that the read values of every entry fit kReadsRingBytes
    const uint64_t entry_words = (uint64_t)S * kv.read_slot * kv.field_words;
    uint64_t dw = kDoneWords;
    if (sv->cfg.keep_reads)
        while (dw > 4096 && dw * entry_words * 4 > kReadsRingBytes) dw >>= 1;
    dw = std::max<uint64_t>(dw, pow2_at_least(2 * (1 + cap_max)));
    lp->done_words = dw;
    int rc;
    if ((rc = host_alloc(&lp->ctl, 1, true)) || (rc = host_alloc(&lp->ring_keys, lp->ring * S, true)) ||
        (rc = host_alloc(&lp->ring_meta, lp->ring * S, true)) || (rc = host_alloc(&lp->done, dw, true)))
        return rc;
    if (sv->cfg.keep_reads && (rc = host_alloc(&lp->reads, dw * entry_words, true))) return rc;
    if ((rc = dev_alloc(&lp->carry_keys, cap_max * S, "device-loop carry")) ||
        (rc = dev_alloc(&lp->carry_meta, cap_max * S, "device-loop carry")) ||
        (rc = dev_alloc(&lp->carry_seq, cap_max, "device-loop carry")))
        return rc;
    std::memset(lp->ctl, 0, sizeof(ServeCtl));
    lp->ctl->tail = lp->ctl->head = sv->cfg.first_seq;
    lp->ctl->cap = sv->cfg.epoch_growth ? floor : cap_max;
    int khz = 0;
    LTPG_HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, L.ctx->device));
    if (khz <= 0) LTPG_FAIL(LTPG_ERR_DEVICE, "device reports no wall-clock rate");
    const uint64_t per_us = std::max<uint64_t>((uint64_t)khz / 1000, 1);
    ServeLoopArgs &a = lp->args;
    a.table = kv.table;
    a.rows = kv.rows;
    a.fields = kv.fields;
    a.fw = kv.field_words;
    a.fsw = kv.field_stride;
    a.max_scan = kv.max_scan;
    a.L = kv.read_slot;
    a.values = L.ctx->cfg.values;
    a.hot = kv.hot_row;
    a.S = S;
    a.ring_keys = lp->ring_keys;
    a.ring_meta = lp->ring_meta;
    a.ring_mask = lp->ring - 1;
    a.ctl = lp->ctl;
    a.done = lp->done;
    a.done_mask = dw - 1;
    a.reads = lp->reads;
    a.carry_keys = lp->carry_keys;
    a.carry_meta = lp->carry_meta;
    a.carry_seq = lp->carry_seq;
    a.cap_max = (uint32_t)cap_max;
    a.floor = (uint32_t)floor;
    a.growth = sv->cfg.epoch_growth;
    a.idle_ticks = kLoopIdleUs * per_us;
    a.stall_ticks = kLoopStallUs * per_us;
    // environment LTPG_SERVE_SLICE_US: a shorter time slice (10 us at least),
    // so that tests see many launches end with requests carried over
    uint64_t slice_us = kLoopSliceUs;
    if (const char *e = std::getenv("LTPG_SERVE_SLICE_US"))
        slice_us = std::min<uint64_t>(std::max<uint64_t>(std::strtoull(e, nullptr, 10), 10), kLoopSliceUs);
    a.slice_ticks = slice_us * per_us;
    a.ctr = L.ctx->counters;
    a.profile = std::getenv("LTPG_SERVE_PROFILE") != nullptr;
    lp->slice_ns = slice_us * 1000;
    lp->ticks_per_us = (double)per_us;
    return LTPG_OK;
}

// Position of an order-entry type in an epoch's batch.
uint32_t type_rank(uint32_t type)
{
    return type == LTPG_TYPE_NEWORDER      ? 0
           : type == LTPG_TYPE_PAYMENT     ? 1
           : type == LTPG_TYPE_ORDERSTATUS ? 2
           : type == LTPG_TYPE_STOCKLEVEL  ? 3
                                           : 4;
}

const uint32_t *req_data(Lane &L, uint64_t seq) { return reinterpret_cast<const uint32_t *>(L.store.data(seq)); }

// Pack an order-entry epoch's members into the flight's pinned blocks
// (lock held): value-major columns of stride C, TIDs (low 32 bits of the
// sequence numbers) in the
// last column.
void pack_tpcc(Lane &L, Flight &f, uint64_t C)
{
    uint32_t at = 0;
    for (uint32_t q = 0; q < f.n_no; q++, at++) {
        const uint32_t *src = req_data(L, f.members[at]);
        for (uint32_t k = 0; k < kNoFields; k++) f.h_no[(uint64_t)k * C + q] = src[k];
        f.h_no[52 * C + q] = (uint32_t)f.members[at];
    }
    for (uint32_t p = 0; p < f.n_pay; p++, at++) {
        const uint32_t *src = req_data(L, f.members[at]);
        for (uint32_t k = 0; k < kPayFields; k++) f.h_pay[(uint64_t)k * C + p] = src[k];
        f.h_pay[9 * C + p] = (uint32_t)f.members[at];
    }
    for (uint32_t o = 0; o < f.n_os; o++, at++) {
        const uint32_t *src = req_data(L, f.members[at]);
        for (uint32_t k = 0; k < kOsFields; k++) f.h_os[(uint64_t)k * C + o] = src[k];
        f.h_os[kOsFields * C + o] = (uint32_t)f.members[at];
    }
    for (uint32_t l = 0; l < f.n_sl; l++, at++) {
        const uint32_t *src = req_data(L, f.members[at]);
        for (uint32_t k = 0; k < kSlFields; k++) f.h_sl[(uint64_t)k * C + l] = src[k];
        f.h_sl[kSlFields * C + l] = (uint32_t)f.members[at];
    }
    for (uint32_t d = 0; d < f.n_dl; d++, at++) {
        const uint32_t *src = req_data(L, f.members[at]);
        for (uint32_t k = 0; k < kDlFields; k++) f.h_dl[(uint64_t)k * C + d] = src[k];
        f.h_dl[kDlFields * C + d] = (uint32_t)f.members[at];
    }
}

// A Delivery batch over value-major columns of stride C (host or device).
ltpg_delivery_batch delivery_view(uint32_t *dl, uint64_t C, uint32_t n)
{
    return ltpg_delivery_batch{n, 0, dl, dl + C, dl + 2 * C, dl + kDlFields * C};
}

// Read-only batches over value-major columns of stride C (host or device).
void readonly_views(uint32_t *os, uint32_t *sl, uint64_t C, uint32_t n_os, uint32_t n_sl,
                    ltpg_orderstatus_batch *ob, ltpg_stocklevel_batch *sb)
{
    *ob = ltpg_orderstatus_batch{n_os, 0, os, os + C, os + 2 * C, os + 3 * C, os + 4 * C, os + kOsFields * C};
    *sb = ltpg_stocklevel_batch{n_sl, 0, sl, sl + C, sl + kSlFields * C};
}

// Take the next host-formed epoch's members (lock held): re-queued first,
// then new arrivals, both in sequence order (every re-queued request arrived
// before every request still new).  An epoch, whose batch TIDs are the low
// 32 bits of the sequence numbers, ends before the 2^32 boundary of the low
// bits, so its TIDs stay in priority order.
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
    const auto fits = [&](uint64_t seq) {
        return f.members.empty() || (seq >> 32) == (f.members.front() >> 32);
    };
    while (f.members.size() < cap && !L.requeue.empty() && fits(L.requeue.top())) {
        f.members.push_back(L.requeue.top());
        L.requeue.pop();
    }
    while (f.members.size() < cap && !L.fresh.empty() && fits(L.fresh.front())) {
        f.members.push_back(L.fresh.front());
        L.fresh.pop_front();
    }
    if (w == 1) {  // NewOrders, Payments, Order-Status, Stock-Level, Deliveries, each in sequence order
        std::vector<uint32_t> perm(f.members.size());
        for (uint32_t j = 0; j < perm.size(); ++j) perm[j] = j;
        std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) {
            return type_rank(L.store.at(f.members[a]).type) < type_rank(L.store.at(f.members[b]).type);
        });
        std::vector<uint64_t> m(perm.size());
        for (uint32_t j = 0; j < perm.size(); ++j) m[j] = f.members[perm[j]];
        f.members.swap(m);
        f.n_no = f.n_pay = f.n_os = f.n_sl = f.n_dl = 0;
        for (uint64_t s : f.members) {
            const uint32_t ty = L.store.at(s).type;
            f.n_no += ty == LTPG_TYPE_NEWORDER;
            f.n_pay += ty == LTPG_TYPE_PAYMENT;
            f.n_os += ty == LTPG_TYPE_ORDERSTATUS;
            f.n_sl += ty == LTPG_TYPE_STOCKLEVEL;
            f.n_dl += ty == LTPG_TYPE_DELIVERY;
        }
    }
}

// Pack, copy and launch one host-formed epoch (lock held: the request store
// may grow under concurrent submissions).
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
            const uint64_t seq = f.members[t];
            const uint8_t *d = L.store.data(seq);
            h_tids[t] = (uint32_t)seq;
            std::memcpy(h_keys + (uint64_t)t * S, d, S * sizeof(uint64_t));
            std::memcpy(h_meta + (uint64_t)t * S, d + 8ull * S, S * sizeof(uint32_t));
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
        const uint32_t n_no = f.n_no, n_pay = f.n_pay, n_os = f.n_os, n_sl = f.n_sl, n_dl = f.n_dl;
        pack_tpcc(L, f, C);
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
        if (n_dl)
            LTPG_HIP_TRY(hipMemcpy2DAsync(f.d_dl, C * 4, f.h_dl, C * 4, (uint64_t)n_dl * 4, kDlFields + 1,
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
        if (n_dl) {
            const ltpg_delivery_batch db = delivery_view(f.d_dl, C, n_dl);
            if ((rc = ltpg_tpcc_submit_delivery(ctx, &db, LTPG_MEM_DEVICE))) return rc;
        }
        if ((rc = ltpg_epoch_run(ctx, LTPG_PHASE_ALL))) return rc;
        if (sv->cfg.keep_reads) {
            const std::pair<uint32_t *, const uint32_t *> reads[5] = {
                {f.h_reads, ctx->tpcc.d.reads_no}, {f.h_reads2, ctx->tpcc.d.reads_pay},
                {f.h_reads3, ctx->tpcc.d.reads_os}, {f.h_reads4, ctx->tpcc.d.reads_sl},
                {f.h_reads5, ctx->tpcc.d.reads_dl}};
            const uint64_t n_words[5] = {(uint64_t)kNoReads * n_no, (uint64_t)kPayReads * n_pay,
                                         (uint64_t)kOsReads * n_os, (uint64_t)kSlReads * n_sl,
                                         (uint64_t)kDlReads * n_dl};
            for (int k = 0; k < 5; k++)
                if (n_words[k])
                    LTPG_HIP_TRY(hipMemcpyAsync(reads[k].first, reads[k].second, n_words[k] * 4,
                                                hipMemcpyDeviceToHost, ctx->stream));
        }
        PublishSet ps{{ctx->tpcc.d.bitmap_no, ctx->tpcc.d.bitmap_pay, ctx->tpcc.d.bitmap_os, ctx->tpcc.d.bitmap_sl,
                       ctx->tpcc.d.bitmap_dl},
                      {f.h_bits, f.h_bits2, f.h_bits3, f.h_bits4, f.h_bits5},
                      {(uint32_t)words(n_no), (uint32_t)words(n_pay), (uint32_t)words(n_os), (uint32_t)words(n_sl),
                       (uint32_t)words(n_dl)}};
        LTPG_HIP_TRY(launch_publish(ps, f.h_seq, ++f.seq, ctx->stream));
    }
    f.launched_ns = now_ns();
    f.idle_since_ns = 0;
    f.busy = true;
    return LTPG_OK;
}

// The plain per-batch form of one epoch, for comparison (ltpg_serve_config.
// plain): the batch is packed on the host in the reference's layout
// (slot-major key-value ops, value-major order-entry blocks), copied by the
// submit, run phase by phase with a synchronisation after each, and its
// bitmaps (and reads) copied back, one step after the other.  Packing
// happens under the lock (the request store may grow); the device work runs
// without it.  On return the flight holds a published epoch, as the served
// path leaves it.
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
            const uint64_t seq = f.members[t];
            const uint8_t *d = L.store.data(seq);
            const uint64_t *k = reinterpret_cast<const uint64_t *>(d);
            const uint32_t *m = reinterpret_cast<const uint32_t *>(d + 8ull * S);
            h_tids[t] = (uint32_t)seq;
            for (uint32_t s = 0; s < S; s++) {
                h_keys[(uint64_t)s * B + t] = k[s];
                h_meta[(uint64_t)s * B + t] = m[s];
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
        const uint32_t n_no = f.n_no, n_pay = f.n_pay, n_os = f.n_os, n_sl = f.n_sl, n_dl = f.n_dl;
        pack_tpcc(L, f, C);
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
        if (!rc && n_dl) {
            const ltpg_delivery_batch db = delivery_view(f.h_dl, C, n_dl);
            rc = ltpg_tpcc_submit_delivery(ctx, &db, LTPG_MEM_HOST);
        }
        if (!rc) rc = run_epoch_stepwise(ctx);
        const uint32_t types[5] = {LTPG_TYPE_NEWORDER, LTPG_TYPE_PAYMENT, LTPG_TYPE_ORDERSTATUS, LTPG_TYPE_STOCKLEVEL,
                                   LTPG_TYPE_DELIVERY};
        const uint32_t counts[5] = {n_no, n_pay, n_os, n_sl, n_dl};
        const uint32_t per[5] = {kNoReads, kPayReads, kOsReads, kSlReads, kDlReads};
        uint64_t *bits[5] = {f.h_bits, f.h_bits2, f.h_bits3, f.h_bits4, f.h_bits5};
        uint32_t *reads[5] = {f.h_reads, f.h_reads2, f.h_reads3, f.h_reads4, f.h_reads5};
        for (int k = 0; k < 5 && !rc; k++) {
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

// A committed request: its completion record, its read values (keep_reads),
// and its record's release.
void commit_request(ltpg_server *sv, Lane &L, uint64_t seq, uint64_t epoch, uint64_t t_now,
                    std::vector<uint32_t> *reads)
{
    Rec &r = L.store.at(seq);
    sv->committed[r.type]++;
    sv->outstanding--;
    r.state = kCommitted;
    ltpg_serve_done d;
    d.id = r.id;
    d.type = r.type;
    d.tid = (uint32_t)seq;
    d.tid_hi = (uint32_t)(seq >> 32);
    d.attempts = r.attempts;
    d.epoch = epoch;
    d.arrival_ns = r.arrival;
    d.commit_ns = t_now;
    sv->done.push_back(d);
    if (reads) sv->reads[r.id] = std::move(*reads);
}

// Each read's field, each scan's rows inside the table, in slot order, of
// a key-value request with metas m and keys k; src(s, q) is row q of slot s.
template <typename Src>
std::vector<uint32_t> kv_request_reads(const KvState &kv, uint32_t S, const uint64_t *k, const uint32_t *m, Src src)
{
    std::vector<uint32_t> v;
    const uint32_t fw = kv.field_words;
    for (uint32_t s = 0; s < S; s++) {
        const uint32_t kind = LTPG_KV_KIND(m[s]);
        const uint64_t row = k[s];
        if (row >= kv.rows || LTPG_KV_FIELD(m[s]) >= kv.fields) continue;
        uint64_t n = 0;
        if (kind == LTPG_KV_READ)
            n = 1;
        else if (kind == LTPG_KV_SCAN && LTPG_KV_SCAN_LEN(m[s]) <= kv.max_scan)
            n = std::min<uint64_t>(LTPG_KV_SCAN_LEN(m[s]), kv.rows - row);
        for (uint64_t q = 0; q < n; q++) {
            const uint32_t *p = src(s, q);
            v.insert(v.end(), p, p + fw);
        }
    }
    return v;
}

// Read a finished host-formed epoch's results (lock held).
void complete(ltpg_server *sv, int w, Flight &f)
{
    Lane &L = sv->lane[w];
    const uint64_t t_now = now_ns();
    const uint32_t B = (uint32_t)f.members.size();
    const uint32_t S = sv->cfg.kv_ops;
    uint64_t commits = 0;
    for (uint32_t t = 0; t < B; t++) {
        const uint64_t seq = f.members[t];
        Rec &r = L.store.at(seq);
        r.attempts++;
        sv->attempts[r.type]++;
        bool ok;
        // order-entry members: NewOrders, Payments, Order-Status, Stock-Level, Deliveries
        const uint32_t c1 = f.n_no, c2 = c1 + f.n_pay, c3 = c2 + f.n_os, c4 = c3 + f.n_sl;
        if (w == 0) ok = bit(f.h_bits, t);
        else if (t < c1) ok = bit(f.h_bits, t);
        else if (t < c2) ok = bit(f.h_bits2, t - c1);
        else if (t < c3) ok = bit(f.h_bits3, t - c2);
        else if (t < c4) ok = bit(f.h_bits4, t - c3);
        else ok = bit(f.h_bits5, t - c4);
        if (!ok) {
            L.requeue.push(seq);
            continue;
        }
        commits++;
        std::vector<uint32_t> v;
        if (sv->cfg.keep_reads) {
            if (w == 0) {
                const KvState &kv = L.ctx->kv;
                const uint8_t *d = L.store.data(seq);
                v = kv_request_reads(kv, S, reinterpret_cast<const uint64_t *>(d),
                                     reinterpret_cast<const uint32_t *>(d + 8ull * S), [&](uint32_t s, uint64_t q) {
                                         return f.h_reads + ((((uint64_t)s * B) + t) * kv.read_slot + q) * kv.field_words;
                                     });
            } else if (t < c1) {
                for (uint32_t k = 0; k < kNoReads; k++) v.push_back(f.h_reads[(uint64_t)k * f.n_no + t]);
            } else if (t < c2) {
                const uint32_t p = t - c1;
                for (uint32_t k = 0; k < kPayReads; k++) v.push_back(f.h_reads2[(uint64_t)k * f.n_pay + p]);
            } else if (t < c3) {
                const uint32_t o = t - c2;
                for (uint32_t k = 0; k < kOsReads; k++) v.push_back(f.h_reads3[(uint64_t)k * f.n_os + o]);
            } else if (t < c4) {
                const uint32_t l = t - c3;
                for (uint32_t k = 0; k < kSlReads; k++) v.push_back(f.h_reads4[(uint64_t)k * f.n_sl + l]);
            } else {
                const uint32_t d = t - c4;
                for (uint32_t k = 0; k < kDlReads; k++) v.push_back(f.h_reads5[(uint64_t)k * f.n_dl + d]);
            }
        }
        commit_request(sv, L, seq, f.epoch, t_now, sv->cfg.keep_reads ? &v : nullptr);
    }
    L.last_commits = commits;
    L.store.retire();
    f.busy = false;
}

// The flight's epoch has been published (its results are in host memory).
bool published(const Flight &f) { return ld_acquire(f.h_seq) == f.seq; }

// State of a host-formed epoch in flight: hipSuccess once published,
// hipErrorNotReady while pending, else the error that ends it.  A device
// error never publishes: after 2 ms the stream is asked, and an idle stream
// whose epoch stays unpublished for 100 ms counts as an error (measured
// from when the stream was first seen idle, with the publication re-read
// after the stream query).
hipError_t flight_state(ltpg_server *sv, Lane &L, Flight &f)
{
    if (published(f)) return sv->completed_epochs++ == sv->fail_epoch ? hipErrorUnknown : hipSuccess;
    const uint64_t t = now_ns();
    if (t - f.launched_ns < 2000000ull) return hipErrorNotReady;
    const hipError_t q = hipStreamQuery(L.ctx->stream);
    if (published(f)) return sv->completed_epochs++ == sv->fail_epoch ? hipErrorUnknown : hipSuccess;
    if (q == hipErrorNotReady) {
        f.idle_since_ns = 0;
        return q;
    }
    if (q != hipSuccess) return q;
    if (!f.idle_since_ns) f.idle_since_ns = t;
    return t - f.idle_since_ns > 100000000ull ? hipErrorUnknown : hipErrorNotReady;
}

bool any_flight_done(ltpg_server *sv)
{
    for (auto &L : sv->lane) {
        if (!L.ctx) continue;
        if (L.loop && L.loop->launched &&
            (ld_acquire(&L.loop->ctl->done_tail) != L.loop->done_head || ld_acquire(&L.loop->ctl->running) == 0))
            return true;
        for (auto &f : L.flights)
            if (f.busy && published(f)) return true;
    }
    return false;
}

// Drain the device loop's done ring (lock held): every published epoch
// record, in order.  Returns whether there was one.
bool drain_loop(ltpg_server *sv)
{
    Lane &L = sv->lane[0];
    Loop &lp = *L.loop;
    const uint64_t tail = ld_acquire(&lp.ctl->done_tail);
    if (tail == lp.done_head) return false;
    const uint64_t t_now = now_ns();
    const uint64_t mask = lp.done_words - 1;
    const uint32_t S = sv->cfg.kv_ops;
    const KvState &kv = L.ctx->kv;
    const uint64_t entry_words = (uint64_t)S * kv.read_slot * kv.field_words;
    uint64_t h = lp.done_head;
    while (h < tail) {
        const uint64_t hdr = lp.done[h & mask];
        const uint32_t B = (uint32_t)(hdr & 0xffffu);
        if (sv->completed_epochs++ == sv->fail_epoch) {
            fail(sv, LTPG_ERR_DEVICE, "serving epoch: injected failure (LTPG_TEST_FAIL_EPOCH)");
            return true;
        }
        const uint64_t epoch = L.epochs++;
        L.max_epoch = std::max<uint64_t>(L.max_epoch, B);
        std::vector<uint64_t> ids;
        for (uint32_t t = 0; t < B; t++) {
            const uint64_t x = h + 1 + t;
            const uint64_t wd = lp.done[x & mask];
            const uint64_t seq = wd & ~kDoneCommitted;
            if (!L.store.has(seq)) {
                fail(sv, LTPG_ERR_DEVICE, "device loop published an unknown request");
                return true;
            }
            Rec &r = L.store.at(seq);
            r.attempts++;
            sv->attempts[LTPG_TYPE_KV]++;
            if (sv->cfg.record_epochs) ids.push_back(r.id);
            if (!(wd & kDoneCommitted)) continue;
            if (sv->cfg.keep_reads) {
                const uint8_t *d = L.store.data(seq);
                std::vector<uint32_t> v = kv_request_reads(
                    kv, S, reinterpret_cast<const uint64_t *>(d), reinterpret_cast<const uint32_t *>(d + 8ull * S),
                    [&](uint32_t s, uint64_t q) {
                        return lp.reads + (x & mask) * entry_words + ((uint64_t)s * kv.read_slot + q) * kv.field_words;
                    });
                commit_request(sv, L, seq, epoch, t_now, &v);
            } else {
                commit_request(sv, L, seq, epoch, t_now, nullptr);
            }
        }
        if (sv->cfg.record_epochs) L.log.push_back(std::move(ids));
        h += 1 + B;
    }
    lp.done_head = h;
    st_release(&lp.ctl->done_head, h);
    L.store.retire();
    return true;
}

// Launch, watch and relaunch the device loop (lock held).  Returns whether
// anything changed.
bool tend_loop(ltpg_server *sv)
{
    Lane &L = sv->lane[0];
    Loop &lp = *L.loop;
    if (lp.launched) {
        if (ld_acquire(&lp.ctl->running) == 0) {
            lp.launched = false;
            const uint64_t code = lp.ctl->exit_code;
            lp.exits[code < 5 ? code : 0]++;
            if (sv->trace)
                std::fprintf(stderr, "[serve] device loop exit %llu after %.3f ms: head %llu of %llu, carry %llu, "
                                     "%llu epochs\n",
                             (unsigned long long)lp.ctl->exit_code, (now_ns() - lp.launched_ns) * 1e-6,
                             (unsigned long long)lp.ctl->head, (unsigned long long)L.next_seq,
                             (unsigned long long)lp.ctl->n_carry, (unsigned long long)lp.ctl->epochs);
            if (sv->trace && lp.args.profile && lp.ctl->prof_epochs) {
                static const char *names[kServeProfParts] = {"control", "admit", "execute", "mark", "check",
                                                             "unused", "write", "publish", "requeue", "idle"};
                std::fprintf(stderr, "[serve] profile over %llu epochs, us per epoch:",
                             (unsigned long long)lp.ctl->prof_epochs);
                for (int k = 0; k < kServeProfParts; ++k)
                    std::fprintf(stderr, " %s %.2f", names[k],
                                 lp.ctl->prof[k] / lp.ticks_per_us / (double)lp.ctl->prof_epochs);
                std::fprintf(stderr, "\n");
            }
            return true;
        }
        // the kernel ends its own slice; past that, ask the stream
        if (now_ns() - lp.launched_ns > lp.slice_ns + 1000000000ull) {
            const hipError_t q = hipStreamQuery(L.ctx->stream);
            if (q == hipErrorNotReady) return false;
            if (ld_acquire(&lp.ctl->running) == 0) {
                lp.launched = false;
                return true;
            }
            fail(sv, LTPG_ERR_DEVICE,
                 std::string("device loop ended without finishing: ") + hipGetErrorString(q == hipSuccess ? hipErrorUnknown : q));
            lp.launched = false;
            return true;
        }
        return false;
    }
    const bool pending = ld_acquire(&lp.ctl->head) != L.next_seq || ld_acquire(&lp.ctl->n_carry) != 0;
    if (!pending || sv->error || sv->cancel || ld_acquire(&lp.ctl->stop)) return false;
    st_release(&lp.ctl->running, 1);
    const hipError_t e = launch_kv_serve_loop(lp.args, L.ctx->stream);
    if (e != hipSuccess) {
        st_release(&lp.ctl->running, 0);
        fail(sv, LTPG_ERR_DEVICE, std::string("device loop launch: ") + hipGetErrorString(e));
        return true;
    }
    lp.launched = true;
    lp.launched_ns = now_ns();
    lp.launches++;
    if (sv->trace)
        std::fprintf(stderr, "[serve] device loop launch %llu: head %llu of %llu, carry %llu\n",
                     (unsigned long long)lp.launches, (unsigned long long)lp.ctl->head, (unsigned long long)L.next_seq,
                     (unsigned long long)lp.ctl->n_carry);
    return true;
}

// Drop every request of a lane that is not in an epoch (cancel; lock held):
// new and re-queued ones of a host-formed lane, or, once the device loop has
// stopped, every uncommitted one of its lane (in the ring or carried).
void drop_request(ltpg_server *sv, Lane &L, uint64_t seq)
{
    Rec &r = L.store.at(seq);
    if (r.state != kPending) return;
    r.state = kDropped;
    sv->dropped++;
    sv->outstanding--;
}

void drop_pending(ltpg_server *sv, Lane &L)
{
    if (L.loop) {
        for (uint64_t s = L.store.base(); s < L.store.end(); ++s) drop_request(sv, L, s);
        L.loop->ctl->n_carry = 0;
        L.loop->ctl->head = L.next_seq;  // nothing of the ring is taken any more
    }
    for (uint64_t s : L.fresh) drop_request(sv, L, s);
    L.fresh.clear();
    while (!L.requeue.empty()) {
        drop_request(sv, L, L.requeue.top());
        L.requeue.pop();
    }
    L.store.retire();
}

void serve_loop(ltpg_server *sv)
{
    (void)hipSetDevice((sv->lane[0].ctx ? sv->lane[0].ctx : sv->lane[1].ctx)->device);
    std::unique_lock<std::mutex> lk(sv->mu);
    Loop *lp = sv->lane[0].loop;
    bool dropped_all = false;
    while (true) {
        bool progress = false, inflight = false;
        if ((sv->stop && (sv->cancel || sv->outstanding == 0 || sv->error)) && lp && lp->launched)
            st_release(&lp->ctl->stop, 1);  // the device loop ends at its next epoch boundary
        if (lp) {
            progress |= drain_loop(sv);
            progress |= tend_loop(sv);  // launches only while work is pending and nothing stops it
            inflight |= lp->launched;
        }
        if (sv->cancel && !dropped_all && !(lp && lp->launched)) {
            // the device loop (if any) has stopped: what it had not committed
            // is dropped with the rest; epochs in flight finish and their
            // aborts are dropped as they complete
            if (lp) progress |= drain_loop(sv);
            for (auto &L : sv->lane)
                if (L.ctx) drop_pending(sv, L);
            if (sv->trace) std::fprintf(stderr, "[serve] cancel: %llu dropped\n", (unsigned long long)sv->dropped);
            dropped_all = true;
            progress = true;
        }
        for (int w = 0; w < 2; w++) {
            Lane &L = sv->lane[w];
            if (!L.ctx || (w == 0 && lp)) continue;
            // results first: re-queued aborts join the next epoch formed
            for (auto &f : L.flights) {
                if (!f.busy) continue;
                const hipError_t q = flight_state(sv, L, f);
                if (q == hipSuccess) {
                    complete(sv, w, f);
                    if (sv->cancel) drop_pending(sv, L);  // aborts of a cancelled run are dropped too
                    progress = true;
                } else if (q != hipErrorNotReady) {
                    fail(sv, LTPG_ERR_DEVICE, std::string("serving epoch: ") + hipGetErrorString(q) +
                                                  (sv->completed_epochs - 1 == sv->fail_epoch
                                                       ? " (injected failure, LTPG_TEST_FAIL_EPOCH)" : ""));
                    f.busy = false;
                    progress = true;
                }
            }
            // start epochs on free slots, in slot order so epochs stay ordered
            for (auto &f : L.flights) {
                if (f.busy || (L.fresh.empty() && L.requeue.empty()) || sv->error) continue;
                take_members(sv, w, f);
                if (f.members.empty()) continue;
                f.epoch = L.epochs++;
                L.max_epoch = std::max<uint64_t>(L.max_epoch, f.members.size());
                if (sv->cfg.record_epochs) {
                    std::vector<uint64_t> ids;
                    for (uint64_t s : f.members) ids.push_back(L.store.at(s).id);
                    L.log.push_back(std::move(ids));
                }
                const int rc = sv->cfg.plain ? launch_plain(sv, w, f, lk) : launch(sv, w, f);
                if (rc) fail(sv, rc, ltpg_last_error());
                progress = true;
            }
            for (auto &f : L.flights) inflight |= f.busy;
        }
        if (progress) sv->cv.notify_all();
        bool pending = false;
        for (auto &L : sv->lane) pending |= !L.fresh.empty() || !L.requeue.empty();
        if (lp) pending |= ld_acquire(&lp->ctl->head) != sv->lane[0].next_seq || ld_acquire(&lp->ctl->n_carry) != 0;
        if (sv->stop && !inflight && (!pending || sv->error || sv->cancel)) break;
        if (!progress) {
            lk.unlock();
            if (inflight) {
                // Poll the epochs in flight rather than sleep: a sleep lasts
                // at least a scheduler tick (~50 us), as long as a whole
                // small epoch on the device.  Only this thread touches the
                // flights, so their results are read without the lock; a
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

// Record a new request of lane w (lock held): its sequence number.
uint64_t new_request(ltpg_server *sv, Lane &L, uint32_t type, uint64_t t)
{
    const uint64_t seq = L.next_seq++;
    Rec &r = L.store.push();
    r.id = sv->next_id++;
    r.arrival = t;
    r.type = type;
    return seq;
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
    c->kv_loop = 1;
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
        if (L.loop && L.loop->ctl) st_release(&L.loop->ctl->stop, 1);
        (void)hipStreamSynchronize(L.ctx->stream);
        if (L.copy) (void)hipStreamSynchronize(L.copy);
        for (auto &f : L.flights) free_flight(f);
        free_loop(L.loop);
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
    if (cfg->first_seq > (1ull << 62)) LTPG_FAIL(LTPG_ERR_INVALID, "first_seq must be below 2^62");
    ltpg_server *sv = new (std::nothrow) ltpg_server();
    if (!sv) LTPG_FAIL(LTPG_ERR_NOMEM, "host allocation of server failed");
    sv->cfg = *cfg;
    if (const char *e = std::getenv("LTPG_TEST_FAIL_EPOCH")) sv->fail_epoch = std::strtoull(e, nullptr, 10);
    sv->trace = std::getenv("LTPG_SERVE_TRACE") != nullptr;
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
        L.next_seq = cfg->first_seq;
        const bool loop = w == 0 && cfg->kv_loop && !cfg->plain && !(L.ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS) &&
                          kv_serve_loop_ok(L.ctx->kv, cfg->kv_ops);
        // the request fields the host keeps: order-entry requests and
        // host-formed key-value epochs are packed from them, and kept reads
        // are cut to each request's read slots
        const size_t pay = w == 1 ? kNoFields * sizeof(uint32_t)
                                  : (!loop || cfg->keep_reads ? 12ull * cfg->kv_ops : 0);
        L.store.init(cfg->first_seq, pay);
        if (loop) {
            if ((rc = alloc_loop(sv))) break;
            continue;
        }
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
    Lane &L = sv->lane[0];
    if (!L.ctx) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server has no key-value context");
    const uint64_t t = now_ns();
    const uint64_t S = sv->cfg.kv_ops;
    ClientLock g(sv);
    if (sv->stop) LTPG_FAIL(LTPG_ERR_STATE, "server is stopping");
    if (first_id) *first_id = sv->next_id;
    if (!L.loop) {
        for (uint32_t i = 0; i < n; i++) {
            const uint64_t seq = new_request(sv, L, LTPG_TYPE_KV, t);
            uint8_t *d = L.store.data(seq);
            std::memcpy(d, keys + i * S, 8 * S);
            std::memcpy(d + 8 * S, meta + i * S, 4 * S);
            L.fresh.push_back(seq);
        }
        sv->outstanding += n;
        return LTPG_OK;
    }
    // the device loop: requests go straight into the ring, in runs of
    // consecutive slots; a full ring waits (without the lock) for the
    // kernel to take requests
    Loop &lp = *L.loop;
    uint32_t i = 0;
    while (i < n) {
        if (sv->error) LTPG_FAIL(sv->error, sv->error_msg);
        const uint64_t head = ld_acquire(&lp.ctl->head);
        const uint64_t room = lp.ring - (L.next_seq - head);
        if (room == 0) {
            g.lk.unlock();
            const uint64_t t_wait = now_ns();
            while (ld_acquire(&lp.ctl->head) == head && !sv->error) {
                if (now_ns() - t_wait > 30000000000ull) {
                    g.lk.lock();
                    LTPG_FAIL(LTPG_ERR_DEVICE, "request ring stayed full for 30 s");
                }
                std::this_thread::yield();
            }
            g.lk = ClientLock::lock_for(sv);
            continue;
        }
        const uint64_t slot = L.next_seq & (lp.ring - 1);
        const uint64_t k = std::min<uint64_t>({room, (uint64_t)(n - i), lp.ring - slot});
        std::memcpy(lp.ring_keys + slot * S, keys + (uint64_t)i * S, 8 * S * k);
        std::memcpy(lp.ring_meta + slot * S, meta + (uint64_t)i * S, 4 * S * k);
        for (uint64_t j = 0; j < k; j++) {
            const uint64_t seq = new_request(sv, L, LTPG_TYPE_KV, t);
            if (uint8_t *d = L.store.data(seq)) {
                std::memcpy(d, keys + (i + j) * S, 8 * S);
                std::memcpy(d + 8 * S, meta + (i + j) * S, 4 * S);
            }
        }
        sv->outstanding += k;
        st_release(&lp.ctl->tail, L.next_seq);
        i += (uint32_t)k;
    }
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
    if (sv->stop) LTPG_FAIL(LTPG_ERR_STATE, "server is stopping");
    if (first_id) *first_id = sv->next_id;
    Lane &L = sv->lane[1];
    for (uint32_t q = 0; q < n_no; q++) {
        const uint64_t seq = new_request(sv, L, LTPG_TYPE_NEWORDER, t);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.store.data(seq));
        const uint32_t *heads[7] = {no->W_ID, no->D_ID, no->C_ID, no->O_ID, no->N_O_ID, no->O_OL_CNT, no->O_OL_ID};
        uint32_t k = 0;
        for (auto h : heads) d[k++] = h[q];
        for (const uint32_t *a : {no->OL_I_ID, no->OL_SUPPLY_W_ID, no->OL_QUANTITY})
            for (uint32_t l = 0; l < LTPG_TPCC_MAX_LINES; l++) d[k++] = a[(uint64_t)l * no->line_stride + q];
        L.fresh.push_back(seq);
    }
    for (uint32_t p = 0; p < n_pay; p++) {
        const uint64_t seq = new_request(sv, L, LTPG_TYPE_PAYMENT, t);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.store.data(seq));
        const uint32_t *cols[9] = {pay->W_ID, pay->D_ID, pay->C_ID, pay->C_LAST, pay->isName,
                                   pay->C_D_ID, pay->C_W_ID, pay->H_AMOUNT, pay->H_ID};
        uint32_t k = 0;
        for (auto c : cols) d[k++] = c[p];
        L.fresh.push_back(seq);
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
    if (sv->stop) LTPG_FAIL(LTPG_ERR_STATE, "server is stopping");
    if (first_id) *first_id = sv->next_id;
    Lane &L = sv->lane[1];
    for (uint32_t o = 0; o < n_os; o++) {
        const uint64_t seq = new_request(sv, L, LTPG_TYPE_ORDERSTATUS, t);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.store.data(seq));
        uint32_t k = 0;
        for (const uint32_t *c : {os->W_ID, os->D_ID, os->C_ID, os->O_ID, os->OL_ID}) d[k++] = c[o];
        L.fresh.push_back(seq);
    }
    for (uint32_t l = 0; l < n_sl; l++) {
        const uint64_t seq = new_request(sv, L, LTPG_TYPE_STOCKLEVEL, t);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.store.data(seq));
        d[0] = sl->W_ID[l];
        d[1] = sl->S_ID[l];
        L.fresh.push_back(seq);
    }
    sv->outstanding += n_os + n_sl;
    return LTPG_OK;
}

int ltpg_serve_submit_delivery(ltpg_server *sv, const ltpg_delivery_batch *dl, uint64_t *first_id)
{
    if (!sv) LTPG_FAIL(LTPG_ERR_INVALID, "null server");
    if (!sv->lane[1].ctx) LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "server has no order-entry context");
    if (sv->lane[1].ctx->cfg.flags & LTPG_FLAG_CONFLICT_COUNTS)
        LTPG_FAIL(LTPG_ERR_UNSUPPORTED, "Deliveries need the fast check (no conflict counts)");
    const uint32_t n = dl ? dl->n : 0;
    if (n && (!dl->O_ID || !dl->NO_W_ID || !dl->NO_O_ID)) LTPG_FAIL(LTPG_ERR_INVALID, "bad Delivery arrays");
    const uint64_t t = now_ns();
    ClientLock g(sv);
    if (sv->stop) LTPG_FAIL(LTPG_ERR_STATE, "server is stopping");
    if (first_id) *first_id = sv->next_id;
    Lane &L = sv->lane[1];
    for (uint32_t i = 0; i < n; i++) {
        const uint64_t seq = new_request(sv, L, LTPG_TYPE_DELIVERY, t);
        uint32_t *d = reinterpret_cast<uint32_t *>(L.store.data(seq));
        d[0] = dl->O_ID[i];
        d[1] = dl->NO_W_ID[i];
        d[2] = dl->NO_O_ID[i];
        L.fresh.push_back(seq);
    }
    sv->outstanding += n;
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
    if (it == sv->reads.end()) LTPG_FAIL(LTPG_ERR_STATE, "request has not committed, or its values were read");
    const auto &v = it->second;
    if (n_words) *n_words = (uint32_t)v.size();
    if (out) {
        std::copy(v.begin(), v.begin() + std::min<size_t>(v.size(), max_words), out);
        sv->reads.erase(it);  // delivered: freed
    }
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

int ltpg_serve_device_stats_get(ltpg_server *sv, ltpg_serve_device_stats *st)
{
    if (!sv || !st) LTPG_FAIL(LTPG_ERR_INVALID, "null argument");
    if (st->size != sizeof(ltpg_serve_device_stats)) LTPG_FAIL(LTPG_ERR_INVALID, "ltpg_serve_device_stats.size mismatch");
    ClientLock g(sv);
    const uint32_t size = st->size;
    std::memset(st, 0, sizeof(*st));
    st->size = size;
    const Loop *lp = sv->lane[0].loop;
    if (!lp) return LTPG_OK;
    st->kv_loop = 1;
    st->launches = lp->launches;
    // the kernel writes these at each exit; while it runs they hold the
    // previous launches' values
    st->epochs = ld_acquire(&lp->ctl->epochs);
    const double ns_per_tick = 1000.0 / lp->ticks_per_us;
    st->busy_ns = (uint64_t)(lp->ctl->busy_ticks * ns_per_tick);
    for (int k = 0; k < 5; k++) st->exits[k] = lp->exits[k];
    for (int k = 0; k < kServeProfParts; k++) st->prof_ns[k] = (uint64_t)(lp->ctl->prof[k] * ns_per_tick);
    st->prof_epochs = lp->ctl->prof_epochs;
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
