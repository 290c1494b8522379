// A C++ caller of include/ltpg.h, built with g++ against libltpg.so the way
// an application links it:
//
//   g++ -std=c++17 -Iinclude tests/abi/integration_loops.cpp -Lltpg_amd -lltpg
//
// It runs the key-value loop and the order-entry loop as a port of
// YCSB/0/ycsb.cu:78-152 and TPCC-allopt-final/tp.cu:136-261 would write them
// (reference-layout query blocks, host-resident batches reused right after
// each submit, column-major snapshots), then one served key-value segment,
// and dumps every input and output as raw little-endian arrays into a folder
// so that a test can replay the same epochs on the oracle.
//
//   integration_loops --host-only        host calls only (no device needed)
//   integration_loops <out-dir>          the loops on device 0
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ltpg.h"

namespace {

constexpr uint64_t YCSB_SIZE = 50000;   // rows
constexpr uint32_t YCSB_COLUMN = 10;    // u32 fields
constexpr uint32_t YCSB_READ_SIZE = 4;
constexpr uint32_t YCSB_WRITE_SIZE = 4;
constexpr uint32_t YCSB_OP_SIZE = YCSB_READ_SIZE + YCSB_WRITE_SIZE;
constexpr uint32_t BATCH_SIZE = 6000;
constexpr int KV_EPOCHS = 3;

constexpr uint32_t WAREHOUSE_SIZE = 2;
constexpr uint32_t TPCC_BATCH = 3000;
constexpr int TPCC_EPOCHS = 3;

constexpr uint32_t SERVE_REQUESTS = 2000;
constexpr uint32_t SERVE_OPS = 4;

int fail(const char *what)
{
    std::fprintf(stderr, "%s failed: %s\n", what, ltpg_last_error());
    return 1;
}

#define CHECK(call)                      \
    do {                                 \
        if ((call) != LTPG_OK) return fail(#call); \
    } while (0)

std::string g_dir;

template <typename T>
bool dump(const std::string &name, const T *p, size_t n)
{
    const std::string path = g_dir + "/" + name;
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, sizeof(T), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

template <typename T>
bool dump(const std::string &name, const std::vector<T> &v)
{
    return dump(name, v.data(), v.size());
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// The reference's YCSB_A_QUERY (YCSB/0/Datastructure.h:36-41): slot-major
// LOC_R[BATCH_SIZE * op + NO], LOC_W likewise, TID[NO]; row ids here are u64.
struct ycsb_a_query {
    std::vector<uint64_t> LOC_R, LOC_W;
    std::vector<uint32_t> TID;
};

// Half of the ops go to a 2000-row hot set so that the epochs abort some
// transactions.
void make_ycsb(ycsb_a_query &q, uint64_t &seed, uint32_t tid_base)
{
    q.LOC_R.resize(YCSB_READ_SIZE * BATCH_SIZE);
    q.LOC_W.resize(YCSB_WRITE_SIZE * BATCH_SIZE);
    q.TID.resize(BATCH_SIZE);
    auto key = [&]() {
        const uint64_t x = splitmix(seed);
        return (x & 1) ? (x >> 8) % 2000 : (x >> 8) % YCSB_SIZE;
    };
    for (auto &k : q.LOC_R) k = key();
    for (auto &k : q.LOC_W) k = key();
    for (uint32_t no = 0; no < BATCH_SIZE; no++) q.TID[no] = tid_base + no;
}

int host_only()
{
    if (ltpg_abi_version() != LTPG_ABI_VERSION) {
        std::fprintf(stderr, "abi version %d, header %d\n", ltpg_abi_version(), LTPG_ABI_VERSION);
        return 1;
    }
    ltpg_config cfg;
    ltpg_config_init(&cfg);
    if (cfg.size != sizeof(ltpg_config)) return fail("ltpg_config_init size");
    ltpg_serve_config sc;
    ltpg_serve_config_init(&sc);
    if (sc.size != sizeof(ltpg_serve_config)) return fail("ltpg_serve_config_init size");
    ltpg_kv_gen g;
    ltpg_kv_gen_init(&g);
    g.seed = 7;
    g.rows = 1000;
    g.fields = 10;
    g.n_txns = 64;
    g.n_slots = 4;
    g.min_ops = 1;
    g.max_ops = 4;
    std::vector<uint64_t> keys(256);
    std::vector<uint32_t> meta(256), tids(64);
    CHECK(ltpg_kv_generate(&g, keys.data(), meta.data(), tids.data(), 1));
    for (uint32_t i = 0; i < 256; i++)
        if ((meta[i] & 3) != LTPG_KV_NONE && keys[i] >= 1000) return fail("generated key range");
    ltpg_tpcc_gen_config tc;
    ltpg_tpcc_gen_config_init(&tc);
    tc.warehouses = 2;
    ltpg_tpcc_gen *tg;
    CHECK(ltpg_tpcc_gen_create(&tc, &tg));
    std::vector<uint32_t> no(53 * 100), pay(10 * 100);
    uint32_t n_no = 0, n_pay = 0;
    CHECK(ltpg_tpcc_gen_next(tg, 100, 0, no.data(), 100, pay.data(), 100, &n_no, &n_pay));
    CHECK(ltpg_tpcc_gen_destroy(tg));
    if (n_no + n_pay != 100) return fail("ltpg_tpcc_gen_next counts");
    ltpg_neworder_batch nb;
    ltpg_neworder_view(no.data(), 100, n_no, &nb);
    if (nb.TID != no.data() + 52 * 100 || nb.OL_I_ID != no.data() + 7 * 100) return fail("ltpg_neworder_view");
    std::printf("host-only ok: abi %d, devices %d\n", ltpg_abi_version(), ltpg_device_count());
    return 0;
}

// YCSB/0/ycsb.cu:78-152 over the C-ABI.
int kv_loop()
{
    ltpg_config cfg;
    ltpg_config_init(&cfg);
    cfg.workload = LTPG_WORKLOAD_KV;
    cfg.kv_rows = YCSB_SIZE;
    cfg.kv_fields = YCSB_COLUMN;
    cfg.kv_field_bytes = 4;
    cfg.kv_max_slots = YCSB_OP_SIZE;
    cfg.max_txns = BATCH_SIZE;
    cfg.values = LTPG_VALUES_RMW;  // committed writes change the table, so the dump pins write-back
    ltpg_ctx *ctx;
    CHECK(ltpg_create(&cfg, &ctx));

    // the host snapshot is column-major u32 (YCSB/0/Datastructure.h:16-19)
    std::vector<uint32_t> snapshot(YCSB_SIZE * YCSB_COLUMN);
    for (uint32_t c = 0; c < YCSB_COLUMN; c++)
        for (uint64_t r = 0; r < YCSB_SIZE; r++) snapshot[c * YCSB_SIZE + r] = (c << 24) ^ (uint32_t)(r * 2654435761u);
    CHECK(ltpg_kv_table_load(ctx, snapshot.data(), YCSB_SIZE, LTPG_LAYOUT_COLUMN_MAJOR));
    if (!dump("kv_init.u32", snapshot)) return fail("dump");

    uint64_t seed = 0x4C545047;
    std::vector<uint64_t> keys(YCSB_OP_SIZE * BATCH_SIZE);
    std::vector<uint32_t> meta(YCSB_OP_SIZE * BATCH_SIZE);
    std::vector<uint32_t> tids(BATCH_SIZE);
    std::vector<uint64_t> bitmap((BATCH_SIZE + 63) / 64);
    std::vector<uint32_t> reads(YCSB_OP_SIZE * BATCH_SIZE);
    for (int epoch = 0; epoch < KV_EPOCHS; epoch++) {
        ycsb_a_query q;
        make_ycsb(q, seed, epoch * BATCH_SIZE);
        for (uint32_t op = 0; op < YCSB_READ_SIZE; op++)
            for (uint32_t no = 0; no < BATCH_SIZE; no++) {
                keys[op * BATCH_SIZE + no] = q.LOC_R[BATCH_SIZE * op + no];
                meta[op * BATCH_SIZE + no] = LTPG_KV_META(LTPG_KV_READ, (op + epoch) % YCSB_COLUMN);
            }
        for (uint32_t op = 0; op < YCSB_WRITE_SIZE; op++)
            for (uint32_t no = 0; no < BATCH_SIZE; no++) {
                keys[(YCSB_READ_SIZE + op) * BATCH_SIZE + no] = q.LOC_W[BATCH_SIZE * op + no];
                meta[(YCSB_READ_SIZE + op) * BATCH_SIZE + no] = LTPG_KV_META(LTPG_KV_WRITE, (op * 3 + epoch) % YCSB_COLUMN);
            }
        tids = q.TID;
        const std::string e = std::to_string(epoch);
        if (!dump("kv_keys" + e + ".u64", keys) || !dump("kv_meta" + e + ".u32", meta) ||
            !dump("kv_tids" + e + ".u32", tids))
            return fail("dump");
        ltpg_kv_batch b = {BATCH_SIZE, YCSB_OP_SIZE, keys.data(), meta.data(), tids.data()};
        CHECK(ltpg_kv_submit(ctx, &b, LTPG_MEM_HOST));
        // the batch has been taken: the caller's arrays are free again
        std::fill(keys.begin(), keys.end(), ~0ull);
        std::fill(meta.begin(), meta.end(), 0xFFFFFFFFu);
        std::fill(tids.begin(), tids.end(), 0u);
        CHECK(ltpg_epoch_run(ctx, LTPG_PHASE_ALL));
        CHECK(ltpg_sync(ctx));
        uint32_t commits = 0;
        CHECK(ltpg_commit_count(ctx, LTPG_TYPE_KV, &commits));
        CHECK(ltpg_commit_bitmap(ctx, LTPG_TYPE_KV, bitmap.data(), bitmap.size()));
        CHECK(ltpg_kv_read_values(ctx, reads.data(), reads.size()));
        if (!dump("kv_commits" + e + ".u32", &commits, 1) || !dump("kv_bitmap" + e + ".u64", bitmap) ||
            !dump("kv_reads" + e + ".u32", reads))
            return fail("dump");
        std::printf("kv epoch %d: %u of %u committed\n", epoch, commits, BATCH_SIZE);
    }
    uint64_t commits = 0, attempts = 0;
    CHECK(ltpg_commit_totals(ctx, LTPG_TYPE_KV, &commits, &attempts, 0));
    const uint64_t totals[2] = {commits, attempts};
    std::vector<uint32_t> table(YCSB_SIZE * YCSB_COLUMN);
    CHECK(ltpg_kv_table_read(ctx, table.data(), YCSB_SIZE, LTPG_LAYOUT_COLUMN_MAJOR));
    if (!dump("kv_totals.u64", totals, 2) || !dump("kv_final.u32", table)) return fail("dump");
    // a few rows through the row read-out
    const uint64_t some[3] = {0, 1999, YCSB_SIZE - 1};
    std::vector<uint32_t> rows(3 * YCSB_COLUMN);
    CHECK(ltpg_kv_table_read_rows(ctx, some, 3, rows.data()));
    for (int i = 0; i < 3; i++)
        for (uint32_t c = 0; c < YCSB_COLUMN; c++)
            if (rows[i * YCSB_COLUMN + c] != table[c * YCSB_SIZE + some[i]]) return fail("ltpg_kv_table_read_rows");
    CHECK(ltpg_destroy(ctx));
    return 0;
}

// TPCC-allopt-final/tp.cu:136-261 over the C-ABI, the blocks packed as the
// reference's NEWORDER_QUERY / PAYMENT_QUERY (Datastructure.h:211-249).
int tpcc_loop()
{
    ltpg_config cfg;
    ltpg_config_init(&cfg);
    cfg.workload = LTPG_WORKLOAD_TPCC;
    cfg.tpcc_warehouses = WAREHOUSE_SIZE;
    cfg.max_txns = TPCC_BATCH;
    cfg.rules = LTPG_RULES_REFERENCE;
    cfg.values = LTPG_VALUES_RMW;
    ltpg_ctx *ctx;
    CHECK(ltpg_create(&cfg, &ctx));
    const uint64_t seed = 0x4C545047;
    CHECK(ltpg_tpcc_populate(ctx, seed));

    ltpg_tpcc_gen_config gc;
    ltpg_tpcc_gen_config_init(&gc);
    gc.seed = seed;
    gc.warehouses = WAREHOUSE_SIZE;
    ltpg_tpcc_gen *gen;
    CHECK(ltpg_tpcc_gen_create(&gc, &gen));
    const uint32_t NEWORDER_CNT = TPCC_BATCH, PAYMENT_CNT = TPCC_BATCH;
    std::vector<uint32_t> neworder_query(53 * NEWORDER_CNT), payment_query(10 * PAYMENT_CNT);
    for (int epoch = 0; epoch < TPCC_EPOCHS; epoch++) {
        uint32_t n_no = 0, n_pay = 0;
        CHECK(ltpg_tpcc_gen_next(gen, TPCC_BATCH, epoch * TPCC_BATCH, neworder_query.data(), NEWORDER_CNT,
                                 payment_query.data(), PAYMENT_CNT, &n_no, &n_pay));
        const std::string e = std::to_string(epoch);
        const uint32_t counts[2] = {n_no, n_pay};
        if (!dump("no_block" + e + ".u32", neworder_query) || !dump("pay_block" + e + ".u32", payment_query) ||
            !dump("tpcc_n" + e + ".u32", counts, 2))
            return fail("dump");
        ltpg_neworder_batch no;
        ltpg_payment_batch pay;
        ltpg_neworder_view(neworder_query.data(), NEWORDER_CNT, n_no, &no);
        ltpg_payment_view(payment_query.data(), PAYMENT_CNT, n_pay, &pay);
        CHECK(ltpg_tpcc_submit(ctx, &no, &pay, LTPG_MEM_HOST));
        std::fill(neworder_query.begin(), neworder_query.end(), 0xFFFFFFFFu);
        std::fill(payment_query.begin(), payment_query.end(), 0xFFFFFFFFu);
        CHECK(ltpg_epoch_run(ctx, LTPG_PHASE_ALL));
        CHECK(ltpg_sync(ctx));
        uint32_t c[2];
        CHECK(ltpg_commit_count(ctx, LTPG_TYPE_NEWORDER, &c[0]));
        CHECK(ltpg_commit_count(ctx, LTPG_TYPE_PAYMENT, &c[1]));
        std::vector<uint64_t> nb((n_no + 63) / 64 + 1), pb((n_pay + 63) / 64 + 1);
        CHECK(ltpg_commit_bitmap(ctx, LTPG_TYPE_NEWORDER, nb.data(), (n_no + 63) / 64));
        CHECK(ltpg_commit_bitmap(ctx, LTPG_TYPE_PAYMENT, pb.data(), (n_pay + 63) / 64));
        nb.pop_back();
        pb.pop_back();
        if (!dump("tpcc_commits" + e + ".u32", c, 2) || !dump("no_bitmap" + e + ".u64", nb) ||
            !dump("pay_bitmap" + e + ".u64", pb))
            return fail("dump");
        std::printf("order-entry epoch %d: %u of %u NewOrder, %u of %u Payment committed\n", epoch, c[0], n_no,
                    c[1], n_pay);
    }
    CHECK(ltpg_tpcc_gen_destroy(gen));
    for (uint32_t t = 0; t < LTPG_TPCC_NTABLES; t++) {
        uint64_t rows;
        uint32_t cols;
        CHECK(ltpg_tpcc_table_shape(ctx, t, &rows, &cols));
        std::vector<uint32_t> host(rows * cols);
        CHECK(ltpg_tpcc_table_read(ctx, t, host.data(), rows, LTPG_LAYOUT_COLUMN_MAJOR));
        if (!dump("tpcc_table" + std::to_string(t) + ".u32", host)) return fail("dump");
    }
    CHECK(ltpg_destroy(ctx));
    return 0;
}

// The served form: requests submitted one call at a time, polled to the end.
int serve_segment()
{
    ltpg_config cfg;
    ltpg_config_init(&cfg);
    cfg.workload = LTPG_WORKLOAD_KV;
    cfg.kv_rows = 4096;
    cfg.kv_fields = 4;
    cfg.kv_field_bytes = 4;
    cfg.kv_max_slots = SERVE_OPS;
    cfg.max_txns = 1024;
    cfg.values = LTPG_VALUES_RMW;
    ltpg_ctx *ctx;
    CHECK(ltpg_create(&cfg, &ctx));
    std::vector<uint32_t> zero(4096 * 4, 0);
    CHECK(ltpg_kv_table_load(ctx, zero.data(), 4096, LTPG_LAYOUT_ROW_MAJOR));
    CHECK(ltpg_sync(ctx));
    ltpg_serve_config sc;
    ltpg_serve_config_init(&sc);
    sc.kv_ops = SERVE_OPS;
    ltpg_server *sv;
    CHECK(ltpg_serve_start(ctx, nullptr, &sc, &sv));
    uint64_t seed = 99;
    std::vector<uint64_t> keys(SERVE_REQUESTS * SERVE_OPS);
    std::vector<uint32_t> meta(SERVE_REQUESTS * SERVE_OPS);
    for (uint32_t i = 0; i < SERVE_REQUESTS * SERVE_OPS; i++) {
        const uint64_t x = splitmix(seed);
        keys[i] = (x >> 8) % 64;  // 64 rows: heavy contention, many re-queues
        meta[i] = LTPG_KV_META((x & 3) == 0 ? LTPG_KV_WRITE : LTPG_KV_READ, (x >> 4) % 4);
    }
    for (uint32_t i = 0; i < SERVE_REQUESTS; i += 100) {
        uint64_t first = 0;
        CHECK(ltpg_serve_submit_kv(sv, 100, keys.data() + i * SERVE_OPS, meta.data() + i * SERVE_OPS, &first));
        if (first != i) return fail("request ids in submission order");
    }
    CHECK(ltpg_serve_drain(sv));
    std::vector<ltpg_serve_done> done(SERVE_REQUESTS + 1);
    uint32_t got = 0, n = 0;
    do {
        CHECK(ltpg_serve_poll(sv, done.data() + got, SERVE_REQUESTS + 1 - got, &n));
        got += n;
    } while (n);
    if (got != SERVE_REQUESTS) {
        std::fprintf(stderr, "served %u of %u\n", got, SERVE_REQUESTS);
        return 1;
    }
    // (id, seq, epoch, attempts) per committed request
    std::vector<uint64_t> rec;
    for (uint32_t i = 0; i < got; i++) {
        const ltpg_serve_done &d = done[i];
        if (d.type != LTPG_TYPE_KV || d.attempts == 0 || d.commit_ns < d.arrival_ns) return fail("done record");
        rec.insert(rec.end(), {d.id, ((uint64_t)d.tid_hi << 32) | d.tid, d.epoch, d.attempts});
    }
    ltpg_serve_stats st;
    CHECK(ltpg_serve_stats_get(sv, &st));
    CHECK(ltpg_serve_stop(sv));
    std::vector<uint32_t> table(4096 * 4);
    CHECK(ltpg_kv_table_read(ctx, table.data(), 4096, LTPG_LAYOUT_ROW_MAJOR));
    if (!dump("serve_keys.u64", keys) || !dump("serve_meta.u32", meta) || !dump("serve_done.u64", rec) ||
        !dump("serve_final.u32", table))
        return fail("dump");
    std::printf("served %u requests in %llu epochs\n", got, (unsigned long long)st.epochs[0]);
    CHECK(ltpg_destroy(ctx));
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s --host-only | <out-dir>\n", argv[0]);
        return 2;
    }
    if (!std::strcmp(argv[1], "--host-only")) return host_only();
    g_dir = argv[1];
    if (ltpg_device_count() < 1) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    if (int rc = host_only()) return rc;
    if (int rc = kv_loop()) return rc;
    if (int rc = tpcc_loop()) return rc;
    if (int rc = serve_segment()) return rc;
    std::printf("all loops done\n");
    return 0;
}
