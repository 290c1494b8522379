// gen_tpcc.cpp - seeded order-entry (NewOrder / Payment) batch generator.
//
// Replaces MakeNeworder / MakePayment and the batch filler of
// LTPG/TPCC-allopt-final/Query.h:16-212,404-525, whose draws come from
// rand() after srand(time(0)) (Random.h:12,52) on many threads and so
// cannot be reproduced.  Here every draw is SplitMix64 of a counter keyed by
// (seed, global transaction index, field, attempt), so a batch is a pure
// function of the configuration and the number of transactions generated
// before it.
//
// The reference's commented-out options are switchable and on by default:
// NURand items (Query.h:105, Random.h non_uniform_distribution without the
// run-time constant C), remote supply warehouses (Query.h:116-125) and remote
// customers (Query.h:189-201), and H_AMOUNT uniform in [1, 5] (Query.h:207).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "ltpg.h"

namespace ltpg {
void set_error(const std::string &msg);
}

struct ltpg_tpcc_gen {
    ltpg_tpcc_gen_config cfg;
    uint64_t next_txn = 0;         // global index of the next transaction
    std::vector<uint32_t> noid;    // per warehouse: next N_O_ID  (Query.h:54-63)
    std::vector<uint32_t> oid;     //                next O_ID    (Query.h:65-75)
    std::vector<uint32_t> oolid;   //                next O_OL_ID (Query.h:77-88)
    std::vector<uint32_t> hid;     //                next H_ID    (Query.h:157-166)
};

namespace {

constexpr uint32_t kOrdersPerWarehouse = 30000;     // NEWORDER / ORDER rows per W
constexpr uint32_t kLinesPerWarehouse = 450000;     // ORDERLINE rows per W
constexpr uint32_t kHistoryPerWarehouse = 30000;

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Draw k of field f of global transaction txn.
inline uint64_t draw(uint64_t seed, uint64_t txn, uint32_t f, uint32_t k)
{
    uint64_t h = splitmix64(seed ^ 0x545043432D47454Eull);
    h = splitmix64(h ^ txn);
    return splitmix64(h ^ (((uint64_t)f << 32) | k));
}

inline uint64_t below(uint64_t x, uint64_t n) { return (uint64_t)(((unsigned __int128)x * n) >> 64); }

inline uint32_t uniform(uint64_t x, uint32_t a, uint32_t b) { return a + (uint32_t)below(x, (uint64_t)b - a + 1); }

// Field ids of the draws.
enum : uint32_t {
    F_TYPE = 0, F_W, F_D, F_C, F_CNT, F_QTY, F_ITEM_A, F_ITEM_B, F_REMOTE, F_RW,
    F_ISNAME, F_NAME, F_ORD, F_CREMOTE, F_CW, F_CD, F_AMOUNT,
};

// Index of a warehouse other than home, uniform (W > 1).
inline uint32_t other_warehouse(uint64_t x, uint32_t home, uint32_t W)
{
    const uint32_t r = (uint32_t)below(x, W - 1);
    return r < home ? r : r + 1;
}

}  // namespace

extern "C" {

void ltpg_tpcc_gen_config_init(ltpg_tpcc_gen_config *c)
{
    if (!c) return;
    memset(c, 0, sizeof(*c));
    c->seed = 0x4C545047ull;
    c->warehouses = 64;
    c->neworder_pct = 50;
    c->remote_supply_pct = 1;
    c->remote_customer_pct = 15;
    c->byname_pct = 60;
    c->nurand_items = 1;
    c->min_lines = 5;
    c->max_lines = 15;
    c->max_amount = 5;
    c->max_quantity = 10;
}

int ltpg_tpcc_gen_create(const ltpg_tpcc_gen_config *c, ltpg_tpcc_gen **out)
{
    if (!c || !out) {
        ltpg::set_error("null argument");
        return LTPG_ERR_INVALID;
    }
    *out = nullptr;
    if (c->warehouses == 0 || c->warehouses > 65536 || c->neworder_pct > 100 || c->remote_supply_pct > 100 ||
        c->remote_customer_pct > 100 || c->byname_pct > 100 || c->min_lines < 1 ||
        c->max_lines > LTPG_TPCC_MAX_LINES || c->min_lines > c->max_lines || c->max_amount == 0 ||
        c->max_quantity == 0) {
        ltpg::set_error("invalid order-entry generator configuration");
        return LTPG_ERR_INVALID;
    }
    ltpg_tpcc_gen *g = new (std::nothrow) ltpg_tpcc_gen();
    if (!g) {
        ltpg::set_error("host allocation of generator failed");
        return LTPG_ERR_NOMEM;
    }
    g->cfg = *c;
    g->noid.assign(c->warehouses, 0);
    g->oid.assign(c->warehouses, 0);
    g->oolid.assign(c->warehouses, 0);
    g->hid.assign(c->warehouses, 0);
    *out = g;
    return LTPG_OK;
}

int ltpg_tpcc_gen_destroy(ltpg_tpcc_gen *g)
{
    delete g;
    return LTPG_OK;
}

void ltpg_neworder_view(const uint32_t *block, uint32_t cap, uint32_t n, ltpg_neworder_batch *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->n = n;
    o->line_stride = cap;
    if (!block) return;
    const uint64_t C = cap;
    o->W_ID = block;
    o->D_ID = block + C;
    o->C_ID = block + 2 * C;
    o->O_ID = block + 3 * C;
    o->N_O_ID = block + 4 * C;
    o->O_OL_CNT = block + 5 * C;
    o->O_OL_ID = block + 6 * C;
    o->OL_I_ID = block + 7 * C;
    o->OL_SUPPLY_W_ID = block + 22 * C;
    o->OL_QUANTITY = block + 37 * C;
    o->TID = block + 52 * C;
}

void ltpg_payment_view(const uint32_t *block, uint32_t cap, uint32_t n, ltpg_payment_batch *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->n = n;
    if (!block) return;
    const uint64_t C = cap;
    o->W_ID = block;
    o->D_ID = block + C;
    o->C_ID = block + 2 * C;
    o->C_LAST = block + 3 * C;
    o->isName = block + 4 * C;
    o->C_D_ID = block + 5 * C;
    o->C_W_ID = block + 6 * C;
    o->H_AMOUNT = block + 7 * C;
    o->H_ID = block + 8 * C;
    o->TID = block + 9 * C;
}

int ltpg_tpcc_gen_next(ltpg_tpcc_gen *g, uint32_t n_txns, uint32_t tid_base, uint32_t *no_block,
                       uint32_t no_cap, uint32_t *pay_block, uint32_t pay_cap, uint32_t *n_neworder,
                       uint32_t *n_payment)
{
    if (!g) {
        ltpg::set_error("null generator");
        return LTPG_ERR_INVALID;
    }
    const ltpg_tpcc_gen_config &c = g->cfg;
    const uint32_t n_no = (uint32_t)((uint64_t)n_txns * c.neworder_pct / 100);
    const uint32_t n_pay = n_txns - n_no;
    if (n_no > no_cap || n_pay > pay_cap) {
        ltpg::set_error("batch capacity too small for the NewOrder / Payment split");
        return LTPG_ERR_INVALID;
    }
    if ((n_no && !no_block) || (n_pay && !pay_block)) {
        ltpg::set_error("null batch block");
        return LTPG_ERR_INVALID;
    }
    const uint32_t W = c.warehouses;
    const uint64_t seed = c.seed;
    ltpg_neworder_batch no;
    ltpg_payment_batch pay;
    ltpg_neworder_view(no_block, no_cap, n_no, &no);
    ltpg_payment_view(pay_block, pay_cap, n_pay, &pay);
    uint32_t *NW = const_cast<uint32_t *>(no.W_ID), *ND = const_cast<uint32_t *>(no.D_ID);
    uint32_t *NC = const_cast<uint32_t *>(no.C_ID), *NOI = const_cast<uint32_t *>(no.O_ID);
    uint32_t *NNO = const_cast<uint32_t *>(no.N_O_ID), *NCNT = const_cast<uint32_t *>(no.O_OL_CNT);
    uint32_t *NOL = const_cast<uint32_t *>(no.O_OL_ID), *NI = const_cast<uint32_t *>(no.OL_I_ID);
    uint32_t *NSW = const_cast<uint32_t *>(no.OL_SUPPLY_W_ID), *NQ = const_cast<uint32_t *>(no.OL_QUANTITY);
    uint32_t *NT = const_cast<uint32_t *>(no.TID);
    uint32_t *PW = const_cast<uint32_t *>(pay.W_ID), *PD = const_cast<uint32_t *>(pay.D_ID);
    uint32_t *PC = const_cast<uint32_t *>(pay.C_ID), *PL = const_cast<uint32_t *>(pay.C_LAST);
    uint32_t *PN = const_cast<uint32_t *>(pay.isName), *PCD = const_cast<uint32_t *>(pay.C_D_ID);
    uint32_t *PCW = const_cast<uint32_t *>(pay.C_W_ID), *PA = const_cast<uint32_t *>(pay.H_AMOUNT);
    uint32_t *PH = const_cast<uint32_t *>(pay.H_ID), *PT = const_cast<uint32_t *>(pay.TID);

    // Arrival order: exactly n_no NewOrders among n_txns, each arrival being
    // a NewOrder with probability (NewOrders left) / (arrivals left)
    // (selection sampling), so every interleaving is equally likely.
    // slot[i] = q (NewOrder) or n_txns + p (Payment) of arrival i.
    std::vector<uint32_t> slot(n_txns);
    {
        uint32_t left_no = n_no, q = 0, p = 0;
        for (uint32_t i = 0; i < n_txns; i++) {
            const bool is_no = below(draw(seed, g->next_txn + i, F_TYPE, 0), n_txns - i) < left_no;
            if (is_no) left_no--;
            slot[i] = is_no ? q++ : n_txns + p++;
        }
    }
    // Every field but the per-warehouse sequence numbers is a pure function
    // of (seed, arrival): drawn in parallel.
    auto fields = [&](uint32_t i0, uint32_t i1) {
        for (uint32_t i = i0; i < i1; i++) {
            const uint64_t t = g->next_txn + i;
            if (slot[i] < n_txns) {
                const uint32_t q = slot[i];
                const uint32_t w = uniform(draw(seed, t, F_W, 0), 0, W - 1);
                NW[q] = w;
                ND[q] = uniform(draw(seed, t, F_D, 0), 0, 9);
                NC[q] = uniform(draw(seed, t, F_C, 0), 0, 2999);
                const uint32_t cnt = uniform(draw(seed, t, F_CNT, 0), c.min_lines, c.max_lines);
                NCNT[q] = cnt;
                for (uint32_t l = 0; l < LTPG_TPCC_MAX_LINES; l++) {
                    const uint64_t at = (uint64_t)l * no_cap + q;
                    if (l >= cnt) {
                        NI[at] = NSW[at] = NQ[at] = 0;
                        continue;
                    }
                    // distinct items within the order (Query.h:103-113): redraw
                    // with the next attempt number until unused
                    uint32_t item = 0;
                    for (uint32_t k = 0;; k++) {
                        if (c.nurand_items) {
                            const uint32_t a = uniform(draw(seed, t, F_ITEM_A, l * 4096 + k), 0, 8191);
                            const uint32_t b = uniform(draw(seed, t, F_ITEM_B, l * 4096 + k), 1, 100000);
                            item = (a | b) % 100000;  // NURand(8191, 1, 100000) - 1
                        } else {
                            item = uniform(draw(seed, t, F_ITEM_A, l * 4096 + k), 0, 99999);
                        }
                        bool dup = false;
                        for (uint32_t j = 0; j < l; j++) dup |= NI[(uint64_t)j * no_cap + q] == item;
                        if (!dup) break;
                    }
                    NI[at] = item;
                    uint32_t sw = w;
                    if (W > 1 && below(draw(seed, t, F_REMOTE, l), 100) < c.remote_supply_pct)
                        sw = other_warehouse(draw(seed, t, F_RW, l), w, W);
                    NSW[at] = sw;
                    NQ[at] = uniform(draw(seed, t, F_QTY, l), 1, c.max_quantity);
                }
                NT[q] = tid_base + i;
            } else {
                const uint32_t p = slot[i] - n_txns;
                const uint32_t w = uniform(draw(seed, t, F_W, 0), 0, W - 1);
                PW[p] = w;
                PD[p] = uniform(draw(seed, t, F_D, 0), 0, 9);
                if (below(draw(seed, t, F_ISNAME, 0), 100) < c.byname_pct) {
                    PN[p] = 1;
                    PC[p] = 0;
                    PL[p] = (uniform(draw(seed, t, F_NAME, 0), 0, 19) << 8) + uniform(draw(seed, t, F_ORD, 0), 0, 149);
                } else {
                    PN[p] = 0;
                    PC[p] = uniform(draw(seed, t, F_C, 0), 0, 2999);
                    PL[p] = 20;
                }
                uint32_t cw = w;
                if (W > 1 && below(draw(seed, t, F_CREMOTE, 0), 100) < c.remote_customer_pct)
                    cw = other_warehouse(draw(seed, t, F_CW, 0), w, W);
                PCW[p] = cw;
                PCD[p] = uniform(draw(seed, t, F_CD, 0), 0, 9);
                PA[p] = uniform(draw(seed, t, F_AMOUNT, 0), 1, c.max_amount);
                PT[p] = tid_base + i;
            }
        }
    };
    const uint32_t nt = n_txns < 4096 ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (nt == 1) {
        fields(0, n_txns);
    } else {
        std::vector<std::thread> th;
        const uint32_t per = (n_txns + nt - 1) / nt;
        for (uint32_t k = 0; k < nt; k++) {
            const uint32_t i0 = std::min(n_txns, k * per), i1 = std::min(n_txns, i0 + per);
            th.emplace_back(fields, i0, i1);
        }
        for (auto &x : th) x.join();
    }
    // Per-warehouse sequence numbers in arrival order, wrapping as
    // Query.h:54-88,157-166 do; an order's lines stay inside its
    // warehouse's ORDERLINE range.
    for (uint32_t i = 0; i < n_txns; i++) {
        if (slot[i] < n_txns) {
            const uint32_t q = slot[i], w = NW[q], cnt = NCNT[q];
            NNO[q] = g->noid[w];
            g->noid[w] = g->noid[w] + 1 >= kOrdersPerWarehouse ? 0 : g->noid[w] + 1;
            NOI[q] = g->oid[w];
            g->oid[w] = g->oid[w] + 1 >= kOrdersPerWarehouse ? 0 : g->oid[w] + 1;
            if (g->oolid[w] + cnt > kLinesPerWarehouse) g->oolid[w] = 0;
            NOL[q] = g->oolid[w];
            g->oolid[w] += cnt;
            if (g->oolid[w] >= kLinesPerWarehouse) g->oolid[w] = 0;
        } else {
            const uint32_t p = slot[i] - n_txns, w = PW[p];
            PH[p] = g->hid[w];
            g->hid[w] = g->hid[w] + 1 >= kHistoryPerWarehouse ? 0 : g->hid[w] + 1;
        }
    }
    g->next_txn += n_txns;
    if (n_neworder) *n_neworder = n_no;
    if (n_payment) *n_payment = n_pay;
    return LTPG_OK;
}

}  // extern "C"
