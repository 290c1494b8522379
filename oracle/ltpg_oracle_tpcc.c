/*
 * ltpg_oracle_tpcc.c - order-entry (NewOrder + Payment) rules, restated on
 * the CPU.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * Restates LTPG/TPCC-allopt-final: the op slots of Execute_neworder.h:
 * 14-423 and Execute_payment.h:39-387 as dispatched by their execute()
 * functions, the check slots as dispatched by check() (Execute_neworder.h:
 * 620-680, Execute_payment.h:547-595), __REGISTER / __CHECK
 * (Genericfunction.h:108-414), the commit rule (writeback_0,
 * Execute_neworder.h:681-697, Execute_payment.h:596-612), the by-name
 * customer lookup of KERNEL_CID (tp.cu:11-23, Execute_payment.h:13-36) and
 * the table contents of Database.h:104-327.  Tables are the reference's
 * column-major snapshot, tbl[col * rows + row].
 *
 * Every register of every slot is kept, including the ones no check ever
 * consults, exactly as the reference registers them.  Registers hold the
 * priority pair TID << 32 | batch position (NewOrder q -> q, Payment p ->
 * n_neworder + p); lower wins.
 *
 * Build-defined where the reference is silent or degenerate:
 *   - CID: the k-th (0-based, k = C_LAST & 0xff) customer, in row order, of
 *     district (C_W_ID * 10 + C_D_ID) whose C_LAST column equals
 *     C_LAST >> 8; its C_ID column.  No such customer, or no such district:
 *     C_ID 0.  (The reference's warp scan miscounts, Genericfunction.h:
 *     46-49.)
 *   - Rows are computed in 64 bits; an op slot whose row is outside its
 *     table is skipped by every phase and counted once in range_errors.
 *   - Values (the reference writes back the staged old value, so its tables
 *     never change):
 *       values == 0 (copy-back): no table changes.
 *       values == 1 (RMW), for committed transactions:
 *         NEWORDER row  <- (O_ID, D_ID, W_ID), written only by the
 *                          lowest-priority committed inserter of that row,
 *                          the last one in priority order (under the
 *                          reference rules two inserters of one row do not
 *                          conflict and may both commit);
 *         ORDER row     <- (O_ID, D_ID, W_ID);
 *         each line i in line order: STOCK (S_QUANTITY, S_YTD, S_ORDER_CNT)
 *           <- (q >= qty + 10 ? q - qty : q - qty + 91, S_YTD + qty,
 *               S_ORDER_CNT + 1); ORDERLINE row <- (O_ID, D_ID, W_ID, i,
 *               I_ID, SUPPLY_W_ID, 0, qty, qty * I_PRICE, I_DATA);
 *         CUSTOMER C_BALANCE -= H_AMOUNT, C_DATA += H_AMOUNT;
 *         HISTORY row   <- (C_ID, C_D_ID, C_W_ID, D_ID, W_ID);
 *         W_YTD and D_YTD += H_AMOUNT (the hot-row reduce; Execute_payment.h:
 *           96,161 stage it, Genericfunction.h:431-518 never applies it).
 *   - Read values: only committed transactions' reads are delivered, from
 *     the table before the epoch's write-back; a slot whose row is outside
 *     its table, or a line beyond O_OL_CNT, reads 0.
 *   - Read-only transactions (the Order-Status and Stock-Level slots of
 *     TPCC-alltxn/Execute.h:1592-1770,2155-2231; the reference never fills
 *     their query sets): Order-Status registers R on CUSTOMER row
 *     W*30000 + D*3000 + C, ORDER row W*30000 + O_ID and ORDERLINE row
 *     W*450000 + OL_ID and counts raw against W on each (check_0..check_2);
 *     Stock-Level registers R on STOCK row W*100000 + S_ID and counts raw
 *     against its W (check_0).  With no writes they commit (Aria's rule,
 *     war = waw = 0) and read CUSTOMER col 2, ORDER col 0, ORDERLINE col 0 /
 *     STOCK col 2 from the snapshot.  Batch positions follow the NewOrders
 *     and Payments: Order-Status o at n_no + n_pay + o, Stock-Level k at
 *     n_no + n_pay + n_os + k.
 *   - Deliveries (DELIVERYQUERY_NAMESPACE, TPCC-alltxn/Execute.h:1772-2153;
 *     the reference declares DELIVERY_SET but never fills it).  Each
 *     Delivery runs every op slot of its execute() chain (:1929-1977) and
 *     every check slot its check() chain reaches (:2108-2149): the chain
 *     tests wID % CHECK_WARP == 5 and == 6 twice, so check_4_0 and
 *     check_5_0 never run.  (The reference also derives the transaction from
 *     wID / 12 and wID / 8 but the slot from wID % EXECUTE_WARP and
 *     wID % CHECK_WARP, so on its grid most Deliveries would reach only some
 *     slots; that is launch geometry of an unfed stub, not a rule, and is
 *     not reproduced -- as for Order-Status and Stock-Level.)  Rows keep the
 *     reference's formulas, which omit the warehouse offset except for
 *     NEWORDER: op_0_0 W on NEWORDER row NO_W_ID * 30000 + NO_O_ID; op_1_0
 *     R and op_2_0 W on ORDER row O_ID; op_3_0 W and op_4_0 R on ORDERLINE
 *     row O_ID; op_5_0 W on CUSTOMER row O_ID.  Checks: check_0_0 raw
 *     (NEWORDER R), check_0_1 raw (NEWORDER W), check_1_0 war (ORDER W),
 *     check_2_0 raw (ORDER R), check_2_1 waw (ORDER W), check_3_0 raw
 *     (ORDERLINE R), check_3_1 waw (ORDERLINE W), check_5_1 waw (CUSTOMER
 *     W).  The reference has no commit rule or write-back for them: Aria's
 *     rule as for the other types; committed reads ORDER col 3 (op_1_1),
 *     ORDERLINE col 3 (op_4_1); with values == 1, NEWORDER row <- (0, 0, 0)
 *     when the Delivery is the row's last committed writer in priority order
 *     (NewOrder inserters included), ORDER col 5 <- TID % 10 + 1, ORDERLINE
 *     col 6 <- TID, CUSTOMER col 5 += ORDERLINE col 8 and col 8 += 1.
 *     Positions: Delivery d at n_no + n_pay + n_os + n_sl + d.
 */
#include "ltpg_oracle.h"

#include <stdlib.h>
#include <string.h>

enum { WH = 0, DI = 1, CU = 2, HI = 3, NO = 4, OR = 5, OL = 6, ST = 7, IT = 8, NT = 9 };
enum { R = 0, W = 1 };

static const uint32_t kCols[NT] = {9, 11, 21, 5, 3, 8, 10, 17, 5};

void oracle_tpcc_shape(uint32_t warehouses, uint32_t table, uint64_t *rows, uint32_t *cols)
{
    /* TPCC-allopt-final/Predefine.h:48-65 */
    static const uint64_t per_w[NT] = {1, 10, 30000, 30000, 30000, 30000, 450000, 100000, 0};
    uint64_t r = table == IT ? 100000 : per_w[table % NT] * warehouses;
    if (rows) *rows = table < NT ? r : 0;
    if (cols) *cols = table < NT ? kCols[table] : 0;
}

/* ---- loader (Database.h:104-327) --------------------------------------- */

static uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* Uniform draw in [a, b] for (seed, stream, row): the reference's
 * random.uniform_dist(a, b) at that row (Database.h:169,224,247,311-312). */
uint32_t oracle_tpcc_pop_uniform(uint64_t seed, uint32_t stream, uint64_t row, uint32_t a, uint32_t b)
{
    uint64_t h = splitmix64(splitmix64(seed ^ 0x504F50554C415445ull) ^ ((uint64_t)stream << 48) ^ row);
    return a + (uint32_t)(((unsigned __int128)h * ((uint64_t)b - a + 1)) >> 64);
}

void oracle_tpcc_populate(uint32_t Wn, uint64_t seed, uint32_t *const tbl[NT])
{
    uint64_t rows[NT];
    for (uint32_t t = 0; t < NT; t++) oracle_tpcc_shape(Wn, t, &rows[t], NULL);
#define SET(t, c, r, v) tbl[t][(uint64_t)(c) * rows[t] + (r)] = (uint32_t)(v)
    for (uint32_t t = 0; t < NT; t++) memset(tbl[t], 0, rows[t] * kCols[t] * sizeof(uint32_t));
    for (uint64_t i = 0; i < rows[WH]; i++) { /* create_warehouse :104-129 */
        SET(WH, 0, i, i); SET(WH, 1, i, 1); SET(WH, 2, i, 1); SET(WH, 6, i, 10);
    }
    for (uint64_t i = 0; i < rows[DI]; i++) { /* create_district :131-158 */
        SET(DI, 0, i, i % 10); SET(DI, 1, i, i / 10); SET(DI, 2, i, 1); SET(DI, 3, i, 1);
    }
    for (uint64_t i = 0; i < rows[CU]; i++) { /* create_customer :160-210 */
        SET(CU, 0, i, i % 3000); SET(CU, 1, i, i / 30000); SET(CU, 2, i, (i / 3000) % 10);
        for (uint32_t c = 3; c <= 9; c++) SET(CU, c, i, 100000);
        SET(CU, 10, i, oracle_tpcc_pop_uniform(seed, CU, i, 0, 1));
        SET(CU, 13, i, i % 20);
    }
    /* HISTORY and NEWORDER start empty (create_history / create_neworder
     * are empty, :211-216) */
    for (uint64_t i = 0; i < rows[OR]; i++) { /* create_order :217-237 */
        SET(OR, 0, i, i % 30000); SET(OR, 1, i, i / 3000); SET(OR, 2, i, i / 30000);
        SET(OR, 3, i, oracle_tpcc_pop_uniform(seed, OR, i, 0, 2999));
    }
    for (uint64_t i = 0; i < rows[OL]; i++) { /* create_orderline :238-263 */
        SET(OL, 0, i, i % 450000); SET(OL, 1, i, i / 45000); SET(OL, 2, i, i / 450000);
        SET(OL, 4, i, oracle_tpcc_pop_uniform(seed, OL, i, 0, 99999)); SET(OL, 5, i, i / 450000);
    }
    for (uint64_t i = 0; i < rows[ST]; i++) { /* create_stock :264-305 */
        SET(ST, 0, i, i % 100000); SET(ST, 1, i, i / 100000); SET(ST, 2, i, 100000); SET(ST, 3, i, 100000);
        for (uint32_t d = 1; d <= 10; d++) SET(ST, 5 + d, i, d);
    }
    for (uint64_t i = 0; i < rows[IT]; i++) { /* create_item :306-320 */
        SET(IT, 0, i, i); SET(IT, 1, i, 1000);
        SET(IT, 2, i, oracle_tpcc_pop_uniform(seed, IT, i, 1, 40000));
        SET(IT, 3, i, oracle_tpcc_pop_uniform(seed, IT + 1, i, 1, 10));
        SET(IT, 4, i, i);
    }
#undef SET
}

/* ---- epoch state --------------------------------------------------------- */

struct oracle_tpcc {
    uint32_t W;
    uint64_t rows[NT];
    uint64_t *reg[NT][2];      /* all-ones between epochs */
    uint64_t *touched;         /* (table * 2 + kind) << 56 | row, to reset */
    size_t n_touched, cap_touched;
    uint64_t *nwin;            /* NEWORDER rows: last committed inserter + 1, or 0 */
};

oracle_tpcc *oracle_tpcc_create(uint32_t warehouses)
{
    oracle_tpcc *o = (oracle_tpcc *)calloc(1, sizeof(*o));
    if (!o) return NULL;
    o->W = warehouses;
    for (uint32_t t = 0; t < NT; t++) {
        oracle_tpcc_shape(warehouses, t, &o->rows[t], NULL);
        for (int k = 0; k < 2; k++) {
            o->reg[t][k] = (uint64_t *)malloc(o->rows[t] * sizeof(uint64_t));
            if (!o->reg[t][k]) { oracle_tpcc_destroy(o); return NULL; }
            memset(o->reg[t][k], 0xff, o->rows[t] * sizeof(uint64_t));
        }
    }
    o->nwin = (uint64_t *)malloc(o->rows[NO] * sizeof(uint64_t));
    if (!o->nwin) { oracle_tpcc_destroy(o); return NULL; }
    memset(o->nwin, 0, o->rows[NO] * sizeof(uint64_t));
    return o;
}

void oracle_tpcc_destroy(oracle_tpcc *o)
{
    if (!o) return;
    for (uint32_t t = 0; t < NT; t++)
        for (int k = 0; k < 2; k++) free(o->reg[t][k]);
    free(o->nwin);
    free(o->touched);
    free(o);
}

static void touch(oracle_tpcc *o, uint32_t t, uint32_t k, uint64_t row)
{
    if (o->n_touched == o->cap_touched) {
        size_t cap = o->cap_touched ? 2 * o->cap_touched : 1u << 16;
        uint64_t *p = (uint64_t *)realloc(o->touched, cap * sizeof(uint64_t));
        if (!p) abort();
        o->touched = p;
        o->cap_touched = cap;
    }
    o->touched[o->n_touched++] = ((uint64_t)(t * 2 + k) << 56) | row;
}

/* __REGISTER (Genericfunction.h:108-251): atomicMin of the priority; the
 * isFrequent slots (Payment 1 and 3) register nothing (:122-134,149-161). */
static void reg_min(oracle_tpcc *o, uint32_t t, uint32_t k, uint64_t row, uint64_t me)
{
    uint64_t *r = &o->reg[t][k][row];
    if (*r == ~0ull) touch(o, t, k, row);
    if (me < *r) *r = me;
}

/* __CHECK (Genericfunction.h:252-414): 1 iff the register is held by a
 * higher-priority transaction. */
static uint32_t lower(const oracle_tpcc *o, uint32_t t, uint32_t k, uint64_t row, uint64_t me)
{
    return o->reg[t][k][row] < me ? 1u : 0u;
}

typedef struct {
    uint32_t n, cap;
    const uint32_t *W_ID, *D_ID, *C_ID, *O_ID, *N_O_ID, *O_OL_CNT, *O_OL_ID, *OL_I_ID, *OL_SUPPLY_W_ID,
        *OL_QUANTITY, *TID;
} nview;
typedef struct {
    uint32_t n;
    const uint32_t *W_ID, *D_ID, *C_ID, *C_LAST, *isName, *C_D_ID, *C_W_ID, *H_AMOUNT, *H_ID, *TID;
} pview;

static nview nview_of(const uint32_t *b, uint32_t cap, uint32_t n)
{
    nview v;
    const uint64_t C = cap;
    v.n = n; v.cap = cap;
    v.W_ID = b; v.D_ID = b + C; v.C_ID = b + 2 * C; v.O_ID = b + 3 * C; v.N_O_ID = b + 4 * C;
    v.O_OL_CNT = b + 5 * C; v.O_OL_ID = b + 6 * C; v.OL_I_ID = b + 7 * C; v.OL_SUPPLY_W_ID = b + 22 * C;
    v.OL_QUANTITY = b + 37 * C; v.TID = b + 52 * C;
    return v;
}

static pview pview_of(const uint32_t *b, uint32_t cap, uint32_t n)
{
    pview v;
    const uint64_t C = cap;
    v.n = n;
    v.W_ID = b; v.D_ID = b + C; v.C_ID = b + 2 * C; v.C_LAST = b + 3 * C; v.isName = b + 4 * C;
    v.C_D_ID = b + 5 * C; v.C_W_ID = b + 6 * C; v.H_AMOUNT = b + 7 * C; v.H_ID = b + 8 * C; v.TID = b + 9 * C;
    return v;
}

/* NewOrder op slots (Execute_neworder.h): rows of slots 0..4 and of line i's
 * ITEM / STOCK / ORDERLINE slots (5+i, 20+i, 35+i). */
typedef struct { uint64_t wh, di, cu, no, od; } nhead;

static nhead nrows(const nview *v, uint32_t q)
{
    nhead h;
    const uint64_t w = v->W_ID[q], d = v->D_ID[q];
    h.wh = w;                                   /* op_0_0 :22 */
    h.di = w * 10 + d;                          /* op_1_0 :50 */
    h.cu = w * 30000 + d * 3000 + v->C_ID[q];   /* op_2_0 :81 */
    h.no = w * 30000 + v->N_O_ID[q];            /* op_3_0 :111 */
    h.od = w * 30000 + v->O_ID[q];              /* op_4_0 :141 */
    return h;
}

static uint32_t nlines(const nview *v, uint32_t q)
{
    uint32_t c = v->O_OL_CNT[q];
    return c < 15 ? c : 15; /* roles exist for ID < 15 only (:338-423) */
}

static uint64_t item_row(const nview *v, uint32_t q, uint32_t i) { return v->OL_I_ID[(uint64_t)i * v->cap + q]; }
static uint64_t stock_row(const nview *v, uint32_t q, uint32_t i)
{
    return (uint64_t)v->OL_SUPPLY_W_ID[(uint64_t)i * v->cap + q] * 100000 + v->OL_I_ID[(uint64_t)i * v->cap + q];
}
static uint64_t line_row(const nview *v, uint32_t q, uint32_t i)
{
    return (uint64_t)v->W_ID[q] * 450000 + v->O_OL_ID[q] + i; /* op_7_0 :266 */
}

static int ok(const oracle_tpcc *o, uint32_t t, uint64_t row) { return row < o->rows[t]; }

static uint32_t get(uint32_t *const tbl[NT], const oracle_tpcc *o, uint32_t t, uint32_t col, uint64_t row)
{
    return tbl[t][(uint64_t)col * o->rows[t] + row];
}

static uint32_t *at(uint32_t *const tbl[NT], const oracle_tpcc *o, uint32_t t, uint32_t col, uint64_t row)
{
    return &tbl[t][(uint64_t)col * o->rows[t] + row];
}

/* KERNEL_CID / get_c_id (tp.cu:11-23, Execute_payment.h:13-36) with the
 * k-th-holder semantics stated above. */
static uint32_t find_cid(const oracle_tpcc *o, uint32_t *const tbl[NT], uint32_t cw, uint32_t cd, uint32_t c_last)
{
    const uint64_t d = (uint64_t)cw * 10 + cd;
    if (d >= o->rows[DI]) return 0;
    const uint32_t name = c_last >> 8, k = c_last & 0xffu;
    uint32_t seen = 0;
    for (uint64_t r = d * 3000; r < d * 3000 + 3000; r++) {
        if (get(tbl, o, CU, 13, r) != name) continue;
        if (seen++ == k) return get(tbl, o, CU, 0, r);
    }
    return 0;
}

int oracle_tpcc_epoch(oracle_tpcc *o, uint32_t *const tbl[NT], uint32_t rules, uint32_t values,
                      const uint32_t *no_block, uint32_t no_cap, uint32_t n_no, const uint32_t *pay_block,
                      uint32_t pay_cap, uint32_t n_pay, oracle_tpcc_result *res)
{
    return oracle_tpcc_epoch_ro(o, tbl, rules, values, no_block, no_cap, n_no, pay_block, pay_cap, n_pay, NULL, 0,
                                0, NULL, 0, 0, res, NULL);
}

/* Read-only transactions' rows: Order-Status op_0_0 / op_1_0 / op_2_0
 * (TPCC-alltxn/Execute.h:1594-1660), Stock-Level op_0_0 (:2157-2171). */
static void os_rows(const uint32_t *b, uint32_t cap, uint32_t o, uint64_t r[3])
{
    const uint64_t C = cap, w = b[o];
    r[0] = w * 30000 + (uint64_t)b[C + o] * 3000 + b[2 * C + o]; /* CUSTOMER  */
    r[1] = w * 30000 + b[3 * C + o];                             /* ORDER     */
    r[2] = w * 450000 + b[4 * C + o];                            /* ORDERLINE */
}

static uint64_t sl_row(const uint32_t *b, uint32_t cap, uint32_t k)
{
    return (uint64_t)b[k] * 100000 + b[(uint64_t)cap + k];
}

int oracle_tpcc_epoch_ro(oracle_tpcc *o, uint32_t *const tbl[NT], uint32_t rules, uint32_t values,
                         const uint32_t *no_block, uint32_t no_cap, uint32_t n_no, const uint32_t *pay_block,
                         uint32_t pay_cap, uint32_t n_pay, const uint32_t *os_block, uint32_t os_cap,
                         uint32_t n_os, const uint32_t *sl_block, uint32_t sl_cap, uint32_t n_sl,
                         oracle_tpcc_result *res, oracle_tpcc_ro_result *ro)
{
    return oracle_tpcc_epoch_all(o, tbl, rules, values, no_block, no_cap, n_no, pay_block, pay_cap, n_pay, os_block,
                                 os_cap, n_os, sl_block, sl_cap, n_sl, NULL, 0, 0, res, ro, NULL);
}

/* Delivery d's rows: NEWORDER (op_0_0, TPCC-alltxn/Execute.h:1780-1784) and
 * the O_ID row of ORDER / ORDERLINE / CUSTOMER (op_1_0 .. op_5_0). */
static uint64_t dl_new_row(const uint32_t *b, uint32_t cap, uint32_t d)
{
    return (uint64_t)b[(uint64_t)cap + d] * 30000 + b[2 * (uint64_t)cap + d];
}

int oracle_tpcc_epoch_all(oracle_tpcc *o, uint32_t *const tbl[NT], uint32_t rules, uint32_t values,
                          const uint32_t *no_block, uint32_t no_cap, uint32_t n_no, const uint32_t *pay_block,
                          uint32_t pay_cap, uint32_t n_pay, const uint32_t *os_block, uint32_t os_cap,
                          uint32_t n_os, const uint32_t *sl_block, uint32_t sl_cap, uint32_t n_sl,
                          const uint32_t *dl_block, uint32_t dl_cap, uint32_t n_dl, oracle_tpcc_result *res,
                          oracle_tpcc_ro_result *ro, oracle_tpcc_dl_result *dl)
{
    static const uint32_t os_tables[3] = {CU, OR, OL};
    static const uint32_t os_cols[3] = {2, 0, 0};
    const nview nv = nview_of(no_block, no_cap, n_no);
    const pview pv = pview_of(pay_block, pay_cap, n_pay);
    const uint64_t os_base = (uint64_t)n_no + n_pay, sl_base = os_base + n_os, dl_base = sl_base + n_sl;
    if (n_dl && !dl) return -1;
    uint32_t errors = 0;
    uint32_t *cid = res->pay_cid ? res->pay_cid : (uint32_t *)malloc(((size_t)n_pay + 1) * sizeof(uint32_t));
    if (!cid) return -1;

    /* CID phase */
    for (uint32_t p = 0; p < n_pay; p++)
        cid[p] = pv.isName[p] ? find_cid(o, tbl, pv.C_W_ID[p], pv.C_D_ID[p], pv.C_LAST[p]) : pv.C_ID[p];

    /* EXECUTE: registration (NEWORDER_NAMESPACE::execute :338-423,
     * PAYMENT_NAMESPACE::execute :304-386) */
    for (uint32_t q = 0; q < n_no; q++) {
        const uint64_t me = ((uint64_t)nv.TID[q] << 32) | q;
        const nhead h = nrows(&nv, q);
        if (ok(o, WH, h.wh)) reg_min(o, WH, R, h.wh, me); else errors++;       /* op_0_0 */
        if (ok(o, DI, h.di)) reg_min(o, DI, R, h.di, me); else errors++;       /* op_1_0 */
        if (ok(o, CU, h.cu)) reg_min(o, CU, R, h.cu, me); else errors++;       /* op_2_0 */
        if (ok(o, NO, h.no)) reg_min(o, NO, rules == 0 ? R : W, h.no, me);    /* op_3_0 */
        else errors++;
        if (ok(o, OR, h.od)) reg_min(o, OR, W, h.od, me); else errors++;       /* op_4_0 */
        for (uint32_t i = 0; i < nlines(&nv, q); i++) {
            const uint64_t ir = item_row(&nv, q, i), sr = stock_row(&nv, q, i), lr = line_row(&nv, q, i);
            if (ok(o, IT, ir)) reg_min(o, IT, R, ir, me); else errors++;       /* op_5_0 */
            if (ok(o, ST, sr)) reg_min(o, ST, W, sr, me); else errors++;       /* op_6_0 */
            if (ok(o, OL, lr)) reg_min(o, OL, W, lr, me); else errors++;       /* op_7_0 */
        }
    }
    for (uint32_t p = 0; p < n_pay; p++) {
        const uint64_t me = ((uint64_t)pv.TID[p] << 32) | (n_no + p);
        const uint64_t wh = pv.W_ID[p], di = (uint64_t)pv.W_ID[p] * 10 + pv.D_ID[p];
        const uint64_t cu = (uint64_t)pv.C_W_ID[p] * 30000 + (uint64_t)pv.C_D_ID[p] * 3000 + cid[p];
        const uint64_t hi = (uint64_t)pv.W_ID[p] * 30000 + pv.H_ID[p];
        if (ok(o, WH, wh)) reg_min(o, WH, R, wh, me); else errors++;           /* op_0_0 */
        if (!ok(o, WH, wh)) errors++;                                          /* op_1_0: isFrequent */
        if (ok(o, DI, di)) reg_min(o, DI, R, di, me); else errors++;           /* op_2_0 */
        if (!ok(o, DI, di)) errors++;                                          /* op_3_0: isFrequent */
        if (ok(o, CU, cu)) reg_min(o, CU, R, cu, me); else errors++;           /* op_4_0 */
        if (ok(o, CU, cu)) reg_min(o, CU, W, cu, me); else errors++;           /* op_5_0 */
        if (ok(o, HI, hi)) reg_min(o, HI, W, hi, me); else errors++;           /* op_6_0 */
    }
    for (uint32_t i = 0; i < n_os; i++) {  /* Order-Status op_0_0, op_1_0, op_2_0 */
        const uint64_t me = ((uint64_t)os_block[5 * (uint64_t)os_cap + i] << 32) | (os_base + i);
        uint64_t r[3];
        os_rows(os_block, os_cap, i, r);
        for (int k = 0; k < 3; k++) {
            if (ok(o, os_tables[k], r[k])) reg_min(o, os_tables[k], R, r[k], me); else errors++;
        }
    }
    for (uint32_t i = 0; i < n_sl; i++) {  /* Stock-Level op_0_0 */
        const uint64_t me = ((uint64_t)sl_block[2 * (uint64_t)sl_cap + i] << 32) | (sl_base + i);
        const uint64_t st = sl_row(sl_block, sl_cap, i);
        if (ok(o, ST, st)) reg_min(o, ST, R, st, me); else errors++;
    }

    for (uint32_t d = 0; d < n_dl; d++) {  /* Delivery op_0_0 .. op_5_0 (:1774-1915) */
        const uint64_t me = ((uint64_t)dl_block[3 * (uint64_t)dl_cap + d] << 32) | (dl_base + d);
        const uint64_t nr = dl_new_row(dl_block, dl_cap, d), od = dl_block[d];
        if (ok(o, NO, nr)) reg_min(o, NO, W, nr, me); else errors++;           /* op_0_0 */
        if (ok(o, OR, od)) reg_min(o, OR, R, od, me); else errors++;           /* op_1_0 */
        if (ok(o, OR, od)) reg_min(o, OR, W, od, me); else errors++;           /* op_2_0 */
        if (ok(o, OL, od)) reg_min(o, OL, W, od, me); else errors++;           /* op_3_0 */
        if (ok(o, OL, od)) reg_min(o, OL, R, od, me); else errors++;           /* op_4_0 */
        if (ok(o, CU, od)) reg_min(o, CU, W, od, me); else errors++;           /* op_5_0 */
    }

    /* CHECK (check dispatch as it runs: NewOrder role 2 does nothing, role 3
     * runs check_2_0, so check_3_0 never runs, Execute_neworder.h:635-646;
     * INTENDED rules run it). */
    memset(res->no_bitmap, 0, ((size_t)n_no + 63) / 64 * 8);
    memset(res->pay_bitmap, 0, ((size_t)n_pay + 63) / 64 * 8);
    res->no_commits = res->pay_commits = 0;
    for (uint32_t q = 0; q < n_no; q++) {
        const uint64_t me = ((uint64_t)nv.TID[q] << 32) | q;
        const nhead h = nrows(&nv, q);
        uint32_t raw = 0, war = 0, waw = 0;
        if (ok(o, WH, h.wh)) raw += lower(o, WH, W, h.wh, me);                  /* check_0_0 */
        if (ok(o, DI, h.di)) raw += lower(o, DI, W, h.di, me);                  /* check_1_0 */
        if (ok(o, CU, h.cu)) raw += lower(o, CU, W, h.cu, me);                  /* check_2_0 */
        if (ok(o, NO, h.no)) {
            if (rules != 0) war += lower(o, NO, R, h.no, me);                   /* check_3_0 */
            waw += lower(o, NO, W, h.no, me);                                   /* check_3_1 */
        }
        if (ok(o, OR, h.od)) {
            war += lower(o, OR, R, h.od, me);                                   /* check_4_0 */
            waw += lower(o, OR, W, h.od, me);                                   /* check_4_1 */
        }
        for (uint32_t i = 0; i < nlines(&nv, q); i++) {
            const uint64_t ir = item_row(&nv, q, i), sr = stock_row(&nv, q, i), lr = line_row(&nv, q, i);
            if (ok(o, IT, ir)) raw += lower(o, IT, W, ir, me);                  /* check_5_0 */
            if (ok(o, ST, sr)) {
                war += lower(o, ST, R, sr, me);                                 /* check_6_0 */
                waw += lower(o, ST, W, sr, me);                                 /* check_6_1 */
            }
            if (ok(o, OL, lr)) {
                war += lower(o, OL, R, lr, me);                                 /* check_7_0 */
                waw += lower(o, OL, W, lr, me);                                 /* check_7_1 */
            }
        }
        res->no_conflicts[q] = (raw & 0xffu) | ((war & 0xffu) << 8) | ((waw & 0xffu) << 16);
        if (waw == 0 && (raw == 0 || war == 0)) {
            res->no_bitmap[q / 64] |= 1ull << (q % 64);
            res->no_commits++;
            if (ok(o, NO, h.no) && me + 1 > o->nwin[h.no]) o->nwin[h.no] = me + 1;
        }
    }
    for (uint32_t p = 0; p < n_pay; p++) {
        const uint64_t me = ((uint64_t)pv.TID[p] << 32) | (n_no + p);
        const uint64_t wh = pv.W_ID[p], di = (uint64_t)pv.W_ID[p] * 10 + pv.D_ID[p];
        const uint64_t cu = (uint64_t)pv.C_W_ID[p] * 30000 + (uint64_t)pv.C_D_ID[p] * 3000 + cid[p];
        const uint64_t hi = (uint64_t)pv.W_ID[p] * 30000 + pv.H_ID[p];
        uint32_t raw = 0, war = 0, waw = 0;
        if (ok(o, WH, wh)) raw += lower(o, WH, W, wh, me);                      /* check_0_0 */
        /* check_1_0 / check_1_1 / check_3_0 / check_3_1: isFrequent, no-op */
        if (ok(o, DI, di)) raw += lower(o, DI, W, di, me);                      /* check_2_0 */
        if (ok(o, CU, cu)) {
            raw += lower(o, CU, W, cu, me);                                     /* check_4_0 */
            war += lower(o, CU, R, cu, me);                                     /* check_5_0 */
            waw += lower(o, CU, W, cu, me);                                     /* check_5_1 */
        }
        if (ok(o, HI, hi)) {
            war += lower(o, HI, R, hi, me);                                     /* check_6_0 */
            waw += lower(o, HI, W, hi, me);                                     /* check_6_1 */
        }
        res->pay_conflicts[p] = (raw & 0xffu) | ((war & 0xffu) << 8) | ((waw & 0xffu) << 16);
        if (waw == 0 && (raw == 0 || war == 0)) {
            res->pay_bitmap[p / 64] |= 1ull << (p % 64);
            res->pay_commits++;
        }
    }

    /* Read-only transactions: raw against W on each row (check_0..check_2,
     * TPCC-alltxn/Execute.h:1710-1745; check_0, :2203-2214); they commit */
    if (ro) {
        memset(ro->os_bitmap, 0, ((size_t)n_os + 63) / 64 * 8);
        memset(ro->sl_bitmap, 0, ((size_t)n_sl + 63) / 64 * 8);
        ro->os_commits = ro->sl_commits = 0;
        for (uint32_t i = 0; i < n_os; i++) {
            const uint64_t me = ((uint64_t)os_block[5 * (uint64_t)os_cap + i] << 32) | (os_base + i);
            uint64_t r[3];
            os_rows(os_block, os_cap, i, r);
            uint32_t raw = 0;
            for (int k = 0; k < 3; k++)
                if (ok(o, os_tables[k], r[k])) raw += lower(o, os_tables[k], W, r[k], me);
            ro->os_conflicts[i] = raw & 0xffu;
            ro->os_bitmap[i / 64] |= 1ull << (i % 64); /* waw == 0 and war == 0 */
            ro->os_commits++;
            if (ro->os_reads)
                for (int k = 0; k < 3; k++)
                    ro->os_reads[(uint64_t)k * n_os + i] =
                        ok(o, os_tables[k], r[k]) ? get(tbl, o, os_tables[k], os_cols[k], r[k]) : 0;
        }
        for (uint32_t i = 0; i < n_sl; i++) {
            const uint64_t me = ((uint64_t)sl_block[2 * (uint64_t)sl_cap + i] << 32) | (sl_base + i);
            const uint64_t st = sl_row(sl_block, sl_cap, i);
            ro->sl_conflicts[i] = ok(o, ST, st) ? lower(o, ST, W, st, me) : 0;
            ro->sl_bitmap[i / 64] |= 1ull << (i % 64);
            ro->sl_commits++;
            if (ro->sl_reads) ro->sl_reads[i] = ok(o, ST, st) ? get(tbl, o, ST, 2, st) : 0;
        }
    }

    /* Deliveries' check slots as the chain reaches them (:2108-2149) */
    if (dl) {
        memset(dl->bitmap, 0, ((size_t)n_dl + 63) / 64 * 8);
        dl->commits = 0;
        for (uint32_t d = 0; d < n_dl; d++) {
            const uint64_t me = ((uint64_t)dl_block[3 * (uint64_t)dl_cap + d] << 32) | (dl_base + d);
            const uint64_t nr = dl_new_row(dl_block, dl_cap, d), od = dl_block[d];
            uint32_t raw = 0, war = 0, waw = 0;
            if (ok(o, NO, nr)) {
                raw += lower(o, NO, R, nr, me);                                 /* check_0_0 */
                raw += lower(o, NO, W, nr, me);                                 /* check_0_1 */
            }
            if (ok(o, OR, od)) {
                war += lower(o, OR, W, od, me);                                 /* check_1_0 */
                raw += lower(o, OR, R, od, me);                                 /* check_2_0 */
                waw += lower(o, OR, W, od, me);                                 /* check_2_1 */
            }
            if (ok(o, OL, od)) {
                raw += lower(o, OL, R, od, me);                                 /* check_3_0 */
                waw += lower(o, OL, W, od, me);                                 /* check_3_1 */
            }
            /* check_4_0 (war, ORDERLINE W) and check_5_0 (raw, CUSTOMER R)
             * sit behind the repeated == 5 / == 6 tests and never run */
            if (ok(o, CU, od)) waw += lower(o, CU, W, od, me);                  /* check_5_1 */
            dl->conflicts[d] = (raw & 0xffu) | ((war & 0xffu) << 8) | ((waw & 0xffu) << 16);
            if (waw == 0 && (raw == 0 || war == 0)) {
                dl->bitmap[d / 64] |= 1ull << (d % 64);
                dl->commits++;
                if (ok(o, NO, nr) && me + 1 > o->nwin[nr]) o->nwin[nr] = me + 1;
                if (dl->reads) {
                    dl->reads[d] = ok(o, OR, od) ? get(tbl, o, OR, 3, od) : 0;                  /* op_1_1 */
                    dl->reads[(uint64_t)n_dl + d] = ok(o, OL, od) ? get(tbl, o, OL, 3, od) : 0; /* op_4_1 */
                }
            }
        }
    }

    /* Read values of committed transactions, from the snapshot */
    if (res->no_reads) {
        for (uint32_t q = 0; q < n_no; q++) {
            if (!((res->no_bitmap[q / 64] >> (q % 64)) & 1ull)) continue;
            const nhead h = nrows(&nv, q);
            uint32_t v[51];
            memset(v, 0, sizeof(v));
            if (ok(o, WH, h.wh)) v[0] = get(tbl, o, WH, 1, h.wh);               /* op_0_1 :39 */
            if (ok(o, DI, h.di)) {                                              /* op_1_1 :68-69 */
                v[1] = get(tbl, o, DI, 2, h.di);
                v[2] = get(tbl, o, DI, 4, h.di);
            }
            if (ok(o, CU, h.cu)) {                                              /* op_2_1 :99-101 */
                v[3] = get(tbl, o, CU, 4, h.cu);
                v[4] = get(tbl, o, CU, 10, h.cu);
                v[5] = get(tbl, o, CU, 13, h.cu);
            }
            for (uint32_t i = 0; i < nlines(&nv, q); i++) {                     /* op_5_1 :200-202 */
                const uint64_t ir = item_row(&nv, q, i);
                if (!ok(o, IT, ir)) continue;
                for (uint32_t k = 0; k < 3; k++) v[6 + 3 * i + k] = get(tbl, o, IT, 2 + k, ir);
            }
            for (uint32_t k = 0; k < 51; k++) res->no_reads[(uint64_t)k * n_no + q] = v[k];
        }
    }
    if (res->pay_reads) {
        static const uint32_t cu_cols[12] = {3, 4, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20};
        for (uint32_t p = 0; p < n_pay; p++) {
            if (!((res->pay_bitmap[p / 64] >> (p % 64)) & 1ull)) continue;
            const uint64_t wh = pv.W_ID[p], di = (uint64_t)pv.W_ID[p] * 10 + pv.D_ID[p];
            const uint64_t cu = (uint64_t)pv.C_W_ID[p] * 30000 + (uint64_t)pv.C_D_ID[p] * 3000 + cid[p];
            uint32_t v[24];
            memset(v, 0, sizeof(v));
            if (ok(o, WH, wh))
                for (uint32_t k = 0; k < 6; k++) v[k] = get(tbl, o, WH, 3 + k, wh);        /* op_0_1/0_2 */
            if (ok(o, DI, di))
                for (uint32_t k = 0; k < 6; k++) v[6 + k] = get(tbl, o, DI, 5 + k, di);    /* op_2_1/2_2 */
            if (ok(o, CU, cu))
                for (uint32_t k = 0; k < 12; k++) v[12 + k] = get(tbl, o, CU, cu_cols[k], cu); /* op_4_1..4_4 */
            for (uint32_t k = 0; k < 24; k++) res->pay_reads[(uint64_t)k * n_pay + p] = v[k];
        }
    }

    /* WRITEBACK (write_back dispatch, Execute_neworder.h:698-850,
     * Execute_payment.h:613-736) with the value rules stated above */
    if (values != 0) {
        for (uint32_t q = 0; q < n_no; q++) {
            if (!((res->no_bitmap[q / 64] >> (q % 64)) & 1ull)) continue;
            const uint64_t me = ((uint64_t)nv.TID[q] << 32) | q;
            const nhead h = nrows(&nv, q);
            const uint32_t oid = nv.O_ID[q], did = nv.D_ID[q], wid = nv.W_ID[q];
            if (ok(o, NO, h.no) && o->nwin[h.no] == me + 1) {
                *at(tbl, o, NO, 0, h.no) = oid;
                *at(tbl, o, NO, 1, h.no) = did;
                *at(tbl, o, NO, 2, h.no) = wid;
            }
            if (ok(o, OR, h.od)) {
                *at(tbl, o, OR, 0, h.od) = oid;
                *at(tbl, o, OR, 1, h.od) = did;
                *at(tbl, o, OR, 2, h.od) = wid;
            }
            for (uint32_t i = 0; i < nlines(&nv, q); i++) {
                const uint64_t ir = item_row(&nv, q, i), sr = stock_row(&nv, q, i), lr = line_row(&nv, q, i);
                const uint32_t qty = nv.OL_QUANTITY[(uint64_t)i * nv.cap + q];
                if (ok(o, ST, sr)) {
                    uint32_t *sq = at(tbl, o, ST, 2, sr);
                    *sq = *sq >= qty + 10 ? *sq - qty : *sq - qty + 91;
                    *at(tbl, o, ST, 3, sr) += qty;
                    *at(tbl, o, ST, 4, sr) += 1;
                }
                if (ok(o, OL, lr)) {
                    const uint32_t price = ok(o, IT, ir) ? get(tbl, o, IT, 2, ir) : 0;
                    const uint32_t data = ok(o, IT, ir) ? get(tbl, o, IT, 3, ir) : 0;
                    const uint32_t vals[10] = {oid, did, wid, i, (uint32_t)ir,
                                               nv.OL_SUPPLY_W_ID[(uint64_t)i * nv.cap + q], 0, qty, qty * price, data};
                    for (uint32_t c = 0; c < 10; c++) *at(tbl, o, OL, c, lr) = vals[c];
                }
            }
        }
        for (uint32_t p = 0; p < n_pay; p++) {
            if (!((res->pay_bitmap[p / 64] >> (p % 64)) & 1ull)) continue;
            const uint64_t wh = pv.W_ID[p], di = (uint64_t)pv.W_ID[p] * 10 + pv.D_ID[p];
            const uint64_t cu = (uint64_t)pv.C_W_ID[p] * 30000 + (uint64_t)pv.C_D_ID[p] * 3000 + cid[p];
            const uint64_t hi = (uint64_t)pv.W_ID[p] * 30000 + pv.H_ID[p];
            const uint32_t amt = pv.H_AMOUNT[p];
            if (ok(o, CU, cu)) {
                *at(tbl, o, CU, 5, cu) -= amt;
                *at(tbl, o, CU, 9, cu) += amt;
            }
            if (ok(o, HI, hi)) {
                const uint32_t vals[5] = {cid[p], pv.C_D_ID[p], pv.C_W_ID[p], pv.D_ID[p], pv.W_ID[p]};
                for (uint32_t c = 0; c < 5; c++) *at(tbl, o, HI, c, hi) = vals[c];
            }
            if (ok(o, WH, wh)) *at(tbl, o, WH, 2, wh) += amt;
            if (ok(o, DI, di)) *at(tbl, o, DI, 3, di) += amt;
        }
        /* Deliveries: a committed Delivery is the only committed writer of
         * its ORDER, ORDERLINE and CUSTOMER rows (waw), so the ORDERLINE
         * amount it adds is the snapshot's */
        for (uint32_t d = 0; d < n_dl; d++) {
            if (!((dl->bitmap[d / 64] >> (d % 64)) & 1ull)) continue;
            const uint32_t tid = dl_block[3 * (uint64_t)dl_cap + d];
            const uint64_t me = ((uint64_t)tid << 32) | (dl_base + d);
            const uint64_t nr = dl_new_row(dl_block, dl_cap, d), od = dl_block[d];
            if (ok(o, NO, nr) && o->nwin[nr] == me + 1)
                for (uint32_t c = 0; c < 3; c++) *at(tbl, o, NO, c, nr) = 0;
            if (ok(o, OR, od)) *at(tbl, o, OR, 5, od) = tid % 10 + 1;
            if (ok(o, OL, od)) *at(tbl, o, OL, 6, od) = tid;
            if (ok(o, CU, od)) {
                *at(tbl, o, CU, 5, od) += ok(o, OL, od) ? get(tbl, o, OL, 8, od) : 0;
                *at(tbl, o, CU, 8, od) += 1;
            }
        }
    }

    /* clear_LOG (Database.h:78-102) for the registers this epoch touched */
    for (size_t k = 0; k < o->n_touched; k++) {
        const uint64_t e = o->touched[k];
        const uint32_t tk = (uint32_t)(e >> 56);
        o->reg[tk / 2][tk % 2][e & ((1ull << 56) - 1)] = ~0ull;
    }
    o->n_touched = 0;
    for (uint32_t q = 0; q < n_no; q++) {
        const nhead h = nrows(&nv, q);
        if (ok(o, NO, h.no)) o->nwin[h.no] = 0;
    }
    for (uint32_t d = 0; d < n_dl; d++) {
        const uint64_t nr = dl_new_row(dl_block, dl_cap, d);
        if (ok(o, NO, nr)) o->nwin[nr] = 0;
    }
    res->range_errors = errors;
    if (!res->pay_cid) free(cid);
    return 0;
}
