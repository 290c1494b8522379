/*
 * ltpg_oracle.h - CPU restatement of LTPG's deterministic concurrency-control
 * rules.  TEST INFRASTRUCTURE ONLY: this is the checker that the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline
 * leg).  Nothing in the product (ltpg_amd/, libltpg) links or calls it.
 *
 * Single-threaded, plain C99, integer arithmetic only.  Every function cites
 * the reference lines it restates.  Semantics that the reference leaves
 * undefined or gets wrong are fixed as documented at each function; the
 * value-mode and rule-set constants are those of include/ltpg.h.
 */
#ifndef LTPG_ORACLE_H
#define LTPG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Result of one key-value epoch. */
typedef struct oracle_kv_result {
    uint64_t *commit_bitmap;   /* [ceil(B/64)]                                */
    uint32_t *conflicts;       /* [B] raw | war << 8 | waw << 16               */
    uint32_t *read_values;     /* [S*B*L*field_words], L = max(1, max_scan),
                                  or NULL; only read and scan slots of
                                  committed transactions are written       */
    uint32_t commit_count;
    uint32_t range_errors;     /* ops whose key is >= rows (skipped)          */
} oracle_kv_result;

/*
 * One key-value epoch over a row-major table (rows x fields x field_words
 * u32), updated in place.  Batch is slot-major: slot s of txn t at
 * keys/meta[s*B + t].  values: 0 = copy-back, 1 = RMW (+ (TID|1) per word).
 * reg_scratch must hold 2*rows u64 (R and W registers of priority pairs
 * TID << 32 | batch position).  hot_row: UINT64_MAX, or the row whose ops
 * never register or conflict and whose committed writes add commutatively.
 * max_scan: the longest scan accepted; read_values holds max(1, max_scan)
 * fields per op slot.
 */
int oracle_kv_epoch(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                    uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                    const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                    oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan);

/* The same epoch with the RMW delta of transaction t given as deltas[t]
 * instead of tids[t] | 1 (the served form's key-value requests: priorities
 * are ranks inside the epoch, deltas come from the 64-bit arrival number). */
int oracle_kv_epoch_deltas(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                           uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                           const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                           oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan,
                           const uint32_t *deltas);

/* The four phases separately (used by tests that pin each one). */
void oracle_kv_register(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                        const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                        uint64_t *reg_r, uint64_t *reg_w, uint64_t hot_row, uint32_t max_scan);
void oracle_kv_check(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                     const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                     const uint64_t *reg_r, const uint64_t *reg_w, uint32_t *conflicts, uint64_t hot_row,
                     uint32_t max_scan);
uint32_t oracle_commit(uint32_t B, const uint32_t *conflicts, uint64_t *bitmap);
void oracle_kv_writeback(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                         uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                         const uint32_t *tids, const uint64_t *bitmap, uint32_t values);

/* ---- order-entry (NewOrder + Payment) --------------------------------- */

/* Result of one order-entry epoch.  Bitmaps and counters are indexed within
 * each type; read values are value-major (value v of txn q at v*n + q):
 * 51 per NewOrder, 24 per Payment.  Pointers marked "or NULL" may be NULL. */
typedef struct oracle_tpcc_result {
    uint64_t *no_bitmap;      /* [ceil(n_no/64)]                              */
    uint64_t *pay_bitmap;     /* [ceil(n_pay/64)]                             */
    uint32_t *no_conflicts;   /* [n_no] raw | war << 8 | waw << 16            */
    uint32_t *pay_conflicts;  /* [n_pay]                                      */
    uint32_t *no_reads;       /* [51 * n_no] or NULL                          */
    uint32_t *pay_reads;      /* [24 * n_pay] or NULL                         */
    uint32_t *pay_cid;        /* [n_pay] or NULL: C_ID each Payment used      */
    uint32_t no_commits, pay_commits;
    uint32_t range_errors;    /* op slots whose row is outside its table      */
} oracle_tpcc_result;

/* Results of an epoch's read-only transactions (Order-Status, Stock-Level):
 * bitmaps, raw counters (war = waw = 0) and read values, value-major (3 per
 * Order-Status: CUSTOMER col 2, ORDER col 0, ORDERLINE col 0; 1 per
 * Stock-Level: STOCK col 2).  Read-value pointers may be NULL. */
typedef struct oracle_tpcc_ro_result {
    uint64_t *os_bitmap, *sl_bitmap;
    uint32_t *os_conflicts, *sl_conflicts;
    uint32_t *os_reads, *sl_reads;
    uint32_t os_commits, sl_commits;
} oracle_tpcc_ro_result;

/* Results of an epoch's Deliveries: bitmap, raw | war << 8 | waw << 16
 * counters and read values, value-major (2 per Delivery: ORDER col 3,
 * ORDERLINE col 3).  reads may be NULL. */
typedef struct oracle_tpcc_dl_result {
    uint64_t *bitmap;
    uint32_t *conflicts;
    uint32_t *reads;
    uint32_t commits;
} oracle_tpcc_dl_result;

typedef struct oracle_tpcc oracle_tpcc;

/* Rows and columns of table t (0..8: WAREHOUSE .. ITEM) at W warehouses. */
void oracle_tpcc_shape(uint32_t warehouses, uint32_t table, uint64_t *rows, uint32_t *cols);
/* Uniform draw in [a, b] of the loader for (seed, stream, row). */
uint32_t oracle_tpcc_pop_uniform(uint64_t seed, uint32_t stream, uint64_t row, uint32_t a, uint32_t b);
/* Fill the nine column-major tables with the loader's contents. */
void oracle_tpcc_populate(uint32_t warehouses, uint64_t seed, uint32_t *const tables[9]);
oracle_tpcc *oracle_tpcc_create(uint32_t warehouses);
void oracle_tpcc_destroy(oracle_tpcc *o);
/* One epoch over the column-major tables (updated in place).  Batches are
 * the reference's packed NEWORDER_QUERY / PAYMENT_QUERY blocks of capacity
 * no_cap / pay_cap (53 * no_cap and 10 * pay_cap u32).  rules: 0 reference,
 * 1 intended; values: 0 copy-back, 1 RMW. */
int oracle_tpcc_epoch(oracle_tpcc *o, uint32_t *const tables[9], uint32_t rules, uint32_t values,
                      const uint32_t *no_block, uint32_t no_cap, uint32_t n_no,
                      const uint32_t *pay_block, uint32_t pay_cap, uint32_t n_pay,
                      oracle_tpcc_result *res);
/* The same epoch with read-only transactions: Order-Status blocks of 6 x
 * os_cap u32 (W_ID, D_ID, C_ID, O_ID, OL_ID, TID) and Stock-Level blocks of
 * 3 x sl_cap u32 (W_ID, S_ID, TID). */
int oracle_tpcc_epoch_ro(oracle_tpcc *o, uint32_t *const tables[9], uint32_t rules, uint32_t values,
                         const uint32_t *no_block, uint32_t no_cap, uint32_t n_no,
                         const uint32_t *pay_block, uint32_t pay_cap, uint32_t n_pay,
                         const uint32_t *os_block, uint32_t os_cap, uint32_t n_os,
                         const uint32_t *sl_block, uint32_t sl_cap, uint32_t n_sl,
                         oracle_tpcc_result *res, oracle_tpcc_ro_result *ro);
/* The same epoch with Deliveries as well: blocks of 4 x dl_cap u32 (O_ID,
 * NO_W_ID, NO_O_ID, TID).  ro may be NULL when n_os = n_sl = 0, dl when
 * n_dl = 0. */
int oracle_tpcc_epoch_all(oracle_tpcc *o, uint32_t *const tables[9], uint32_t rules, uint32_t values,
                          const uint32_t *no_block, uint32_t no_cap, uint32_t n_no,
                          const uint32_t *pay_block, uint32_t pay_cap, uint32_t n_pay,
                          const uint32_t *os_block, uint32_t os_cap, uint32_t n_os,
                          const uint32_t *sl_block, uint32_t sl_cap, uint32_t n_sl,
                          const uint32_t *dl_block, uint32_t dl_cap, uint32_t n_dl,
                          oracle_tpcc_result *res, oracle_tpcc_ro_result *ro, oracle_tpcc_dl_result *dl);

#ifdef __cplusplus
}
#endif
#endif
