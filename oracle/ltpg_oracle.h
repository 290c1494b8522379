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
    uint32_t *read_values;     /* [S*B*field_words] or NULL; only read slots
                                  of committed transactions are written    */
    uint32_t commit_count;
    uint32_t range_errors;     /* ops whose key is >= rows (skipped)          */
} oracle_kv_result;

/*
 * One key-value epoch over a row-major table (rows x fields x field_words
 * u32), updated in place.  Batch is slot-major: slot s of txn t at
 * keys/meta[s*B + t].  values: 0 = copy-back, 1 = RMW (+ (TID|1) per word).
 * reg_scratch must hold 2*rows u64 (R and W registers of priority pairs
 * TID << 32 | batch position).
 */
int oracle_kv_epoch(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                    uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                    const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                    oracle_kv_result *res);

/* The four phases separately (used by tests that pin each one). */
void oracle_kv_register(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                        const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                        uint64_t *reg_r, uint64_t *reg_w);
void oracle_kv_check(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                     const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                     const uint64_t *reg_r, const uint64_t *reg_w, uint32_t *conflicts);
uint32_t oracle_commit(uint32_t B, const uint32_t *conflicts, uint64_t *bitmap);
void oracle_kv_writeback(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                         uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                         const uint32_t *tids, const uint64_t *bitmap, uint32_t values);

#ifdef __cplusplus
}
#endif
#endif
