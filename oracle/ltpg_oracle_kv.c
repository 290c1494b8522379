/*
 * ltpg_oracle_kv.c - key-value rules, restated on the CPU.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked into the product).
 *
 * Restates LTPG/YCSB/0 (identical to YCSB/1): Execute.h:11-271 and
 * Genericfunction.h:31-158.  Per op slot s of transaction t
 * (slot-major index s*B + t, as the kernels index LOC_R[YCSB_A_SIZE*OP + NO],
 * Execute.h:20,47):
 *   read  : register R (read_0 -> __REGISTER, Execute.h:25;
 *           Genericfunction.h:71-85), stage the field (read_1 -> __READ,
 *           Execute.h:36; Genericfunction.h:31-45), check W -> raw
 *           (check_r, Execute.h:109-122; __CHECK, Genericfunction.h:101-144)
 *   write : register W (write_0, Execute.h:38-53), check R -> war and
 *           W -> waw (check_w_0 / check_w_1, Execute.h:123-150)
 *   commit: waw == 0 && (raw == 0 || war == 0)  (writeback_0, Execute.h:
 *           184-191)
 *   write-back of committed writes (writeback_w, Execute.h:193-206).
 *
 * Build-defined where the reference is degenerate (its writes store the
 * value they staged, so the table never changes):
 *   values == 0 (copy-back): write-back stores the snapshot value: no change.
 *   values == 1 (RMW): each u32 word of the written field becomes
 *     word + (TID | 1), applied per write op in slot order.
 * A slot whose meta kind is NONE does nothing.  A key >= rows is counted in
 * range_errors and the op does nothing.
 *
 * Hot row (the reference's ZIPFIAN hot key, row 0: skipped by __REGISTER
 * and __CHECK, Genericfunction.h:74-79,88-93,114-119,128-133, and by
 * writeback_w, Execute.h:202; its writes go to the reduce, Execute.h:63-73,
 * 207-224, Genericfunction.h:160-210, which never reaches the table).  Here
 * ops on hot_row never register and never conflict; a committed read of it
 * returns the snapshot; committed writes to it apply their RMW delta as a
 * commutative add (the same value rule as every other write).  hot_row =
 * UINT64_MAX: no hot row.
 *
 * Priority.  The reference compares TIDs only and relies on them being
 * distinct (TID = (epoch & 0xfff) << 20 | seq, TPCC-allopt-final/Query.h:
 * 500).  Here priority is the pair (TID, batch position) in lexicographic
 * order, which is the reference's order whenever TIDs are distinct and a
 * total order when they are not.  Registers hold the minimum pair.
 *
 * Read values.  Only committed transactions' reads are staged: an aborted
 * transaction's reads are discarded (it re-executes in a later epoch), so
 * the build does not produce them.  They come from the snapshot, i.e. the
 * table before this epoch's write-back.
 *
 * Scans (LTPG/YCSB/2: read_0 registers R on rows Loc .. Loc + SCAN_SIZE - 1
 * below the table size, Execute.h:21-32; check_r checks them, Execute.h:
 * 126-134; __READ stages them, Genericfunction.h:44-50).  Build-defined: a
 * SCAN op (kind 3, length (meta >> 24) + 1 <= max_scan) is a read of every
 * row key .. min(key + len, rows) - 1 of its field -- each row registers R
 * and counts raw against W -- and stages each row's field (the reference
 * checks row Loc + 1 for every i and keeps only the last staged row).  A
 * scan whose first row or field is outside the table, or whose length
 * exceeds max_scan, is a range error.  Read values of slot i hold L =
 * max(1, max_scan) fields: row j of a scan at ((i * L) + j) * field_words.
 */
#include "ltpg_oracle.h"

#include <stdlib.h>
#include <string.h>

#define KIND(m) ((m) & 3u)
#define FIELD(m) (((m) >> 8) & 0xffffu)
#define SCANLEN(m) (((m) >> 24) + 1u)
#define K_READ 1u
#define K_WRITE 2u
#define K_SCAN 3u

/* An op takes part in the epoch iff its kind is READ, WRITE or SCAN, its
 * (first) row and field exist and a scan is at most max_scan rows; anything
 * else is skipped by every phase (and counted as a range error unless the
 * slot is empty). */
static int op_valid(uint32_t m, uint64_t row, uint64_t rows, uint32_t fields, uint32_t max_scan)
{
    uint32_t k = KIND(m);
    if (k == K_SCAN && SCANLEN(m) > max_scan) return 0;
    return k != 0 && row < rows && FIELD(m) < fields;
}

/* Rows read by op m at row: 1 for a READ, the clipped range for a SCAN. */
static uint64_t rows_read(uint32_t m, uint64_t row, uint64_t rows)
{
    if (KIND(m) == K_READ) return 1;
    if (KIND(m) != K_SCAN) return 0;
    uint64_t n = SCANLEN(m);
    return row + n <= rows ? n : rows - row;
}

static uint64_t prio(const uint32_t *tids, uint32_t t) { return ((uint64_t)tids[t] << 32) | t; }

/* __REGISTER: atomicMin(&log->YCSB_{R,W}[Loc], TID) -> a plain minimum of
 * the priority pair. */
void oracle_kv_register(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                        const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                        uint64_t *reg_r, uint64_t *reg_w, uint64_t hot_row, uint32_t max_scan)
{
    for (uint32_t s = 0; s < S; s++) {
        for (uint32_t t = 0; t < B; t++) {
            uint64_t i = (uint64_t)s * B + t;
            uint64_t row = keys[i];
            if (!op_valid(meta[i], row, rows, fields, max_scan)) continue;
            if (KIND(meta[i]) == K_WRITE) {
                if (row != hot_row && prio(tids, t) < reg_w[row]) reg_w[row] = prio(tids, t);
                continue;
            }
            uint64_t n = rows_read(meta[i], row, rows);
            for (uint64_t j = 0; j < n; j++)
                if (row + j != hot_row && prio(tids, t) < reg_r[row + j]) reg_r[row + j] = prio(tids, t);
        }
    }
}

/* __CHECK: existed = register; if (TID > existed) mark[NO]++ (Generic-
 * function.h:140-143).  Reads consult W (raw); writes consult R (war) and
 * W (waw). */
void oracle_kv_check(uint64_t rows, uint32_t fields, uint32_t B, uint32_t S,
                     const uint64_t *keys, const uint32_t *meta, const uint32_t *tids,
                     const uint64_t *reg_r, const uint64_t *reg_w, uint32_t *conflicts, uint64_t hot_row,
                     uint32_t max_scan)
{
    for (uint32_t t = 0; t < B; t++) {
        uint32_t raw = 0, war = 0, waw = 0;
        const uint64_t me = prio(tids, t);
        for (uint32_t s = 0; s < S; s++) {
            uint64_t i = (uint64_t)s * B + t;
            uint64_t row = keys[i];
            if (!op_valid(meta[i], row, rows, fields, max_scan)) continue;
            if (KIND(meta[i]) != K_WRITE) {
                uint64_t n = rows_read(meta[i], row, rows);
                for (uint64_t j = 0; j < n; j++)
                    if (row + j != hot_row && me > reg_w[row + j]) raw++;
            } else if (row != hot_row) {
                if (me > reg_r[row]) war++;
                if (me > reg_w[row]) waw++;
            }
        }
        conflicts[t] = (raw & 0xffu) | ((war & 0xffu) << 8) | ((waw & 0xffu) << 16);
    }
}

/* writeback_0 (Execute.h:184-191): commit iff waw == 0 && (raw == 0 ||
 * war == 0); bit t of the bitmap, and the COMMIT count. */
uint32_t oracle_commit(uint32_t B, const uint32_t *conflicts, uint64_t *bitmap)
{
    uint32_t count = 0;
    memset(bitmap, 0, ((size_t)B + 63) / 64 * sizeof(uint64_t));
    for (uint32_t t = 0; t < B; t++) {
        uint32_t raw = conflicts[t] & 0xffu;
        uint32_t war = (conflicts[t] >> 8) & 0xffu;
        uint32_t waw = (conflicts[t] >> 16) & 0xffu;
        if (waw == 0 && (raw == 0 || war == 0)) {
            bitmap[t / 64] |= 1ull << (t % 64);
            count++;
        }
    }
    return count;
}

/* writeback_w / __WRITEBACK (Execute.h:193-206, Genericfunction.h:145-158):
 * committed transactions write their write slots, in slot order. */
static void writeback_deltas(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                             uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                             const uint32_t *tids, const uint64_t *bitmap, uint32_t values,
                             const uint32_t *deltas)
{
    if (values == 0) return; /* copy-back stores the unchanged snapshot value */
    for (uint32_t t = 0; t < B; t++) {
        if (!((bitmap[t / 64] >> (t % 64)) & 1ull)) continue;
        uint32_t delta = deltas ? deltas[t] : (tids[t] | 1u);
        for (uint32_t s = 0; s < S; s++) {
            uint64_t i = (uint64_t)s * B + t;
            uint32_t m = meta[i];
            uint64_t row = keys[i];
            if (KIND(m) != K_WRITE || !op_valid(m, row, rows, fields, 0)) continue;
            uint32_t *w = table + (row * fields + FIELD(m)) * (uint64_t)field_words;
            for (uint32_t j = 0; j < field_words; j++) w[j] += delta;
        }
    }
}

void oracle_kv_writeback(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                         uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                         const uint32_t *tids, const uint64_t *bitmap, uint32_t values)
{
    writeback_deltas(rows, fields, field_words, table, B, S, keys, meta, tids, bitmap, values, NULL);
}

static int kv_epoch(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                    uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                    const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                    oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan, const uint32_t *deltas);

int oracle_kv_epoch(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                    uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                    const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                    oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan)
{
    return kv_epoch(rows, fields, field_words, table, B, S, keys, meta, tids, values, reg_scratch, res,
                    hot_row, max_scan, NULL);
}

int oracle_kv_epoch_deltas(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                           uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                           const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                           oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan,
                           const uint32_t *deltas)
{
    return kv_epoch(rows, fields, field_words, table, B, S, keys, meta, tids, values, reg_scratch, res,
                    hot_row, max_scan, deltas);
}

static int kv_epoch(uint64_t rows, uint32_t fields, uint32_t field_words, uint32_t *table,
                    uint32_t B, uint32_t S, const uint64_t *keys, const uint32_t *meta,
                    const uint32_t *tids, uint32_t values, uint64_t *reg_scratch,
                    oracle_kv_result *res, uint64_t hot_row, uint32_t max_scan, const uint32_t *deltas)
{
    const uint64_t L = max_scan > 1 ? max_scan : 1;
    uint64_t *reg_r = reg_scratch, *reg_w = reg_scratch + rows;
    uint32_t errors = 0;
    /* clear_LOG: every register back to "no registrant" (YCSB/0/Database.h:
     * 45-50 writes 0xFFFFFFFF; here the all-ones priority pair) */
    memset(reg_scratch, 0xff, (size_t)rows * 2 * sizeof(uint64_t));

    /* execute: register */
    oracle_kv_register(rows, fields, B, S, keys, meta, tids, reg_r, reg_w, hot_row, max_scan);
    for (uint64_t i = 0; i < (uint64_t)S * B; i++)
        if (KIND(meta[i]) != 0 && !op_valid(meta[i], keys[i], rows, fields, max_scan)) errors++;

    oracle_kv_check(rows, fields, B, S, keys, meta, tids, reg_r, reg_w, res->conflicts, hot_row, max_scan);
    res->commit_count = oracle_commit(B, res->conflicts, res->commit_bitmap);

    /* committed reads see the snapshot: stage them before any write-back */
    if (res->read_values) {
        for (uint32_t s = 0; s < S; s++) {
            for (uint32_t t = 0; t < B; t++) {
                uint64_t i = (uint64_t)s * B + t;
                if (!((res->commit_bitmap[t / 64] >> (t % 64)) & 1ull)) continue;
                if (KIND(meta[i]) == K_WRITE || !op_valid(meta[i], keys[i], rows, fields, max_scan)) continue;
                uint64_t n = rows_read(meta[i], keys[i], rows);
                for (uint64_t j = 0; j < n; j++) {
                    const uint32_t *src =
                        table + ((keys[i] + j) * fields + FIELD(meta[i])) * (uint64_t)field_words;
                    memcpy(res->read_values + (i * L + j) * field_words, src, field_words * sizeof(uint32_t));
                }
            }
        }
    }
    writeback_deltas(rows, fields, field_words, table, B, S, keys, meta, tids, res->commit_bitmap, values,
                     deltas);
    res->range_errors = errors;
    return 0;
}
