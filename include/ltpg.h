/*
 * ltpg.h - C ABI of the MI355X deterministic concurrency-control path.
 *
 * The path is LTPG's large-batch deterministic concurrency control: a batch
 * (epoch) of transactions goes through four phases,
 *
 *   EXECUTE    each op registers its TID with an atomic minimum on a per-row
 *              register (R for reads, W for writes) and stages the columns it
 *              reads from the table snapshot;
 *   CHECK      each op reads back the opposite register(s) and counts raw /
 *              war / waw conflicts; the transaction commits iff
 *              waw == 0 && (raw == 0 || war == 0);
 *   WRITEBACK  committed transactions write their write set to the table;
 *   CID        (order-entry only) Payment by-name customer resolution.
 *
 * Every call here replaces a piece of the reference's implicit seam: the
 * __global__ kernels of LTPG/TPCC-allopt-final/tp.cu:11-117 and
 * LTPG/YCSB/0/ycsb.cu:9-58, the Database/Query lifecycle of
 * LTPG/TPCC-allopt-final/Database.h:37-102 and Query.h:330-383,532-631, and
 * the epoch loop of tp.cu:136-261.  Each declaration cites the part it
 * replaces.
 *
 * Conventions
 *   - Every function returns int: LTPG_OK (0) or a negative LTPG_ERR_* code.
 *     ltpg_last_error() returns a thread-local message for the last failure.
 *     No C++ exception crosses this boundary.
 *   - A context is single-threaded: the caller serialises calls on one
 *     context.  Separate contexts are independent.
 *   - The library owns every device allocation.  Batches given with
 *     LTPG_MEM_HOST are copied to the device before the call returns (the
 *     caller may reuse its buffers then), into one of two device slots on a
 *     copy stream, so the copy overlaps the previous epoch still running on
 *     the device; pinned host memory copies at full PCIe speed.  Batches
 *     given with LTPG_MEM_DEVICE are read in place and must stay valid until
 *     ltpg_sync() returns.
 *   - TIDs are u32 priorities; a lower TID wins.  The reference relies on
 *     TIDs being distinct within a batch (TID = (epoch & 0xfff) << 20 | seq,
 *     TPCC-allopt-final/Query.h:500); here equal TIDs are ordered by batch
 *     position, so priority is the pair (TID, position).
 */
#ifndef LTPG_H
#define LTPG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LTPG_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
#define LTPG_OK 0
#define LTPG_ERR_INVALID (-1)     /* bad argument or shape                   */
#define LTPG_ERR_NOMEM (-2)       /* device or host allocation failed        */
#define LTPG_ERR_DEVICE (-3)      /* HIP runtime error / no device           */
#define LTPG_ERR_STATE (-4)       /* call out of order (e.g. no batch)       */
#define LTPG_ERR_RANGE (-5)       /* a key / row id outside its table        */
#define LTPG_ERR_BUDGET (-6)      /* admission refused: HBM budget exceeded  */
#define LTPG_ERR_UNSUPPORTED (-7) /* feature not configured in this context  */

/* ---- rule sets (which register each op slot registers and checks) ------ */
/* REFERENCE: the slots that actually dispatch in the reference, including
 *   the NewOrder check dispatch that never runs check_3_0
 *   (TPCC-allopt-final/Execute_neworder.h:635-646) and the NEWORDER insert
 *   registering R (Execute_neworder.h:115).
 * INTENDED: the NEWORDER insert registers W and checks R->war and W->waw. */
#define LTPG_RULES_REFERENCE 0u
#define LTPG_RULES_INTENDED 1u

/* ---- value semantics of write-back ------------------------------------ */
/* COPYBACK: a committed write stores the snapshot value it staged, as the
 *   reference does (TPCC-allopt-final/Genericfunction.h:93-107,416-430):
 *   the table never changes.
 * RMW: key-value writes add (TID | 1) to every u32 word of their field;
 *   order-entry writes store the values of the order-entry rules (NewOrder
 *   inserts, stock update, Payment customer update, W_YTD/D_YTD reduce). */
#define LTPG_VALUES_COPYBACK 0u
#define LTPG_VALUES_RMW 1u

/* ---- memory kinds ------------------------------------------------------ */
#define LTPG_MEM_HOST 0u
#define LTPG_MEM_DEVICE 1u
/* OR'd into the memory kind of ltpg_kv_submit: the batch is request-major,
 * slot s of transaction t at [t * n_slots + s] (each transaction's slots one
 * contiguous run, as a request queue stages them).  The phases read it in
 * place; committed reads keep their slot-major positions.  No counterpart
 * in the reference, whose batches are slot-major only. */
#define LTPG_BATCH_REQUEST_MAJOR 0x10u

/* ---- table layouts ----------------------------------------------------- */
/* COLUMN_MAJOR is the reference's SNAPSHOT layout, tbl[col * rows + row]
 * (TPCC-allopt-final/Datastructure.h:119-130, YCSB/0/Datastructure.h:16-19).
 * ROW_MAJOR is the device layout, tbl[row * cols + col]. */
#define LTPG_LAYOUT_ROW_MAJOR 0u
#define LTPG_LAYOUT_COLUMN_MAJOR 1u

/* ---- workloads --------------------------------------------------------- */
#define LTPG_WORKLOAD_KV 1u
#define LTPG_WORKLOAD_TPCC 2u

/* ---- transaction types (result queries) --------------------------------- */
#define LTPG_TYPE_KV 0u
#define LTPG_TYPE_NEWORDER 1u
#define LTPG_TYPE_PAYMENT 2u
#define LTPG_TYPE_ORDERSTATUS 3u /* read-only order-entry types, see       */
#define LTPG_TYPE_STOCKLEVEL 4u  /* ltpg_tpcc_submit_readonly               */
#define LTPG_TYPE_DELIVERY 5u    /* ltpg_tpcc_submit_delivery               */

/* ---- phases ------------------------------------------------------------ */
#define LTPG_PHASE_CID 1u
#define LTPG_PHASE_EXECUTE 2u
#define LTPG_PHASE_CHECK 4u
#define LTPG_PHASE_WRITEBACK 8u /* committed reads are delivered, then writes applied */
#define LTPG_PHASE_ALL 15u

/* ---- key-value op encoding (meta word of one op slot) ------------------ */
/* bits 0-1: kind; bits 8-23: field index; bits 24-31: scan length - 1.  A
 * slot of kind NONE is empty (ragged transactions).  A READ registers R and
 * stages its field; a WRITE is a read-modify-write of its field that
 * registers W (YCSB/0/Execute.h:11-78).  A SCAN reads the field of rows
 * key .. key + len - 1, clipped at the end of the table: every row is a read
 * for registration and conflicts, and the rows' fields are staged in order
 * (the range reads of YCSB/2/Execute.h:21-32,126-134, whose check consults
 * row key + 1 for every row and whose staging keeps only the last row; here
 * each row is checked and delivered).  Scans need ltpg_config.kv_max_scan >=
 * len; a longer scan is a range error. */
#define LTPG_KV_NONE 0u
#define LTPG_KV_READ 1u
#define LTPG_KV_WRITE 2u
#define LTPG_KV_SCAN 3u
#define LTPG_KV_META(kind, field) ((uint32_t)(kind) | ((uint32_t)(field) << 8))
#define LTPG_KV_SCAN_META(field, len) (LTPG_KV_SCAN | ((uint32_t)(field) << 8) | (((uint32_t)(len) - 1u) << 24))
#define LTPG_KV_KIND(meta) ((meta) & 3u)
#define LTPG_KV_FIELD(meta) (((meta) >> 8) & 0xffffu)
#define LTPG_KV_SCAN_LEN(meta) (((meta) >> 24) + 1u)
#define LTPG_KV_MAX_SCAN 256u

#define LTPG_NO_HOT_ROW UINT64_MAX

/* ---- context flags ------------------------------------------------------ */
/* CONFLICT_COUNTS: CHECK counts raw / war / waw of every transaction (the
 * reference's per-transaction counters) for ltpg_conflict_counters().
 * Without it CHECK skips transactions already known to abort (waw found
 * during EXECUTE) and the counters are not produced. */
#define LTPG_FLAG_CONFLICT_COUNTS 1u

/* ---- order-entry schema (TPCC-allopt-final/Predefine.h:48-77) ---------- */
#define LTPG_TPCC_WAREHOUSE 0u
#define LTPG_TPCC_DISTRICT 1u
#define LTPG_TPCC_CUSTOMER 2u
#define LTPG_TPCC_HISTORY 3u
#define LTPG_TPCC_NEWORDER 4u
#define LTPG_TPCC_ORDER 5u
#define LTPG_TPCC_ORDERLINE 6u
#define LTPG_TPCC_STOCK 7u
#define LTPG_TPCC_ITEM 8u
#define LTPG_TPCC_NTABLES 9u
#define LTPG_TPCC_MAX_LINES 15u

/* ---- context configuration --------------------------------------------- */
typedef struct ltpg_config {
    uint32_t size;          /* sizeof(ltpg_config), set by ltpg_config_init   */
    int32_t device;         /* HIP device ordinal                              */
    uint32_t workload;      /* LTPG_WORKLOAD_KV or LTPG_WORKLOAD_TPCC          */
    uint32_t rules;         /* LTPG_RULES_*                                    */
    uint32_t values;        /* LTPG_VALUES_*                                   */
    uint32_t max_txns;      /* capacity: transactions per batch (<= 2^24)      */
    /* key-value table (YCSB/0/Predefine.h:9-10,21) */
    uint64_t kv_rows;
    uint32_t kv_fields;
    uint32_t kv_field_bytes; /* multiple of 4, at most 1024                    */
    uint32_t kv_max_slots;   /* capacity: op slots per transaction (<= 64)     */
    uint32_t flags;          /* LTPG_FLAG_*                                    */
    uint64_t kv_hot_row;     /* LTPG_NO_HOT_ROW, or a row whose writes go      */
                             /* through the commutative reduce                 */
    /* order-entry (TPCC-allopt-final/Predefine.h:9) */
    uint32_t tpcc_warehouses;
    uint32_t kv_max_scan;    /* longest key-value scan (0: no scans; <= 256);
                                each op slot's read values hold
                                max(1, kv_max_scan) fields                    */
    uint64_t hbm_budget;     /* serving: bytes of access-set buffers admitted  */
} ltpg_config;

/* Fill cfg with defaults (device 0, KV, reference rules, RMW values). */
void ltpg_config_init(ltpg_config *cfg);

/* ---- lifecycle (Database()/Query() ctors, TPCC-allopt-final/Database.h:37-57,
 *      Query.h:330-355; clear_LOG, Database.h:78-102) ----------------------- */
typedef struct ltpg_ctx ltpg_ctx;

int ltpg_create(const ltpg_config *cfg, ltpg_ctx **out);
int ltpg_destroy(ltpg_ctx *ctx);
const char *ltpg_last_error(void);
int ltpg_abi_version(void);
/* Number of HIP devices visible (0 when none).  Never initialises a context. */
int ltpg_device_count(void);

/* Use the caller's hipStream_t (NULL: the context's own stream). */
int ltpg_set_stream(ltpg_ctx *ctx, void *hip_stream);
/* Block until all work queued by this context has finished; returns any
 * deferred error (e.g. LTPG_ERR_RANGE raised by a kernel bounds check). */
int ltpg_sync(ltpg_ctx *ctx);

/* ---- device memory for LTPG_MEM_DEVICE batches (the reference's
 *      cudaMalloc'd query slots, TPCC-allopt-final/Query.h:344-351, and their
 *      H2D copies, Query.h:544-578) ---------------------------------------- */
int ltpg_device_alloc(ltpg_ctx *ctx, uint64_t bytes, void **dptr);
int ltpg_device_free(ltpg_ctx *ctx, void *dptr);
/* Synchronous host -> device / device -> host copies on the context stream. */
int ltpg_copy_to_device(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int ltpg_copy_to_host(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* Device -> device copy by a streaming kernel on the context stream
 * (asynchronous; bytes must be a multiple of 16 and both pointers 16-byte
 * aligned).  Used to snapshot tables and as the known-size copy that
 * calibrates HBM traffic counters. */
int ltpg_device_copy(ltpg_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* Random-access ceilings of this device (no counterpart in the reference:
 * the roofline the register phases are measured against).  Over a scratch
 * array of `bytes` (rounded down to 8) it times n_ops random 8-byte loads,
 * then n_ops random probe + 64-bit atomicMin pairs (the registration step
 * of EXECUTE), each spread over the whole array, and returns the rates in
 * G ops/s.  Synchronous; allocates and frees the array. */
int ltpg_random_access_rates(ltpg_ctx *ctx, uint64_t bytes, uint64_t n_ops, double *load_gops,
                             double *atomic_gops);

/* Random-gather ceiling (no counterpart in the reference: the roofline of
 * the read-out and write-back copies).  Over a scratch array of `bytes`
 * split into slots of stride_bytes it times n_ops copies of field_bytes
 * from random slots into a dense destination, 64 fields per wave step with
 * up to 32 in flight, and returns the rate in GB/s of field bytes read plus
 * written.  Synchronous; allocates and frees both arrays. */
int ltpg_random_gather_rate(ltpg_ctx *ctx, uint64_t bytes, uint32_t field_bytes, uint32_t stride_bytes,
                            uint64_t n_ops, double *gbs);

/* Costs of the ways a multi-phase epoch can be sequenced on this device (no
 * counterpart in the reference; the evidence behind the batched form's use of
 * graphs rather than one persistent kernel): out[0] = microseconds per
 * grid-wide barrier of `blocks` co-resident 256-thread workgroups (a
 * cooperative launch passing `reps` barriers), out[1] = microseconds per
 * kernel boundary inside a HIP graph of `reps` back-to-back `blocks`-wide
 * kernels, out[2] = the same launched one by one on the stream.  Synchronous;
 * LTPG_ERR_DEVICE if the workgroups cannot all be resident. */
int ltpg_sync_costs(ltpg_ctx *ctx, uint32_t blocks, uint32_t reps, double out[3]);

/* Latencies of the steps an epoch of the served key-value device loop
 * chains (no counterpart in the reference; the terms of that loop's
 * dependent-chain roofline), in nanoseconds per step, measured in one
 * 1024-thread workgroup: out_ns[0] an LDS-only workgroup barrier, [1] a
 * dependent LDS load, [2] a dependent LDS atomic with return (every lane
 * running such a chain), [3] a dependent load of pinned host memory (a PCIe
 * round trip), [4] a system-scope release fence plus a store to pinned host
 * memory, back to back.  Synchronous; allocates and frees its buffers. */
int ltpg_latency_costs(ltpg_ctx *ctx, double out_ns[5]);

/* ---- key-value tables (YCSB/0/Database.h:14-131) ------------------------ */
/* Load rows x kv_fields x kv_field_bytes.  COLUMN_MAJOR requires
 * kv_field_bytes == 4 (the reference's u32 snapshot). */
int ltpg_kv_table_load(ltpg_ctx *ctx, const void *host, uint64_t rows, uint32_t layout);
int ltpg_kv_table_read(ltpg_ctx *ctx, void *host, uint64_t rows, uint32_t layout);
/* Device-side fill with the reference's initial contents: every u32 word of
 * field c holds c (YCSB/0/Database.h:72-99). */
int ltpg_kv_table_fill_reference(ltpg_ctx *ctx);
/* Rows rows[0 .. n-1] of the table, all fields, packed into host
 * (n * kv_fields * kv_field_bytes bytes; the reference copies its whole
 * snapshot back, YCSB/0/Database.h:131; here a resident table of up to
 * hundreds of GB is read by row).  Synchronous; LTPG_ERR_RANGE for a row
 * outside the table. */
int ltpg_kv_table_read_rows(ltpg_ctx *ctx, const uint64_t *rows, uint64_t n, uint32_t *host);

/* ---- key-value batch (YCSB_A_QUERY, YCSB/0/Datastructure.h:36-41) ------- */
/* Slot-major, as the reference kernel indexes LOC[B * op + NO]
 * (YCSB/0/Execute.h:20,47): slot s of transaction t is at [s * n_txns + t]. */
typedef struct ltpg_kv_batch {
    uint32_t n_txns;         /* B (0: an empty epoch; the arrays may be NULL)  */
    uint32_t n_slots;        /* S (1..kv_max_slots when B > 0)                 */
    const uint64_t *keys;    /* [S * B] row ids                                */
    const uint32_t *meta;    /* [S * B] LTPG_KV_META(kind, field)              */
    const uint32_t *tids;    /* [B]                                            */
} ltpg_kv_batch;

/* Make `batch` the context's current epoch (copy_to_device,
 * YCSB/0/Query.h:197-216).  Starts a new epoch: registers of the previous
 * epoch become stale without a memset.  `mem`: LTPG_MEM_HOST or
 * LTPG_MEM_DEVICE, optionally | LTPG_BATCH_REQUEST_MAJOR. */
int ltpg_kv_submit(ltpg_ctx *ctx, const ltpg_kv_batch *batch, uint32_t mem);

/* ---- order-entry tables (SNAPSHOT, TPCC-allopt-final/Datastructure.h:119-130;
 *      sizes Predefine.h:48-65) --------------------------------------------- */
/* Rows and columns of table `table_id` (LTPG_TPCC_*) for the context's
 * warehouse count: W, 10W, 30000W, 30000W, 30000W, 30000W, 450000W,
 * 100000W, 100000 rows of 9, 11, 21, 5, 3, 8, 10, 17, 5 u32 columns. */
int ltpg_tpcc_table_shape(ltpg_ctx *ctx, uint32_t table_id, uint64_t *rows, uint32_t *cols);
/* Load / read one table; COLUMN_MAJOR is the reference's snapshot layout
 * (tbl[col * rows + row]).  Loading the customer table rebuilds the
 * by-name index (C_LAST, column 13) used by the CID phase. */
int ltpg_tpcc_table_load(ltpg_ctx *ctx, uint32_t table_id, const uint32_t *host, uint64_t rows,
                         uint32_t layout);
int ltpg_tpcc_table_read(ltpg_ctx *ctx, uint32_t table_id, uint32_t *host, uint64_t rows,
                         uint32_t layout);
/* Fill every table on the device with the reference loader's contents
 * (create_warehouse .. create_item, TPCC-allopt-final/Database.h:104-327);
 * its uniform draws (C_CREDIT, O_C_ID, OL_I_ID, I_PRICE, I_DATA) come from a
 * counter-based hash of (seed, table, row).  Rebuilds the by-name index. */
int ltpg_tpcc_populate(ltpg_ctx *ctx, uint64_t seed);

/* ---- order-entry batches (NEWORDER_QUERY / PAYMENT_QUERY,
 *      TPCC-allopt-final/Datastructure.h:211-249) ---------------------------- */
typedef struct ltpg_neworder_batch {
    uint32_t n;               /* transactions                                  */
    uint32_t line_stride;     /* >= n: line l of txn q is [l * line_stride + q] */
    const uint32_t *W_ID, *D_ID, *C_ID, *O_ID, *N_O_ID, *O_OL_CNT, *O_OL_ID; /* [n] */
    const uint32_t *OL_I_ID, *OL_SUPPLY_W_ID, *OL_QUANTITY; /* [15 * line_stride] */
    const uint32_t *TID;      /* [n]                                           */
} ltpg_neworder_batch;

typedef struct ltpg_payment_batch {
    uint32_t n;
    uint32_t reserved;
    const uint32_t *W_ID, *D_ID, *C_ID, *C_LAST, *isName, *C_D_ID, *C_W_ID, *H_AMOUNT, *H_ID;
    const uint32_t *TID;      /* all [n]                                       */
} ltpg_payment_batch;

/* Views of the reference's packed structs: a NEWORDER_QUERY / PAYMENT_QUERY
 * whose arrays hold `cap` entries each (NEWORDER_CNT / PAYMENT_CNT), of which
 * the first n are used.  The packed NewOrder block is 53 * cap u32 (7 arrays,
 * 3 x 15 line arrays, TID); the packed Payment block is 10 * cap u32. */
void ltpg_neworder_view(const uint32_t *block, uint32_t cap, uint32_t n, ltpg_neworder_batch *out);
void ltpg_payment_view(const uint32_t *block, uint32_t cap, uint32_t n, ltpg_payment_batch *out);

/* Make (neworder, payment) the current epoch (copy_to_device,
 * TPCC-allopt-final/Query.h:532-605).  Either may be NULL or empty.
 * NewOrder q has batch position q, Payment p has position n_neworder + p. */
int ltpg_tpcc_submit(ltpg_ctx *ctx, const ltpg_neworder_batch *neworder,
                     const ltpg_payment_batch *payment, uint32_t mem);
/* Read-only order-entry transactions (the Order-Status and Stock-Level
 * slots of TPCC-alltxn/Execute.h:1592-1770,2155-2231, whose query sets the
 * reference declares but never fills, TPCC-alltxn/Query.h:390-392).
 *   Order-Status o reads CUSTOMER row W_ID*30000 + D_ID*3000 + C_ID (column
 *     2), ORDER row W_ID*30000 + O_ID (column 0) and ORDERLINE row
 *     W_ID*450000 + OL_ID (column 0) (op_0 .. op_2);
 *   Stock-Level k reads STOCK row W_ID*100000 + S_ID (column 2) (op_0).
 * Each read registers R and counts raw against W; having no writes, the
 * transactions always commit (Aria's rule with war = waw = 0), and their R
 * registrations give war to NewOrders that write those rows.  Priorities:
 * Order-Status o has batch position n_neworder + n_payment + o, Stock-Level
 * k position n_neworder + n_payment + n_orderstatus + k.  The query's C_LAST
 * and isName fields are not used by the reference's Order-Status slots and
 * are not taken.  Not available with LTPG_FLAG_CONFLICT_COUNTS. */
typedef struct ltpg_orderstatus_batch {
    uint32_t n;
    uint32_t reserved;
    const uint32_t *W_ID, *D_ID, *C_ID, *O_ID, *OL_ID, *TID;  /* all [n]      */
} ltpg_orderstatus_batch;

typedef struct ltpg_stocklevel_batch {
    uint32_t n;
    uint32_t reserved;
    const uint32_t *W_ID, *S_ID, *TID;                         /* all [n]      */
} ltpg_stocklevel_batch;

/* Add read-only transactions to the batch of the last ltpg_tpcc_submit
 * (which starts every epoch with none).  Either may be NULL or empty; all
 * four types together must fit max_txns. */
int ltpg_tpcc_submit_readonly(ltpg_ctx *ctx, const ltpg_orderstatus_batch *orderstatus,
                              const ltpg_stocklevel_batch *stocklevel, uint32_t mem);

/* Delivery transactions (DeliveryQuery {O_ID, NO_W_ID, NO_O_ID} plus TID,
 * TPCC-alltxn/Datastructure.h:201-206,278-288; the slots of TPCC-alltxn/
 * Execute.h:1772-2153, whose query set the reference declares but never
 * fills).  Rows as the reference computes them, without a warehouse offset
 * except for NEWORDER:
 *   op_0: NEWORDER row NO_W_ID * 30000 + NO_O_ID, writes (the delete);
 *   op_1 / op_2: ORDER row O_ID, reads column 3, writes;
 *   op_3 / op_4: ORDERLINE row O_ID, writes, reads column 3;
 *   op_5: CUSTOMER row O_ID, writes.
 * Checks as the reference dispatches them (its check_4_0 and check_5_0 sit
 * behind repeated conditions and never run): raw from the NEWORDER row's R
 * and W and from the ORDER and ORDERLINE rows' R, war from the ORDER row's
 * W, waw from the ORDER, ORDERLINE and CUSTOMER rows' W.  War implies waw
 * (the same register), so a Delivery commits iff no higher-priority
 * transaction writes its ORDER, ORDERLINE or CUSTOMER row.  Its NEWORDER
 * write conflicts only with NewOrders under LTPG_RULES_INTENDED (their
 * check_3_1); under the reference rules a NewOrder inserting the row and a
 * Delivery deleting it both commit, and the row ends as the last of them in
 * priority order leaves it.  Committed reads: ORDER column 3, ORDERLINE
 * column 3 (value-major, 2 per Delivery).  LTPG_VALUES_RMW (no write-back
 * in the reference): NEWORDER row <- (0, 0, 0); ORDER O_CARRIER_ID (column
 * 5) <- TID % 10 + 1; ORDERLINE OL_DELIVERY_D (column 6) <- TID; CUSTOMER
 * C_BALANCE (column 5) += OL_AMOUNT (ORDERLINE column 8) and
 * C_DELIVERY_CNT (column 8) += 1.  Batch positions follow every other type:
 * Delivery d at n_neworder + n_payment + n_orderstatus + n_stocklevel + d.
 * Not available with LTPG_FLAG_CONFLICT_COUNTS. */
typedef struct ltpg_delivery_batch {
    uint32_t n;
    uint32_t reserved;
    const uint32_t *O_ID, *NO_W_ID, *NO_O_ID, *TID;          /* all [n]      */
} ltpg_delivery_batch;

/* Add Deliveries to the batch of the last ltpg_tpcc_submit (which starts
 * every epoch with none), before or after its read-only transactions; all
 * five types together must fit max_txns. */
int ltpg_tpcc_submit_delivery(ltpg_ctx *ctx, const ltpg_delivery_batch *delivery, uint32_t mem);

/* C_ID used by each Payment of the current batch: the query's C_ID, or,
 * once the CID phase has run, the by-name resolution for isName Payments
 * (KERNEL_CID writes it back into PAYMENT_QUERY.C_ID, TPCC-allopt-final/
 * tp.cu:11-23, Execute_payment.h:13-36): the C_ID column of the k-th
 * (0-based, k = C_LAST & 0xff) customer in row order of district
 * C_W_ID * 10 + C_D_ID whose C_LAST equals C_LAST >> 8, or 0 if none. */
int ltpg_tpcc_payment_cid(ltpg_ctx *ctx, uint32_t *out, uint32_t n);
/* Read values of the last epoch's committed transactions (the read slots of
 * set_local_set, TPCC-allopt-final/Execute_neworder.h:30-202,
 * Execute_payment.h:55-239), value-major: value v of transaction q at
 * out[v * n + q], n = the type's batch size, taken from the snapshot.
 *   NEWORDER (51): W_TAX; D_TAX, D_NEXT_O_ID; C_DISCOUNT, C_CREDIT, C_LAST;
 *     then I_PRICE, I_DATA, I_NAME of lines 0..14 (0 beyond O_OL_CNT).
 *   PAYMENT (24): WAREHOUSE cols 3..8; DISTRICT cols 5..10; CUSTOMER cols
 *     3, 4, 11..20.
 *   ORDERSTATUS (3): CUSTOMER col 2, ORDER col 0, ORDERLINE col 0.
 *   STOCKLEVEL (1): STOCK col 2.
 *   DELIVERY (2): ORDER col 3, ORDERLINE col 3.
 * A slot whose row is outside its table reads 0; aborted transactions'
 * entries are not written. */
int ltpg_tpcc_read_values(ltpg_ctx *ctx, uint32_t type, uint32_t *out, uint64_t n_words);

/* ---- seeded order-entry generator (MakeNeworder / MakePayment,
 *      TPCC-allopt-final/Query.h:16-212, with its commented-out options
 *      enabled: NURand items, remote supply warehouse, remote customer) ---- */
typedef struct ltpg_tpcc_gen_config {
    uint64_t seed;
    uint32_t warehouses;
    uint32_t neworder_pct;        /* share of NewOrder in a batch (exact)      */
    uint32_t remote_supply_pct;   /* per order line (1)                         */
    uint32_t remote_customer_pct; /* per Payment (15)                           */
    uint32_t byname_pct;          /* per Payment (60)                           */
    uint32_t nurand_items;        /* 1: item = NURand(8191, 1, 100000) - 1      */
    uint32_t min_lines, max_lines;/* O_OL_CNT uniform (5, 15)                   */
    uint32_t max_amount;          /* H_AMOUNT uniform in [1, max_amount] (5)    */
    uint32_t max_quantity;        /* OL_QUANTITY uniform in [1, max_quantity]   */
} ltpg_tpcc_gen_config;
typedef struct ltpg_tpcc_gen ltpg_tpcc_gen;

void ltpg_tpcc_gen_config_init(ltpg_tpcc_gen_config *c);
int ltpg_tpcc_gen_create(const ltpg_tpcc_gen_config *c, ltpg_tpcc_gen **out);
int ltpg_tpcc_gen_destroy(ltpg_tpcc_gen *g);
/* Next batch of n_txns transactions (TIDs tid_base .. tid_base + n_txns - 1
 * in arrival order, types interleaved by a seeded shuffle with exact
 * counts) into packed reference-layout blocks of capacity no_cap / pay_cap.
 * The per-warehouse order / history id counters continue across calls
 * (wrapping as the reference does, Query.h:54-88,157-166, with order lines
 * kept inside their warehouse). */
int ltpg_tpcc_gen_next(ltpg_tpcc_gen *g, uint32_t n_txns, uint32_t tid_base, uint32_t *no_block,
                       uint32_t no_cap, uint32_t *pay_block, uint32_t pay_cap, uint32_t *n_neworder,
                       uint32_t *n_payment);

/* ---- launch shape of the key-value kernels (replaces the fixed warp-role
 *      and grid constants, YCSB/0/Predefine.h:27-35) ------------------------ */
/* EXEC_PER_OP: EXECUTE runs one lane per op instead of one lane per
 *   transaction.
 * SPLIT_REGISTERS: a row's R and W registers live in two arrays instead of
 *   one 16-byte pair.
 * SMALL_CHUNKS: CHECK and WRITEBACK blocks own 64 transactions instead of
 *   256 (more blocks, fewer dependent rounds each).
 * LANE_COPY: the read-out and write-back copy each field by one lane with
 *   16-byte loads (fields of 16-byte aligned stride, at most 32 words)
 *   instead of by 32 lanes one word each; faster for small epochs, slower
 *   for large ones.
 * SMALL_EPOCH: an epoch of at most LTPG_SMALL_EPOCH_OPS op slots (n_txns *
 *   n_slots) run with LTPG_PHASE_ALL, in the fast form (no
 *   LTPG_FLAG_CONFLICT_COUNTS), without timing, and with a hot row only if
 *   the table has at most 1024 fields, runs
 *   as one single-workgroup kernel: ops, flags and the W registrations stay
 *   in LDS and the per-row registers in HBM are not touched.  One launch
 *   instead of four or five; other epochs run as before.  Not one of the
 *   autotuner's combinations (LTPG_TUNE_ALL).  The served form runs its
 *   small key-value epochs this way whatever the context's flags.
 * PARTITIONED: EXECUTE of the fast form (no LTPG_FLAG_CONFLICT_COUNTS) bins
 *   the epoch's register operations by row range (8192 rows per bucket) and
 *   keeps each range's W minima in LDS, instead of probing and updating the
 *   per-row registers in HBM at random (EXEC_PER_OP then has no effect).
 *   Two streaming passes over the ops and one over the binned operations
 *   (12 bytes each); the autotuner tries it with SPLIT_REGISTERS,
 *   SMALL_CHUNKS and LANE_COPY.
 *   Needs kv_rows <= 8192 * 16384 and kv_max_slots * max_txns *
 *   max(1, kv_max_scan) < 2^32 (else LTPG_ERR_UNSUPPORTED); the first
 *   selection allocates its buffers, 12 bytes per possible operation.
 * MORE_ITEMS: items per thread -- the per-op EXECUTE (EXEC_PER_OP) takes 4
 *   ops per lane instead of 2 (its LDS fold spans 1024 ops), and the read
 *   marking 2 instead of 1; no effect with PARTITIONED.
 * Results are identical for every combination; may change between any two
 * epochs. */
#define LTPG_TUNE_EXEC_PER_OP 1u
#define LTPG_TUNE_SPLIT_REGISTERS 2u
#define LTPG_TUNE_SMALL_CHUNKS 4u
#define LTPG_TUNE_LANE_COPY 8u
#define LTPG_TUNE_ALL 15u
#define LTPG_TUNE_SMALL_EPOCH 16u
#define LTPG_TUNE_PARTITIONED 32u
#define LTPG_TUNE_MORE_ITEMS 64u
#define LTPG_SMALL_EPOCH_OPS 2048u
int ltpg_kv_set_tuning(ltpg_ctx *ctx, uint32_t flags);
/* Launch shape of an order-entry context's CID, EXECUTE, CHECK and
 * WRITEBACK kernels (no counterpart in the reference, whose grid constants
 * are fixed, TPCC-allopt-final/Predefine.h:35-46).  TPCC_SMALL_BLOCKS: 64-thread
 * workgroups (4 NewOrders or 64 Payments each) instead of 256, which spreads
 * a small batch over four times as many CUs.  Results are identical; may
 * change between any two epochs.  LTPG_ERR_UNSUPPORTED on a key-value
 * context. */
#define LTPG_TUNE_TPCC_SMALL_BLOCKS 128u
int ltpg_tpcc_set_tuning(ltpg_ctx *ctx, uint32_t flags);

/* ---- epoch phases (KERNEL_EXECUTE/CHECK/WRITEBACK, TPCC-allopt-final/
 *      tp.cu:24-117, YCSB/0/ycsb.cu:9-58; launched by tp.cu:154-236) ------- */
/* Run the phases in phase_mask on the current batch, in the order CID,
 * EXECUTE, CHECK, WRITEBACK.  Asynchronous on the context's stream; errors
 * raised by kernels are returned by the next ltpg_sync(). */
int ltpg_epoch_run(ltpg_ctx *ctx, uint32_t phase_mask);
/* Make every per-row register read as "no registrant" (clear_LOG,
 * TPCC-allopt-final/Database.h:78-102).  O(1): registers carry an epoch tag,
 * so this advances the tag instead of writing 0xFFFFFFFF over every row. */
int ltpg_epoch_reset(ltpg_ctx *ctx);

/* ---- results (statistic_query / copy_to_host, TPCC-allopt-final/Query.h:
 *      606-631,697-725; COMMIT_AND_ABORT, YCSB/0/Execute.h:176-192) ------- */
/* Commit bitmap of the last epoch: bit t of word t/64 is 1 iff transaction t
 * of `type` (LTPG_TYPE_KV, or LTPG_TYPE_NEWORDER / LTPG_TYPE_PAYMENT indexed
 * within their own batch) committed.  Synchronises the context. */
int ltpg_commit_bitmap(ltpg_ctx *ctx, uint32_t type, uint64_t *out, uint64_t n_words);
/* Number of committed transactions of `type` in the last epoch. */
int ltpg_commit_count(ltpg_ctx *ctx, uint32_t type, uint32_t *out);
/* Per-transaction conflict counters raw | war << 8 | waw << 16 (the
 * reference's raw/war/waw[NO], TPCC-allopt-final/Datastructure.h:231-233).
 * Needs LTPG_FLAG_CONFLICT_COUNTS (else LTPG_ERR_UNSUPPORTED). */
int ltpg_conflict_counters(ltpg_ctx *ctx, uint32_t type, uint32_t *out, uint32_t n);
/* Read values of the last key-value epoch (set_local_set,
 * YCSB/0/Genericfunction.h:31-45): with L = max(1, kv_max_scan), word w of
 * row i (0 for a READ, 0 .. len - 1 for a SCAN) of slot s of transaction t
 * at out[(((s * B) + t) * L + i) * field_words + w], taken from the snapshot
 * (the table before the epoch's write-back).  Only read and scan slots of
 * committed transactions are written, and of a scan only the rows inside
 * the table: an aborted transaction's reads are discarded. */
int ltpg_kv_read_values(ltpg_ctx *ctx, uint32_t *out, uint64_t n_words);
/* Running totals since the last reset: committed and attempted (checked)
 * transactions of `type`, kept on the device so a pipelined epoch loop needs
 * no per-epoch readback.  reset != 0 zeroes them after reading. */
int ltpg_commit_totals(ltpg_ctx *ctx, uint32_t type, uint64_t *commits, uint64_t *attempts,
                       int reset);
/* Enable (1) or disable (0) per-phase HIP events (off by default).  While
 * enabled every ltpg_epoch_run records events around each phase on the
 * context's stream. */
int ltpg_set_timing(ltpg_ctx *ctx, int enable);
/* Timing slots: CID, EXECUTE, CHECK, WRITEBACK's read-out of committed
 * reads, WRITEBACK's writes. */
#define LTPG_TIMING_SLOTS 5
/* Sum of the device time in milliseconds of each timing slot over the
 * epochs timed since the previous call, and their number; then starts a new
 * measurement.  Synchronises the context. */
int ltpg_phase_times(ltpg_ctx *ctx, float ms[LTPG_TIMING_SLOTS], uint32_t *n_epochs);

/* ---- batched form: many batches of uneven size and mixed types in one
 *      submission (replaces the tp driver's per-epoch loop, TPCC-allopt-
 *      final/tp.cu:136-261, YCSB/0/ycsb.cu:78-152, and its query prefetch
 *      threads, TPCC-allopt-final/Query.h:640-725) --------------------------- */
typedef struct ltpg_batch_desc {
    uint32_t workload;              /* LTPG_WORKLOAD_KV or LTPG_WORKLOAD_TPCC     */
    uint32_t reserved;
    ltpg_kv_batch kv;               /* device pointers (workload KV)              */
    ltpg_neworder_batch neworder;   /* device pointers (workload TPCC)            */
    ltpg_payment_batch payment;
    ltpg_orderstatus_batch orderstatus;  /* device pointers, n may be 0 (workload
                                            TPCC: read-only transactions added to
                                            the epoch, ltpg_tpcc_submit_readonly) */
    ltpg_stocklevel_batch stocklevel;
    ltpg_delivery_batch delivery;    /* device pointers, n may be 0 (workload TPCC,
                                        ltpg_tpcc_submit_delivery)                */
} ltpg_batch_desc;

/* PLAIN: one epoch per batch, in list order, as the reference's driver loop
 *   runs it (TPF/tp.cu:136-261, Y0/ycsb.cu:78-152): each phase kernel
 *   followed by a stream synchronisation, then the commit-count readback.
 * GRAPH: the batches of each workload, in list order, captured once into a
 *   HIP graph (register refill, then per batch: submit, CID / EXECUTE /
 *   CHECK / WRITEBACK, bitmap archive copy); the key-value and order-entry
 *   graphs run concurrently on their contexts' streams, with one host
 *   synchronisation at the end.
 * Both produce identical commit sets and tables: batches of one workload run
 * in list order, and the two workloads share no table. */
#define LTPG_RUN_PLAIN 0u
#define LTPG_RUN_GRAPH 1u

typedef struct ltpg_batched ltpg_batched;

/* A runner over n batches whose arrays stay resident in device memory until
 * ltpg_batched_destroy.  kv_ctx / tpcc_ctx may be NULL when no batch of that
 * workload is listed; each context's capacity bounds its batches. */
int ltpg_batched_create(ltpg_ctx *kv_ctx, ltpg_ctx *tpcc_ctx, const ltpg_batch_desc *descs, uint32_t n,
                        ltpg_batched **out);
/* Run every batch once in `mode`.  Blocks until done; range errors of any
 * batch are returned as LTPG_ERR_RANGE. */
int ltpg_batched_run(ltpg_batched *b, uint32_t mode);
/* Commit bitmap of batch i (type LTPG_TYPE_KV, or LTPG_TYPE_NEWORDER /
 * LTPG_TYPE_PAYMENT for an order-entry batch) from the last run. */
int ltpg_batched_bitmap(ltpg_batched *b, uint32_t batch, uint32_t type, uint64_t *out, uint64_t n_words);
/* Completion time of each batch of the last run, in ms from the run's
 * start (every batch counted as arriving at the start): host clock after the
 * commit readback for PLAIN, the device clock after the batch's last kernel
 * for GRAPH. */
int ltpg_batched_times(ltpg_batched *b, double *ms, uint32_t n);
int ltpg_batched_destroy(ltpg_batched *b);

/* Autotuner.  For every size class (floor(log2(n_txns))) of the runner's
 * key-value batches and of its order-entry batches it times every distinct
 * launch shape on the first batch of the class (reps epochs after one
 * warm-up, HIP events on the context's stream, write-back in copy-back mode
 * so the tables are left as they were), keeps the fastest for every batch of
 * the class, and captures the graphs again.  Key-value shapes: the
 * LTPG_TUNE_ALL combinations, each with and without MORE_ITEMS (EXEC_PER_OP
 * and MORE_ITEMS act only without PARTITIONED), and PARTITIONED with the
 * SPLIT_REGISTERS / SMALL_CHUNKS / LANE_COPY combinations; order-entry
 * shapes: 0 and LTPG_TUNE_TPCC_SMALL_BLOCKS.  One entry per class and
 * workload is written to out, key-value classes first. */
typedef struct ltpg_tuning_entry {
    uint32_t workload;         /* LTPG_WORKLOAD_KV or LTPG_WORKLOAD_TPCC            */
    uint32_t log2_txns;        /* the size class                                    */
    uint32_t flags;            /* chosen LTPG_TUNE_* combination                    */
    uint32_t reserved;
    float ms[64];              /* epoch time of every shape tried, 0 for the others:
                                  key-value at LTPG_TUNE_INDEX(flags), order-entry
                                  [0] (256-thread blocks) and [1] (SMALL_BLOCKS) */
} ltpg_tuning_entry;
#define LTPG_TUNE_INDEX(f) (((f) & LTPG_TUNE_ALL) | (((f) & LTPG_TUNE_PARTITIONED) ? 16u : 0u) | \
                            (((f) & LTPG_TUNE_MORE_ITEMS) ? 32u : 0u))
int ltpg_batched_autotune(ltpg_batched *b, uint32_t reps, ltpg_tuning_entry *out, uint32_t max_out,
                          uint32_t *n_out);
/* One launch shape for every batch of the runner, in place of the
 * autotuner's choice: the key-value bits of flags (LTPG_TUNE_SMALL_EPOCH
 * excluded) for its key-value batches, LTPG_TUNE_TPCC_SMALL_BLOCKS for its
 * order-entry batches; captures the graphs again. */
int ltpg_batched_set_tuning(ltpg_batched *b, uint32_t flags);

/* ---- served form: a request loop over resident tables (replaces the
 *      producer/consumer threads around the tp loop, TPCC-allopt-final/
 *      Query.h:640-725, tp.cu:136-261, with continuous admission) ---------- */
typedef struct ltpg_serve_config {
    uint32_t size;            /* sizeof(ltpg_serve_config)                      */
    uint32_t slots;           /* epochs in flight per workload (1..4; 2 = the
                                 next epoch is packed and copied while one runs) */
    uint32_t kv_ops;          /* op slots per key-value request (<= kv_max_slots) */
    uint32_t max_epoch_txns;  /* cap on transactions per epoch (0 = context max) */
    uint64_t hbm_budget;      /* bytes of per-epoch access-set buffers (batch
                                 slots + read values) admitted per slot; 0 = no cap */
    uint32_t keep_reads;      /* 1: keep committed requests' read values        */
    uint32_t record_epochs;   /* 1: keep every epoch's member list (tests)       */
    uint32_t epoch_growth;    /* > 0: an epoch admits at most epoch_growth x the
                                 commits of the workload's last epoch (at least
                                 epoch_floor) -- under contention epochs commit a short
                                 chain whatever their size; 0 = always the cap.
                                 Default 8. */
    uint32_t plain;           /* 1: run the same admission as the plain
                                 per-batch form instead, for comparison: one
                                 epoch at a time, the batch copied from host
                                 memory in the reference's slot-major layout,
                                 phases, a synchronisation and the commit
                                 bitmap read back before the next epoch is
                                 formed (the driver loop of tp.cu:136-261);
                                 `slots` is ignored.  Default 0. */
    uint32_t epoch_floor;     /* smallest epoch the epoch_growth cap allows
                                 (0 = 32).  Default 32: with 64-op requests an
                                 epoch of 32 fits one single-workgroup kernel
                                 (LTPG_SMALL_EPOCH_OPS) */
    uint32_t kv_loop;         /* 1: key-value epochs form and run on the device
                                 in one persistent single-workgroup kernel
                                 (LTPG_SMALL_EPOCH_OPS op slots per epoch, fast
                                 form); requests reach it through a ring in
                                 pinned host memory and aborted ones stay on
                                 the device.  0: the host forms each epoch.
                                 Ignored with `plain` or conflict counts.
                                 Default 1. */
    uint32_t ring_requests;   /* key-value request ring of the device loop, in
                                 requests (rounded up to a power of two; 0 =
                                 65536); a submit waits while it is full */
    uint64_t first_seq;       /* arrival sequence number of each workload's
                                 first request (default 0) */
} ltpg_serve_config;

/* Priorities of the served form.  Each workload numbers its requests in
 * arrival order with a 64-bit sequence number that never wraps; a request's
 * batch TID is the low 32 bits of it, and no epoch spans a 2^32 boundary of
 * those bits, so within every epoch TID order is arrival order and a
 * re-queued request always precedes every later arrival.  A key-value write
 * adds TID | 1 to each word of its field; a Delivery's values use its TID. */
typedef struct ltpg_serve_done {
    uint64_t id;              /* request id (submission order, both workloads)   */
    uint32_t type;            /* LTPG_TYPE_*                                     */
    uint32_t tid;             /* low 32 bits of the arrival sequence number      */
    uint32_t attempts;        /* epochs the request took part in                */
    uint32_t tid_hi;          /* high 32 bits of the arrival sequence number     */
    uint64_t epoch;           /* serving epoch (per workload) it committed in    */
    uint64_t arrival_ns;      /* steady clock at submission                     */
    uint64_t commit_ns;       /* steady clock when its epoch's result was read  */
} ltpg_serve_done;

typedef struct ltpg_serve_stats {
    uint64_t committed[6];    /* by type (LTPG_TYPE_KV .. LTPG_TYPE_DELIVERY)    */
    uint64_t attempts[6];     /* transaction-epochs by type                     */
    uint64_t epochs[2];       /* key-value, order-entry                         */
    uint64_t max_epoch_txns[2];
} ltpg_serve_stats;

/* The key-value device loop (ltpg_serve_config.kv_loop) as the device
 * measured it.  No counterpart in the reference. */
typedef struct ltpg_serve_device_stats {
    uint32_t size;            /* sizeof(ltpg_serve_device_stats), set by the caller */
    uint32_t kv_loop;         /* 1 when the key-value lane runs the device loop   */
    uint64_t launches;        /* kernel launches (it returns when idle, etc.)      */
    uint64_t epochs;          /* epochs it ran                                    */
    uint64_t busy_ns;         /* device wall-clock time inside those epochs       */
    uint64_t exits[5];        /* launches ended by: 0 other, 1 stop, 2 idle,
                                 3 full done ring, 4 time slice                   */
    uint64_t prof_ns[10];     /* environment LTPG_SERVE_PROFILE set at start:
                                 time in each part of the epochs (control read,
                                 admission, EXECUTE, read marking, CHECK, unused,
                                 read-out and write-back, publication, requeue,
                                 idle) -- a profiled epoch takes a few tenths
                                 of a microsecond longer                         */
    uint64_t prof_epochs;     /* epochs profiled                                  */
} ltpg_serve_device_stats;

typedef struct ltpg_server ltpg_server;

void ltpg_serve_config_init(ltpg_serve_config *c);
/* Start a serving thread over kv_ctx and / or tpcc_ctx (either may be NULL).
 * While it runs, the contexts belong to the server.  LTPG_ERR_BUDGET when
 * the budget admits no transaction. */
int ltpg_serve_start(ltpg_ctx *kv_ctx, ltpg_ctx *tpcc_ctx, const ltpg_serve_config *cfg, ltpg_server **out);
/* n key-value requests, request-major: op k of request i at
 * keys / meta[i * kv_ops + k] (LTPG_KV_META; NONE pads). */
int ltpg_serve_submit_kv(ltpg_server *sv, uint32_t n, const uint64_t *keys, const uint32_t *meta,
                         uint64_t *first_id);
/* Order-entry requests in the batch layouts (TID fields are ignored: the
 * server assigns priorities in arrival order).  NewOrders get ids
 * first_id .. first_id + n_no - 1, then the Payments. */
int ltpg_serve_submit_tpcc(ltpg_server *sv, const ltpg_neworder_batch *neworder,
                           const ltpg_payment_batch *payment, uint64_t *first_id);
/* Read-only order-entry requests (Order-Status, Stock-Level; host arrays,
 * TID fields ignored), which join the order-entry epochs after their
 * NewOrders and Payments as ltpg_tpcc_submit_readonly adds them.  Ids
 * first_id .. for the Order-Status requests, then the Stock-Level ones.
 * LTPG_ERR_UNSUPPORTED for a context with LTPG_FLAG_CONFLICT_COUNTS.  No
 * counterpart in the reference, whose read-only transactions are unfed
 * stubs (TPCC-alltxn/Execute.h:1592-1770,2155-2231). */
int ltpg_serve_submit_readonly(ltpg_server *sv, const ltpg_orderstatus_batch *orderstatus,
                               const ltpg_stocklevel_batch *stocklevel, uint64_t *first_id);
/* Delivery requests (host arrays, TID field ignored), which join the
 * order-entry epochs after every other type as ltpg_tpcc_submit_delivery
 * adds them; ids first_id ..  LTPG_ERR_UNSUPPORTED with
 * LTPG_FLAG_CONFLICT_COUNTS.  No counterpart in the reference (an unfed
 * stub there, TPCC-alltxn/Execute.h:1772-2153). */
int ltpg_serve_submit_delivery(ltpg_server *sv, const ltpg_delivery_batch *delivery, uint64_t *first_id);
/* Up to max completed (committed) requests, oldest first. */
int ltpg_serve_poll(ltpg_server *sv, ltpg_serve_done *out, uint32_t max, uint32_t *n_out);
/* Read values of a committed request (keep_reads): key-value: the field
 * words of its read ops, and of each scan's rows inside the table, in op
 * order; NewOrder 51 / Payment 24 / Order-Status 3 / Stock-Level 1 /
 * Delivery 2 values as ltpg_tpcc_read_values lays them out for one
 * transaction. */
int ltpg_serve_read_values(ltpg_server *sv, uint64_t id, uint32_t *out, uint32_t max_words, uint32_t *n_words);
/* Request ids of serving epoch e of a workload, in batch order (record_epochs). */
int ltpg_serve_epoch(ltpg_server *sv, uint32_t workload, uint64_t e, uint64_t *ids, uint32_t max, uint32_t *n);
/* Block until every submitted request has committed. */
int ltpg_serve_drain(ltpg_server *sv);
int ltpg_serve_stats_get(ltpg_server *sv, ltpg_serve_stats *st);
int ltpg_serve_device_stats_get(ltpg_server *sv, ltpg_serve_device_stats *st);
/* Drain, stop the thread and free the server (the contexts stay). */
int ltpg_serve_stop(ltpg_server *sv);
/* Drop every request not in an epoch yet (new and re-queued, including the
 * aborts of the epochs still in flight), let those epochs finish, then stop
 * and free the server.  *dropped = requests that never committed. */
int ltpg_serve_cancel(ltpg_server *sv, uint64_t *dropped);

/* ---- a key-value epoch that allocates nothing --------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: one key-value epoch (EXECUTE, CHECK with the read-out of
 * committed reads, WRITEBACK; the fast form, no hot row, no scans) over a
 * caller-owned device table, with the same kernels and results as
 * ltpg_epoch_run on a context.  It calls no allocator: the per-row registers,
 * per-transaction flags and counters live in a fixed device array of the
 * library (about 19 MB, held by each GPU it runs on until the process ends).
 *   table        device, rows x fields x field_bytes, row-major, packed;
 *                updated in place (RMW: a write adds TID | 1 to each word)
 *   batch        device arrays, slot-major (slot s of txn t at s*B + t)
 *   bitmap       device, ceil(n_txns / 64) words: the commit bitmap
 *   read_values  device, n_slots * n_txns * field_bytes / 4 words, laid out
 *                as ltpg_kv_read_values (max_scan 0)
 *   result       device, 2 words, or NULL: commit count, ops out of range
 *   hip_stream   the stream the work is queued on (NULL: the null stream)
 * Asynchronous.  Sizes beyond LTPG_NOALLOC_MAX_ROWS / LTPG_NOALLOC_MAX_TXNS
 * are refused with LTPG_ERR_INVALID.  Every call on one GPU uses that GPU's
 * one array: calls that overlap on a GPU (two streams, two threads) race;
 * serialise them. */
#define LTPG_NOALLOC_MAX_ROWS (1u << 20)
#define LTPG_NOALLOC_MAX_TXNS (1u << 20)
int ltpg_kv_epoch_noalloc(uint32_t *table, uint64_t rows, uint32_t fields, uint32_t field_bytes,
                          const ltpg_kv_batch *batch, uint32_t values, uint64_t *bitmap, uint32_t *read_values,
                          uint32_t *result, void *hip_stream);

/* ---- seeded workload generators (replace MakeYCSB / Random.h, whose
 *      rand() after srand(time(0)) cannot be reproduced: YCSB/0/Query.h:
 *      17-121, YCSB/0/Random.h:12,52) ----------------------------------- */
typedef struct ltpg_kv_gen {
    uint64_t seed;
    uint64_t rows;           /* keys are row ids in [0, rows)                  */
    uint32_t fields;         /* field index uniform in [0, fields)             */
    uint32_t n_txns;
    uint32_t n_slots;        /* S of the slot-major output                     */
    uint32_t min_ops;        /* ops per txn uniform in [min_ops, max_ops]      */
    uint32_t max_ops;        /* <= n_slots; slots beyond a txn's count are     */
                             /* LTPG_KV_NONE                                   */
    uint32_t read_pct;       /* percent of ops that are reads                  */
    double theta;            /* Zipf skew; 0 = uniform                         */
    uint32_t tid_base;       /* tids[t] = tid_base + t                         */
    uint32_t distinct_keys;  /* 1: keys within one transaction are distinct    */
    uint64_t txn_base;       /* global index of transaction 0 (stream offset)  */
    uint32_t scan_len;       /* > 0: reads are SCANs of this length (YCSB/2)   */
    uint32_t reserved;
} ltpg_kv_gen;

void ltpg_kv_gen_init(ltpg_kv_gen *g);
/* Fill keys[S*B], meta[S*B], tids[B].  Deterministic in (seed, txn_base + t,
 * slot); uses up to `threads` host threads (0 = hardware concurrency). */
int ltpg_kv_generate(const ltpg_kv_gen *g, uint64_t *keys, uint32_t *meta, uint32_t *tids,
                     uint32_t threads);

#ifdef __cplusplus
}
#endif
#endif /* LTPG_H */
