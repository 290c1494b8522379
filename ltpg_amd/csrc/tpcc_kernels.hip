// tpcc_kernels.hip - order-entry (NewOrder + Payment) kernels for gfx950.
//
// The reference runs one warp "role" per op slot per 32 transactions
// (KERNEL_EXECUTE / KERNEL_CHECK / KERNEL_WRITEBACK,
// LTPG/TPCC-allopt-final/tp.cu:24-117, with the role tables of
// Execute_neworder.h and Execute_payment.h), registering every slot and
// staging every column into a set_local_set array.  Here:
//
//   CID      one thread per Payment; a by-name Payment binary-searches its
//            district's sorted (C_LAST, row) index.
//   EXECUTE  registers only the registers some check can observe.  Under the
//            reference's dispatch (Execute_neworder.h:620-680,
//            Execute_payment.h:547-595) the WAREHOUSE, DISTRICT and ITEM
//            registers and the NEWORDER insert's R register are never
//            consulted, or consulted only on a register nobody lowers, so
//            they cannot change any outcome.  What remains is W of CUSTOMER
//            (Payment), ORDER, STOCK, ORDERLINE, HISTORY and NEWORDER, plus
//            R of CUSTOMER when the raw / war / waw counters are requested.
//            NewOrders run as 16 waves per 64 transactions (15 line roles and
//            a header role, wave-uniform), Payments one thread each.  waw is
//            found here, by displacement (device_common.h).
//   CHECK    every war of a NewOrder or Payment comes from a register that
//            nobody lowers except CUSTOMER R, and a Payment's CUSTOMER raw
//            (check_4_0) reads the same W register as its waw (check_5_1);
//            so raw && war implies waw and the rule waw == 0 && (raw == 0 ||
//            war == 0) is waw == 0, which EXECUTE has already decided.  CHECK
//            reads the doomed flag, writes the commit bitmap and gathers the
//            committed transactions' read values from the snapshot.  With
//            conflict counters it recomputes raw / war / waw from the
//            registers as the reference's check slots do.
//   WRITEBACK one thread per committed transaction, its lines in order.
//
// Tables are row-major, `stride` (cols rounded up to 4) u32 words per row.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_common.h"
#include "ltpg_internal.h"

namespace ltpg {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

enum : int {
    WH = LTPG_TPCC_WAREHOUSE,
    DI = LTPG_TPCC_DISTRICT,
    CU = LTPG_TPCC_CUSTOMER,
    HI = LTPG_TPCC_HISTORY,
    NO = LTPG_TPCC_NEWORDER,
    OR = LTPG_TPCC_ORDER,
    OL = LTPG_TPCC_ORDERLINE,
    ST = LTPG_TPCC_STOCK,
    IT = LTPG_TPCC_ITEM,
};

// Row strides (u32 words) of the device layout.
constexpr uint32_t kStride[kTpccTables] = {12, 12, 24, 8, 4, 8, 12, 20, 8};

__device__ __forceinline__ uint64_t smix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// The loader's uniform draw in [a, b] for (seed, stream, row).
__device__ __forceinline__ uint32_t pop_uniform(uint64_t seed, uint32_t stream, uint64_t row, uint32_t a,
                                                uint32_t b)
{
    const uint64_t h = smix(smix(seed ^ 0x504F50554C415445ull) ^ ((uint64_t)stream << 48) ^ row);
    return a + (uint32_t)__umul64hi(h, (uint64_t)b - a + 1);
}

// ---- loader (create_warehouse .. create_item, Database.h:104-327) -------

template <int TABLE>
__device__ __forceinline__ void row_values(uint64_t i, uint64_t seed, uint32_t (&v)[kStride[TABLE]])
{
#pragma unroll
    for (uint32_t k = 0; k < kStride[TABLE]; k++) v[k] = 0;
    if constexpr (TABLE == WH) {
        v[0] = (uint32_t)i; v[1] = 1; v[2] = 1; v[6] = 10;
    } else if constexpr (TABLE == DI) {
        v[0] = (uint32_t)(i % 10); v[1] = (uint32_t)(i / 10); v[2] = 1; v[3] = 1;
    } else if constexpr (TABLE == CU) {
        v[0] = (uint32_t)(i % 3000); v[1] = (uint32_t)(i / 30000); v[2] = (uint32_t)((i / 3000) % 10);
#pragma unroll
        for (int c = 3; c <= 9; c++) v[c] = 100000;
        v[10] = pop_uniform(seed, CU, i, 0, 1);
        v[13] = (uint32_t)(i % 20);
    } else if constexpr (TABLE == OR) {
        v[0] = (uint32_t)(i % 30000); v[1] = (uint32_t)(i / 3000); v[2] = (uint32_t)(i / 30000);
        v[3] = pop_uniform(seed, OR, i, 0, 2999);
    } else if constexpr (TABLE == OL) {
        v[0] = (uint32_t)(i % 450000); v[1] = (uint32_t)(i / 45000); v[2] = (uint32_t)(i / 450000);
        v[4] = pop_uniform(seed, OL, i, 0, 99999); v[5] = (uint32_t)(i / 450000);
    } else if constexpr (TABLE == ST) {
        v[0] = (uint32_t)(i % 100000); v[1] = (uint32_t)(i / 100000); v[2] = 100000; v[3] = 100000;
#pragma unroll
        for (int dd = 1; dd <= 10; dd++) v[5 + dd] = dd;
    } else if constexpr (TABLE == IT) {
        v[0] = (uint32_t)i; v[1] = 1000;
        v[2] = pop_uniform(seed, IT, i, 1, 40000);
        v[3] = pop_uniform(seed, IT + 1, i, 1, 10);
        v[4] = (uint32_t)i;
    }
    // HISTORY and NEWORDER start empty.
}

template <int TABLE>
__global__ __launch_bounds__(kBlock) void populate_kernel(uint32_t *__restrict__ data, uint64_t rows, uint64_t seed)
{
    constexpr uint32_t S = kStride[TABLE];
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < rows; i += step) {
        uint32_t v[S];
        row_values<TABLE>(i, seed, v);
        uint4 *dst = reinterpret_cast<uint4 *>(data + i * S);
#pragma unroll
        for (uint32_t k = 0; k < S / 4; k++) dst[k] = make_uint4(v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]);
    }
}

// ---- by-name index -------------------------------------------------------

// One workgroup per district: sort its 3000 (C_LAST, row-in-district) keys
// in LDS (bitonic, padded to 4096) so that the k-th holder of a name, in row
// order, is entry lower_bound(name) + k.  The reference keeps C_LAST in row
// order and scans it (Database.h:182, Genericfunction.h:29-77).
constexpr uint32_t kPerDistrict = 3000;
constexpr uint32_t kSortN = 4096;

__global__ __launch_bounds__(kBlock) void name_index_kernel(const uint32_t *__restrict__ customer,
                                                           unsigned long long *__restrict__ index)
{
    __shared__ unsigned long long keys[kSortN];
    const uint64_t base = (uint64_t)blockIdx.x * kPerDistrict;
    for (uint32_t i = threadIdx.x; i < kSortN; i += kBlock)
        keys[i] = i < kPerDistrict
                      ? ((unsigned long long)customer[(base + i) * kStride[CU] + 13] << 32) | i
                      : ~0ull;
    __syncthreads();
    for (uint32_t k = 2; k <= kSortN; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = threadIdx.x; t < kSortN / 2; t += kBlock) {
                const uint32_t a = (t / j) * 2 * j + (t % j), b = a + j;
                const bool up = (a & k) == 0;
                const unsigned long long x = keys[a], y = keys[b];
                if ((x > y) == up) {
                    keys[a] = y;
                    keys[b] = x;
                }
            }
            __syncthreads();
        }
    }
    for (uint32_t i = threadIdx.x; i < kPerDistrict; i += kBlock) index[base + i] = keys[i];
}

// ---- layout conversion ------------------------------------------------------

// to_device: dense host layout (column-major src[c * rows + r] or row-major
// src[r * cols + c]) -> strided device rows (padding zeroed); else back.
__global__ __launch_bounds__(kBlock) void relayout_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                                         uint64_t rows, uint32_t cols, uint32_t stride,
                                                         uint32_t layout, int to_device)
{
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < rows; r += step) {
        for (uint32_t c = 0; c < stride; c++) {
            const uint64_t dense = layout == LTPG_LAYOUT_COLUMN_MAJOR ? (uint64_t)c * rows + r : r * cols + c;
            if (to_device)
                dst[r * stride + c] = c < cols ? src[dense] : 0u;
            else if (c < cols)
                dst[dense] = src[r * stride + c];
        }
    }
}

// ---- CID (KERNEL_CID / get_c_id, tp.cu:11-23, Execute_payment.h:13-36) --

// C_ID of the k-th (0-based, k = C_LAST & 0xff) customer of district
// C_W_ID * 10 + C_D_ID whose C_LAST is C_LAST >> 8; 0 when there is none.
template <uint32_t BLK>
__global__ __launch_bounds__(BLK) void cid_kernel(TpccDev d, PaymentDev pay)
{
    const uint32_t p = blockIdx.x * BLK + threadIdx.x;
    if (p >= pay.n) return;
    uint32_t c = pay.C_ID[p];
    if (pay.isName[p]) {
        c = 0;
        const uint64_t dist = (uint64_t)pay.C_W_ID[p] * 10 + pay.C_D_ID[p];
        if (dist < d.t[DI].rows) {
            const uint32_t cl = pay.C_LAST[p], name = cl >> 8, k = cl & 0xffu;
            const unsigned long long *idx = d.name_index + dist * kPerDistrict;
            const unsigned long long key = (unsigned long long)name << 32;
            uint32_t lo = 0, hi = kPerDistrict;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (idx[mid] < key) lo = mid + 1; else hi = mid;
            }
            const uint32_t j = lo + k;
            if (j < kPerDistrict) {
                const unsigned long long e = idx[j];
                if ((uint32_t)(e >> 32) == name)
                    c = d.t[CU].data[(dist * kPerDistrict + (uint32_t)e) * kStride[CU]];
            }
        }
    }
    d.cid[p] = c;
}

// ---- row formulas (u64; the reference's u32 forms, Execute_neworder.h:22-266,
//      Execute_payment.h:47-286) --------------------------------------------

struct NoRows {
    uint64_t wh, di, cu, no, od;
};

__device__ __forceinline__ NoRows no_rows(const NeworderDev &no, uint32_t q)
{
    NoRows r;
    const uint64_t w = no.W_ID[q], dd = no.D_ID[q];
    r.wh = w;
    r.di = w * 10 + dd;
    r.cu = w * 30000 + dd * 3000 + no.C_ID[q];
    r.no = w * 30000 + no.N_O_ID[q];
    r.od = w * 30000 + no.O_ID[q];
    return r;
}

__device__ __forceinline__ uint32_t no_lines(const NeworderDev &no, uint32_t q)
{
    return min(no.O_OL_CNT[q], (uint32_t)LTPG_TPCC_MAX_LINES);
}

struct PayRows {
    uint64_t wh, di, cu, hi;
};

__device__ __forceinline__ PayRows pay_rows(const TpccDev &d, const PaymentDev &pay, uint32_t p)
{
    PayRows r;
    const uint64_t w = pay.W_ID[p];
    r.wh = w;
    r.di = w * 10 + pay.D_ID[p];
    r.cu = (uint64_t)pay.C_W_ID[p] * 30000 + (uint64_t)pay.C_D_ID[p] * 3000 + d.cid[p];
    r.hi = w * 30000 + pay.H_ID[p];
    return r;
}

// Register addresses.
__device__ __forceinline__ unsigned long long *cust_r(const TpccDev &d, uint64_t row) { return d.regs + 2 * row; }
__device__ __forceinline__ unsigned long long *cust_w(const TpccDev &d, uint64_t row) { return d.regs + 2 * row + 1; }

// ---- EXECUTE ---------------------------------------------------------------

// NewOrders take 16 adjacent lanes each (lane j < 15: line j, lane 15: the
// header slots), BLK / 16 orders per block: an order's consecutive ORDERLINE
// registers are touched by adjacent lanes, and all its roles share a CU.
// Payments take one lane each, in the blocks after the NewOrder blocks.
// BLK is 256, or 64 for the small-blocks launch shape (LTPG_TUNE_TPCC_SMALL_BLOCKS),
// which spreads a small batch over four times as many CUs.
template <bool COUNTS, uint32_t BLK>
__global__ __launch_bounds__(BLK) void execute_kernel(TpccDev d, NeworderDev no, PaymentDev pay, uint32_t tag,
                                                        uint8_t epoch, uint32_t no_blocks, DeviceCounters *ctr)
{
    // the epoch's commit counts start at zero (CHECK adds to them)
    if (blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += BLK)
            ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    uint32_t bad = 0;
    if (blockIdx.x < no_blocks) {
        const uint32_t q = blockIdx.x * (BLK / 16) + threadIdx.x / 16;
        const uint32_t j = threadIdx.x % 16;
        if (q < no.n) {
            const unsigned long long me = priority(tag, no.TID[q], q);
            bool waw = false;
            if (j < LTPG_TPCC_MAX_LINES) {
                // line role: STOCK W (op_6_0), ORDERLINE W (op_7_0); the ITEM
                // slot (op_5_0) registers a register no check consults
                if (j < no_lines(no, q)) {
                    const uint64_t at = (uint64_t)j * no.line_stride + q;
                    const uint32_t item = no.OL_I_ID[at];
                    const uint64_t sr = (uint64_t)no.OL_SUPPLY_W_ID[at] * 100000 + item;
                    const uint64_t lr = (uint64_t)no.W_ID[q] * 450000 + no.O_OL_ID[q] + j;
                    bad += item >= d.t[IT].rows;
                    const bool sok = sr < d.t[ST].rows, lok = lr < d.t[OL].rows;
                    bad += !sok + !lok;
                    // both probes in flight before either atomic
                    unsigned long long *rs = d.regs + d.off_stock + (sok ? sr : 0);
                    unsigned long long *rl = d.regs + d.off_line + (lok ? lr : 0);
                    const unsigned long long cs = probe(rs), cl = probe(rl);
                    if (sok) waw |= finish_write(rs, cs, me, tag, d.doomed, epoch);
                    if (lok) waw |= finish_write(rl, cl, me, tag, d.doomed, epoch);
                }
            } else {
                // header role: slots 0..4
                const NoRows r = no_rows(no, q);
                bad += r.wh >= d.t[WH].rows;
                bad += r.di >= d.t[DI].rows;
                if (r.cu < d.t[CU].rows) {
                    if (COUNTS) register_read(cust_r(d, r.cu), me);
                } else bad++;
                // NEWORDER insert: a write (waw) under INTENDED; under
                // REFERENCE its R register is never consulted
                if (r.no < d.t[NO].rows) {
                    if (d.rules == LTPG_RULES_INTENDED)
                        waw |= register_write(d.regs + d.off_new + r.no, me, tag, d.doomed, epoch);
                } else bad++;
                if (r.od < d.t[OR].rows) waw |= register_write(d.regs + d.off_order + r.od, me, tag, d.doomed, epoch);
                else bad++;
            }
            if (waw) d.doomed[q] = epoch;
        }
    } else {
        const uint32_t p = (blockIdx.x - no_blocks) * BLK + threadIdx.x;
        if (p < pay.n) {
            const unsigned long long me = priority(tag, pay.TID[p], no.n + p);
            const PayRows r = pay_rows(d, pay, p);
            bad += r.wh >= d.t[WH].rows ? 2 : 0;  // slots 0, 1
            bad += r.di >= d.t[DI].rows ? 2 : 0;  // slots 2, 3
            bool waw = false;
            const bool cok = r.cu < d.t[CU].rows, hok = r.hi < d.t[HI].rows;
            bad += (cok ? 0 : 2) + !hok;          // slots 4, 5 and the HISTORY insert
            if (COUNTS && cok) register_read(cust_r(d, r.cu), me);
            // both probes in flight before either atomic
            unsigned long long *rc = cust_w(d, cok ? r.cu : 0);
            unsigned long long *rh = d.regs + d.off_hist + (hok ? r.hi : 0);
            const unsigned long long cc = probe(rc), ch = probe(rh);
            if (cok) waw |= finish_write(rc, cc, me, tag, d.doomed, epoch);
            if (hok) waw |= finish_write(rh, ch, me, tag, d.doomed, epoch);
            if (waw) d.doomed[no.n + p] = epoch;
        }
    }
    if (bad) atomicAdd(&ctr->range_errors, bad);
}

// ---- CHECK + read-out of committed reads -------------------------------------

__device__ __forceinline__ uint32_t lower(const unsigned long long *r, unsigned long long me)
{
    return *r < me ? 1u : 0u;
}

// Committed NewOrder q: W_TAX; D_TAX, D_NEXT_O_ID; C_DISCOUNT, C_CREDIT,
// C_LAST; per line I_PRICE, I_DATA, I_NAME (op_0_1, op_1_1, op_2_1, op_5_1).
__device__ __forceinline__ void gather_neworder(const TpccDev &d, const NeworderDev &no, uint32_t q)
{
    const NoRows r = no_rows(no, q);
    const uint32_t L = no_lines(no, q);
    const uint32_t n = no.n;
    uint32_t *out = d.reads_no + q;
    const bool okw = r.wh < d.t[WH].rows, okd = r.di < d.t[DI].rows, okc = r.cu < d.t[CU].rows;
    const uint32_t *w = d.t[WH].data + (okw ? r.wh : 0) * kStride[WH];
    const uint32_t *di = d.t[DI].data + (okd ? r.di : 0) * kStride[DI];
    const uint32_t *cu = d.t[CU].data + (okc ? r.cu : 0) * kStride[CU];
    uint32_t v[kNoReads];
    v[0] = w[1];
    v[1] = di[2];
    v[2] = di[4];
    v[3] = cu[4];
    v[4] = cu[10];
    v[5] = cu[13];
    // every line slot exists in the batch arrays (15 * line_stride), so the
    // item ids load unconditionally and the loads of all lines overlap
    uint32_t ok_items = 0;
#pragma unroll
    for (uint32_t i = 0; i < LTPG_TPCC_MAX_LINES; i++) {
        const uint32_t item = no.OL_I_ID[(uint64_t)i * no.line_stride + q];
        const bool ok = i < L && item < d.t[IT].rows;
        ok_items |= (uint32_t)ok << i;
        const uint32_t *it = d.t[IT].data + (uint64_t)(ok ? item : 0) * kStride[IT];
        v[6 + 3 * i] = it[2];
        v[7 + 3 * i] = it[3];
        v[8 + 3 * i] = it[4];
    }
    if (!okw) v[0] = 0;
    if (!okd) v[1] = v[2] = 0;
    if (!okc) v[3] = v[4] = v[5] = 0;
#pragma unroll
    for (uint32_t i = 0; i < LTPG_TPCC_MAX_LINES; i++)
        if (!((ok_items >> i) & 1u)) v[6 + 3 * i] = v[7 + 3 * i] = v[8 + 3 * i] = 0;
#pragma unroll
    for (uint32_t k = 0; k < kNoReads; k++) out[(uint64_t)k * n] = v[k];
}

// Committed Payment p: warehouse cols 3..8, district cols 5..10, customer
// cols 3, 4, 11..20 (op_0_1/0_2, op_2_1/2_2, op_4_1..4_4).
__device__ __forceinline__ void gather_payment(const TpccDev &d, const PaymentDev &pay, uint32_t p,
                                               const PayRows &r)
{
    const uint32_t n = pay.n;
    uint32_t *out = d.reads_pay + p;
    const bool okw = r.wh < d.t[WH].rows, okd = r.di < d.t[DI].rows, okc = r.cu < d.t[CU].rows;
    const uint32_t *w = d.t[WH].data + (okw ? r.wh : 0) * kStride[WH];
    const uint32_t *di = d.t[DI].data + (okd ? r.di : 0) * kStride[DI];
    const uint32_t *cu = d.t[CU].data + (okc ? r.cu : 0) * kStride[CU];
    uint32_t v[kPayReads];
#pragma unroll
    for (uint32_t k = 0; k < 6; k++) v[k] = okw ? w[3 + k] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < 6; k++) v[6 + k] = okd ? di[5 + k] : 0u;
    v[12] = okc ? cu[3] : 0u;
    v[13] = okc ? cu[4] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < 10; k++) v[14 + k] = okc ? cu[11 + k] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kPayReads; k++) out[(uint64_t)k * n] = v[k];
}

// Read-only transactions, one thread each, after every write has
// registered (Order-Status op_0..op_2, TPCC-alltxn/Execute.h:1594-1672;
// Stock-Level op_0, :2157-2181; their check slots consult W, :1710-1745,
// 2203-2214).  They register no R: the only use of R in the commit rule is
// war of a row's writer, and a writer without waw is the row's minimum, the
// owner of W; so a read of row x against w = W[x] of this epoch with
// w > me flags w's owner with war directly.  A read-only transaction always
// commits (war = waw = 0); the thread gathers its read values (the table is
// the snapshot until WRITEBACK) and sets its commit bit.
__device__ __forceinline__ void ro_mark(const TpccDev &d, const unsigned long long *reg, unsigned long long me,
                                        uint32_t tag, uint8_t epoch)
{
    const unsigned long long w = *reg;
    if ((uint32_t)(w >> kTagShift) == tag && w > me) d.war[w & kPosMask] = epoch;
}

__global__ __launch_bounds__(kBlock) void readonly_kernel(TpccDev d, ReadOnlyDev ro, uint32_t base, uint32_t tag,
                                                         uint8_t epoch, uint32_t os_waves, DeviceCounters *ctr)
{
    __shared__ uint32_t s_cnt[2];  // Order-Status, Stock-Level transactions (all commit)
    if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t wave = blockIdx.x * kWaves + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    const bool is_os = wave < os_waves;
    const uint32_t i = (is_os ? wave : wave - os_waves) * 64 + lane;
    const bool live = i < (is_os ? ro.n_os : ro.n_sl);
    uint32_t bad = 0;
    if (live && is_os) {
        const uint64_t w = ro.os_W_ID[i];
        const uint64_t cu = w * 30000 + (uint64_t)ro.os_D_ID[i] * 3000 + ro.os_C_ID[i];   // op_0_0
        const uint64_t od = w * 30000 + ro.os_O_ID[i];                                     // op_1_0
        const uint64_t ol = w * 450000 + ro.os_OL_ID[i];                                   // op_2_0
        const unsigned long long me = priority(tag, ro.os_TID[i], base + i);
        const bool okc = cu < d.t[CU].rows, oko = od < d.t[OR].rows, okl = ol < d.t[OL].rows;
        uint32_t v[kOsReads];
        v[0] = okc ? d.t[CU].data[cu * kStride[CU] + 2] : 0u;  // op_0_1
        v[1] = oko ? d.t[OR].data[od * kStride[OR]] : 0u;      // op_1_1
        v[2] = okl ? d.t[OL].data[ol * kStride[OL]] : 0u;      // op_2_1
        if (okc) ro_mark(d, cust_w(d, cu), me, tag, epoch); else ++bad;
        if (oko) ro_mark(d, d.regs + d.off_order + od, me, tag, epoch); else ++bad;
        if (okl) ro_mark(d, d.regs + d.off_line + ol, me, tag, epoch); else ++bad;
#pragma unroll
        for (uint32_t k = 0; k < kOsReads; k++) d.reads_os[(uint64_t)k * ro.n_os + i] = v[k];
    } else if (live) {
        const uint64_t st = (uint64_t)ro.sl_W_ID[i] * 100000 + ro.sl_S_ID[i];             // op_0_0
        const unsigned long long me = priority(tag, ro.sl_TID[i], base + ro.n_os + i);
        const bool oks = st < d.t[ST].rows;
        d.reads_sl[i] = oks ? d.t[ST].data[st * kStride[ST] + 2] : 0u;                     // op_0_1
        if (oks) ro_mark(d, d.regs + d.off_stock + st, me, tag, epoch); else ++bad;
    }
    const unsigned long long bits = __ballot(live);
    if (lane == 0 && bits) {
        if (is_os) d.bitmap_os[wave] = bits;
        else d.bitmap_sl[wave - os_waves] = bits;
        atomicAdd(&s_cnt[is_os ? 0 : 1], (uint32_t)__popcll(bits));
    }
    if (bad) atomicAdd(&ctr->range_errors, bad);
    __syncthreads();
    if (threadIdx.x < 2 && s_cnt[threadIdx.x]) {
        const uint32_t stripe = blockIdx.x % kCounterStripes, c = s_cnt[threadIdx.x];
        const uint32_t type = threadIdx.x == 0 ? LTPG_TYPE_ORDERSTATUS : LTPG_TYPE_STOCKLEVEL;
        atomicAdd(&ctr->commits[stripe][type], c);
        atomicAdd(&ctr->totals[stripe][type], (unsigned long long)c);
        atomicAdd(&ctr->totals[stripe][kAttempts + type], (unsigned long long)c);
    }
}

template <bool COUNTS, uint32_t BLK>
__global__ __launch_bounds__(BLK) void check_kernel(TpccDev d, NeworderDev no, PaymentDev pay, uint32_t tag,
                                                      uint8_t epoch, uint32_t no_waves, bool ro, bool dl,
                                                      DeviceCounters *ctr)
{
    __shared__ uint32_t s_cnt[4];  // NewOrder commits, attempts, Payment commits, attempts
    if (threadIdx.x < 4) s_cnt[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t wave = blockIdx.x * (BLK / 64) + threadIdx.x / 64;
    const uint32_t lane = threadIdx.x % 64;
    bool commit = false, live = false;
    const bool is_no = wave < no_waves;
    if (is_no) {
        const uint32_t q = wave * 64 + lane;
        live = q < no.n;
        if (live) {
            if (COUNTS) {
                // check dispatch as it runs (Execute_neworder.h:620-680); the
                // dropped slots read registers nobody lowers: their count is 0
                const unsigned long long me = priority(tag, no.TID[q], q);
                const NoRows r = no_rows(no, q);
                uint32_t raw = 0, waw = 0;
                if (r.cu < d.t[CU].rows) raw += lower(cust_w(d, r.cu), me);                      // check_2_0
                if (d.rules == LTPG_RULES_INTENDED && r.no < d.t[NO].rows)
                    waw += lower(d.regs + d.off_new + r.no, me);                                 // check_3_1
                if (r.od < d.t[OR].rows) waw += lower(d.regs + d.off_order + r.od, me);          // check_4_1
                const uint32_t L = no_lines(no, q);
                for (uint32_t i = 0; i < L; i++) {
                    const uint64_t at = (uint64_t)i * no.line_stride + q;
                    const uint64_t sr = (uint64_t)no.OL_SUPPLY_W_ID[at] * 100000 + no.OL_I_ID[at];
                    const uint64_t lr = (uint64_t)no.W_ID[q] * 450000 + no.O_OL_ID[q] + i;
                    if (sr < d.t[ST].rows) waw += lower(d.regs + d.off_stock + sr, me);          // check_6_1
                    if (lr < d.t[OL].rows) waw += lower(d.regs + d.off_line + lr, me);           // check_7_1
                }
                d.conf_no[q] = (raw & 0xffu) | ((waw & 0xffu) << 16);
                commit = waw == 0;
            } else {
                commit = d.doomed[q] != epoch;
                // read-only readers of its ORDER / STOCK / ORDERLINE rows
                // give a NewOrder war; it then aborts iff it also has raw,
                // whose only source is a lower Payment of its customer
                // (check_2_0: the WAREHOUSE, DISTRICT and ITEM W registers
                // are never lowered)
                if (ro && commit && d.war[q] == epoch) {
                    const NoRows r = no_rows(no, q);
                    commit = !(r.cu < d.t[CU].rows && *cust_w(d, r.cu) < priority(tag, no.TID[q], q));
                }
                // check_3_1 (waw) reads NEWORDER W, which Deliveries lower
                // without dooming anyone: decided here, against the register
                if (dl && commit) {
                    const NoRows r = no_rows(no, q);
                    commit = !(r.no < d.t[NO].rows && d.regs[d.off_new + r.no] < priority(tag, no.TID[q], q));
                }
            }
            if (commit) {
                gather_neworder(d, no, q);
                // the last committed inserter of a NEWORDER row writes it
                const uint64_t nr = (uint64_t)no.W_ID[q] * 30000 + no.N_O_ID[q];
                if (nr < d.t[NO].rows)
                    atomicMax(d.new_last + nr, ((unsigned long long)epoch << kTagShift) |
                                                   ((unsigned long long)no.TID[q] << kTidShift) | q);
            }
        }
    } else {
        const uint32_t p = (wave - no_waves) * 64 + lane;
        live = p < pay.n;
        if (live) {
            const PayRows r = pay_rows(d, pay, p);
            if (COUNTS) {
                const unsigned long long me = priority(tag, pay.TID[p], no.n + p);
                uint32_t raw = 0, war = 0, waw = 0;
                if (r.cu < d.t[CU].rows) {
                    raw += lower(cust_w(d, r.cu), me);                                           // check_4_0
                    war += lower(cust_r(d, r.cu), me);                                           // check_5_0
                    waw += lower(cust_w(d, r.cu), me);                                           // check_5_1
                }
                if (r.hi < d.t[HI].rows) waw += lower(d.regs + d.off_hist + r.hi, me);           // check_6_1
                d.conf_pay[p] = (raw & 0xffu) | ((war & 0xffu) << 8) | ((waw & 0xffu) << 16);
                commit = waw == 0 && (raw == 0 || war == 0);
            } else {
                commit = d.doomed[no.n + p] != epoch;
            }
            if (commit) gather_payment(d, pay, p, r);
        }
    }
    const unsigned long long bits = __ballot(commit);
    const unsigned long long lv = __ballot(live);
    if (lane == 0 && lv) {
        if (is_no) d.bitmap_no[wave] = bits;
        else d.bitmap_pay[wave - no_waves] = bits;
        atomicAdd(&s_cnt[is_no ? 0 : 2], (uint32_t)__popcll(bits));
        atomicAdd(&s_cnt[is_no ? 1 : 3], (uint32_t)__popcll(lv));
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        const uint32_t k = threadIdx.x;  // 0: NewOrder, 1: Payment
        const uint32_t c = s_cnt[2 * k], a = s_cnt[2 * k + 1];
        if (a) {
            const uint32_t stripe = blockIdx.x % kCounterStripes;
            const uint32_t type = k == 0 ? LTPG_TYPE_NEWORDER : LTPG_TYPE_PAYMENT;
            if (c) {
                atomicAdd(&ctr->commits[stripe][type], c);
                atomicAdd(&ctr->totals[stripe][type], (unsigned long long)c);
            }
            atomicAdd(&ctr->totals[stripe][kAttempts + type], (unsigned long long)a);
        }
    }
}

// ---- Deliveries ----------------------------------------------------------------

// DELIVERYQUERY_NAMESPACE (TPCC-alltxn/Execute.h:1772-2153), one thread per
// Delivery.  Its rows: NEWORDER NO_W_ID * 30000 + NO_O_ID (op_0_0) and
// O_ID, without a warehouse offset, in ORDER, ORDERLINE and CUSTOMER
// (op_1_0 .. op_5_0).
struct DlRows {
    uint64_t no, o;
};

__device__ __forceinline__ DlRows dl_rows(const DeliveryDev &dl, uint32_t i)
{
    return DlRows{(uint64_t)dl.NO_W_ID[i] * 30000 + dl.NO_O_ID[i], (uint64_t)dl.O_ID[i]};
}

// EXECUTE.  W on ORDER, ORDERLINE and CUSTOMER with early waw: the rows its
// check_2_1, check_3_1 and check_5_1 consult, and the others' waw / raw on
// them.  W on NEWORDER as a plain atomicMin: a NewOrder's check_3_1 (waw)
// compares against it in CHECK, and the Delivery's own check_0_1 counts raw,
// which never decides a Delivery (its only war, check_1_0, reads the ORDER W
// register its check_2_1 waw reads).  The R registrations of op_1_0 and
// op_4_0 are not kept: R of ORDER and ORDERLINE is consulted by a NewOrder's
// war (check_4_0, check_7_0), but a Delivery writes both rows it reads, so a
// NewOrder behind it on such a row has waw already; and by a Delivery's raw,
// which decides nothing.  Runs after the NewOrders and Payments registered,
// so no NewOrder ever displaces a Delivery's NEWORDER registration.
__global__ __launch_bounds__(kBlock) void delivery_execute_kernel(TpccDev d, DeliveryDev dl, uint32_t base,
                                                                 uint32_t tag, uint8_t epoch, DeviceCounters *ctr)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    uint32_t bad = 0;
    if (i < dl.n) {
        const DlRows r = dl_rows(dl, i);
        const unsigned long long me = priority(tag, dl.TID[i], base + i);
        const bool nok = r.no < d.t[NO].rows, ook = r.o < d.t[OR].rows, lok = r.o < d.t[OL].rows,
                   cok = r.o < d.t[CU].rows;
        bad += !nok + 2 * !ook + 2 * !lok + !cok;  // slots 0, 1-2, 3-4, 5
        unsigned long long *rn = d.regs + d.off_new + (nok ? r.no : 0);
        unsigned long long *ro = d.regs + d.off_order + (ook ? r.o : 0);
        unsigned long long *rl = d.regs + d.off_line + (lok ? r.o : 0);
        unsigned long long *rc = cust_w(d, cok ? r.o : 0);
        // every probe in flight before the first atomic
        const unsigned long long cn = probe(rn), co = probe(ro), cl = probe(rl), cc = probe(rc);
        if (nok && cn > me) atomicMin(rn, me);
        bool waw = false;
        if (ook) waw |= finish_write(ro, co, me, tag, d.doomed, epoch);
        if (lok) waw |= finish_write(rl, cl, me, tag, d.doomed, epoch);
        if (cok) waw |= finish_write(rc, cc, me, tag, d.doomed, epoch);
        if (waw) d.doomed[base + i] = epoch;
    }
    if (bad) atomicAdd(&ctr->range_errors, bad);
}

// CHECK: a Delivery commits iff it has no waw (its war implies waw, see
// above); committed ones read ORDER col 3 (op_1_1) and ORDERLINE col 3
// (op_4_1) from the snapshot and enter the NEWORDER row's last-writer race.
__global__ __launch_bounds__(kBlock) void delivery_check_kernel(TpccDev d, DeliveryDev dl, uint32_t base,
                                                               uint8_t epoch, DeviceCounters *ctr)
{
    __shared__ uint32_t s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < dl.n;
    const bool commit = live && d.doomed[base + i] != epoch;
    if (commit) {
        const DlRows r = dl_rows(dl, i);
        d.reads_dl[i] = r.o < d.t[OR].rows ? d.t[OR].data[r.o * kStride[OR] + 3] : 0u;
        d.reads_dl[(uint64_t)dl.n + i] = r.o < d.t[OL].rows ? d.t[OL].data[r.o * kStride[OL] + 3] : 0u;
        if (r.no < d.t[NO].rows)
            atomicMax(d.new_last + r.no, ((unsigned long long)epoch << kTagShift) |
                                             ((unsigned long long)dl.TID[i] << kTidShift) | (base + i));
    }
    const unsigned long long bits = __ballot(commit);
    if (threadIdx.x % 64 == 0 && __ballot(live)) {
        d.bitmap_dl[i / 64] = bits;
        atomicAdd(&s_cnt, (uint32_t)__popcll(bits));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t first = blockIdx.x * kBlock;
        const uint32_t a = dl.n - first < (uint32_t)kBlock ? dl.n - first : (uint32_t)kBlock;
        const uint32_t stripe = blockIdx.x % kCounterStripes, c = s_cnt;
        if (c) {
            atomicAdd(&ctr->commits[stripe][LTPG_TYPE_DELIVERY], c);
            atomicAdd(&ctr->totals[stripe][LTPG_TYPE_DELIVERY], (unsigned long long)c);
        }
        atomicAdd(&ctr->totals[stripe][kAttempts + LTPG_TYPE_DELIVERY], (unsigned long long)a);
    }
}

// WRITEBACK (RMW values; the reference writes nothing back for Deliveries):
// the NEWORDER row's last committed writer in priority order deletes it
// (0, 0, 0); ORDER O_CARRIER_ID <- TID % 10 + 1; ORDERLINE OL_DELIVERY_D <-
// TID; CUSTOMER C_BALANCE += OL_AMOUNT, C_DELIVERY_CNT += 1.  A committed
// Delivery is the only committed writer of its ORDER, ORDERLINE and
// CUSTOMER rows, so the OL_AMOUNT it adds is the snapshot's.
__global__ __launch_bounds__(kBlock) void delivery_writeback_kernel(TpccDev d, DeliveryDev dl, uint32_t base,
                                                                   uint8_t epoch)
{
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= dl.n || !((d.bitmap_dl[i / 64] >> (i % 64)) & 1ull)) return;
    const DlRows r = dl_rows(dl, i);
    const uint32_t tid = dl.TID[i];
    const unsigned long long last =
        ((unsigned long long)epoch << kTagShift) | ((unsigned long long)tid << kTidShift) | (base + i);
    if (r.no < d.t[NO].rows && d.new_last[r.no] == last) {
        uint32_t *row = d.t[NO].data + r.no * kStride[NO];
        row[0] = 0u;
        row[1] = 0u;
        row[2] = 0u;
    }
    if (r.o < d.t[OR].rows) d.t[OR].data[r.o * kStride[OR] + 5] = tid % 10u + 1u;
    uint32_t amount = 0;
    if (r.o < d.t[OL].rows) {
        uint32_t *row = d.t[OL].data + r.o * kStride[OL];
        amount = row[8];
        row[6] = tid;
    }
    if (r.o < d.t[CU].rows) {
        uint32_t *c = d.t[CU].data + r.o * kStride[CU];
        c[5] += amount;
        c[8] += 1u;
    }
}

// ---- WRITEBACK ---------------------------------------------------------------

// Committed transactions write their write slots (write_back dispatch,
// Execute_neworder.h:698-850, Execute_payment.h:613-736).  RMW values:
//   NEWORDER <- (O_ID, D_ID, W_ID), by its last committed inserter in
//               priority order only (under the reference rules two inserters
//               of one row do not conflict);
//   ORDER    <- (O_ID, D_ID, W_ID);
//   per line, in line order: STOCK (S_QUANTITY, S_YTD, S_ORDER_CNT) <-
//     (q >= qty + 10 ? q - qty : q - qty + 91, S_YTD + qty, S_ORDER_CNT + 1),
//     ORDERLINE <- (O_ID, D_ID, W_ID, line, I_ID, SUPPLY_W_ID, 0, qty,
//     qty * I_PRICE, I_DATA);
//   CUSTOMER C_BALANCE -= H_AMOUNT, C_DATA += H_AMOUNT;
//   HISTORY  <- (C_ID, C_D_ID, C_W_ID, D_ID, W_ID);
//   W_YTD, D_YTD += H_AMOUNT (commutative hot-row update, no register).
// Copy-back stores the value each write slot staged from the snapshot, so
// the rows are rewritten unchanged and the YTD update is not applied.  The
// committed writer of a row is unique (waw), so every row read here still
// holds its snapshot value, apart from a NewOrder's own earlier lines.
template <uint32_t BLK>
__global__ __launch_bounds__(BLK) void writeback_kernel(TpccDev d, NeworderDev no, PaymentDev pay, uint8_t epoch,
                                                          uint32_t no_blocks, int rmw)
{
    if (blockIdx.x < no_blocks) {
        // 16 lanes per NewOrder, as in EXECUTE: lane j < 15 writes line j,
        // lane 15 the NEWORDER and ORDER rows.  The group of an order is
        // active or idle as a whole, so the shuffles below read live lanes.
        const uint32_t q = blockIdx.x * (BLK / 16) + threadIdx.x / 16;
        const uint32_t j = threadIdx.x % 16;
        if (q >= no.n || !((d.bitmap_no[q / 64] >> (q % 64)) & 1ull)) return;
        const uint32_t oid = no.O_ID[q], did = no.D_ID[q], wid = no.W_ID[q];
        const uint32_t L = no_lines(no, q);
        if (j == 15) {
            const NoRows r = no_rows(no, q);
            const unsigned long long last =
                ((unsigned long long)epoch << kTagShift) | ((unsigned long long)no.TID[q] << kTidShift) | q;
            if (r.no < d.t[NO].rows && d.new_last[r.no] == last) {
                uint32_t *row = d.t[NO].data + r.no * kStride[NO];
                if (rmw) { row[0] = oid; row[1] = did; row[2] = wid; }
                else { const uint32_t a = row[0], b = row[1], c = row[2]; row[0] = a; row[1] = b; row[2] = c; }
            }
            if (r.od < d.t[OR].rows) {
                uint32_t *row = d.t[OR].data + r.od * kStride[OR];
                if (rmw) { row[0] = oid; row[1] = did; row[2] = wid; }
                else { const uint32_t a = row[0], b = row[1], c = row[2]; row[0] = a; row[1] = b; row[2] = c; }
            }
        }
        const bool has = j < L;
        const uint64_t at = (uint64_t)(has ? j : 0) * no.line_stride + q;
        const uint32_t item = no.OL_I_ID[at], sw = no.OL_SUPPLY_W_ID[at], qty = no.OL_QUANTITY[at];
        const uint64_t sr = (uint64_t)sw * 100000 + item;
        const unsigned long long key = has && sr < d.t[ST].rows ? sr : ~0ull;
        // Lines naming the same STOCK row apply in line order: the last of
        // them applies all of them; the others write nothing.
        const int base = (int)(threadIdx.x % 64) & ~15;
        uint32_t match = 0, qk[LTPG_TPCC_MAX_LINES];
        bool later = false;
#pragma unroll
        for (int k = 0; k < (int)LTPG_TPCC_MAX_LINES; k++) {
            const unsigned long long kk = __shfl(key, base + k);
            qk[k] = __shfl(qty, base + k);
            match |= (uint32_t)(kk == key && k <= (int)j) << k;
            later |= kk == key && k > (int)j;
        }
        if (key != ~0ull && !later) {
            uint32_t *st = d.t[ST].data + sr * kStride[ST];
            uint32_t sq = st[2], ytd = st[3], cnt = st[4];
            if (rmw) {
#pragma unroll
                for (int k = 0; k < (int)LTPG_TPCC_MAX_LINES; k++) {
                    if ((match >> k) & 1u) {
                        sq = sq >= qk[k] + 10 ? sq - qk[k] : sq - qk[k] + 91;
                        ytd += qk[k];
                        cnt += 1;
                    }
                }
            }
            st[2] = sq;
            st[3] = ytd;
            st[4] = cnt;
        }
        const uint64_t lr = (uint64_t)wid * 450000 + no.O_OL_ID[q] + j;
        if (has && lr < d.t[OL].rows) {
            uint4 *row = reinterpret_cast<uint4 *>(d.t[OL].data + lr * kStride[OL]);
            uint2 *tail = reinterpret_cast<uint2 *>(row + 2);
            if (rmw) {
                const uint32_t price = d.reads_no[(uint64_t)(6 + 3 * j) * no.n + q];
                const uint32_t data = d.reads_no[(uint64_t)(7 + 3 * j) * no.n + q];
                row[0] = make_uint4(oid, did, wid, j);
                row[1] = make_uint4(item, sw, 0u, qty);
                *tail = make_uint2(qty * price, data);
            } else {
                const uint4 a = row[0], b = row[1];
                const uint2 c = *tail;
                row[0] = a; row[1] = b; *tail = c;
            }
        }
    } else {
        const uint32_t p = (blockIdx.x - no_blocks) * BLK + threadIdx.x;
        if (p >= pay.n || !((d.bitmap_pay[p / 64] >> (p % 64)) & 1ull)) return;
        const PayRows r = pay_rows(d, pay, p);
        const uint32_t amt = pay.H_AMOUNT[p];
        if (r.cu < d.t[CU].rows) {
            uint32_t *c = d.t[CU].data + r.cu * kStride[CU];
            const uint32_t bal = c[5], data = c[9];
            c[5] = rmw ? bal - amt : bal;
            c[9] = rmw ? data + amt : data;
        }
        if (r.hi < d.t[HI].rows) {
            uint4 *row = reinterpret_cast<uint4 *>(d.t[HI].data + r.hi * kStride[HI]);
            uint32_t *tail = reinterpret_cast<uint32_t *>(row + 1);
            if (rmw) {
                row[0] = make_uint4(d.cid[p], pay.C_D_ID[p], pay.C_W_ID[p], pay.D_ID[p]);
                *tail = pay.W_ID[p];
            } else {
                const uint4 a = row[0];
                const uint32_t b = *tail;
                row[0] = a; *tail = b;
            }
        }
    }
}

// W_YTD / D_YTD += H_AMOUNT of committed Payments (the hot-row reduce,
// Genericfunction.h:431-518, applied to the table).  A block sums a chunk of
// Payments into LDS bins (W warehouse bins, then 10W district bins) and adds
// each non-zero bin to the table once, so a hot row takes one global atomic
// per block instead of one per Payment.  More bins than LDS holds: direct
// global atomics.
constexpr uint32_t kYtdBins = 12288;  // 48 KB of LDS
// 1024 threads per block: the block count is kept small (each block ends
// with one global atomic per touched bin), so each block is made wide
// instead, which keeps every thread's chain of dependent loads short
constexpr uint32_t kYtdBlock = 1024;

__global__ __launch_bounds__(kYtdBlock) void ytd_kernel(TpccDev d, PaymentDev pay, uint32_t chunk)
{
    __shared__ uint32_t bins[kYtdBins];
    const uint64_t W = d.t[WH].rows, nd = d.t[DI].rows;
    const bool lds = W + nd <= kYtdBins;
    if (lds)
        for (uint32_t i = threadIdx.x; i < W + nd; i += kYtdBlock) bins[i] = 0;
    __syncthreads();
    const uint64_t p0 = (uint64_t)blockIdx.x * chunk, p1 = min<uint64_t>(pay.n, p0 + chunk);
    for (uint64_t p = p0 + threadIdx.x; p < p1; p += kYtdBlock) {
        if (!((d.bitmap_pay[p / 64] >> (p % 64)) & 1ull)) continue;
        const uint64_t w = pay.W_ID[p], di = w * 10 + pay.D_ID[p];
        const uint32_t amt = pay.H_AMOUNT[p];
        if (w < W) {
            if (lds) atomicAdd(&bins[w], amt);
            else atomicAdd(d.t[WH].data + w * kStride[WH] + 2, amt);
        }
        if (di < nd) {
            if (lds) atomicAdd(&bins[W + di], amt);
            else atomicAdd(d.t[DI].data + di * kStride[DI] + 3, amt);
        }
    }
    __syncthreads();
    if (lds) {
        for (uint32_t i = threadIdx.x; i < W + nd; i += kYtdBlock) {
            const uint32_t v = bins[i];
            if (v == 0) continue;
            if (i < W) atomicAdd(d.t[WH].data + (uint64_t)i * kStride[WH] + 2, v);
            else atomicAdd(d.t[DI].data + (uint64_t)(i - W) * kStride[DI] + 3, v);
        }
    }
}

inline uint32_t grid_for(uint64_t n, uint32_t cap = 65536)
{
    const uint64_t b = (n + kBlock - 1) / kBlock;
    return (uint32_t)(b < 1 ? 1 : (b > cap ? cap : b));
}

inline uint32_t waves64(uint32_t n) { return (n + 63) / 64; }

}  // namespace

hipError_t launch_tpcc_populate(const TpccDev &d, uint64_t seed, hipStream_t s)
{
#define POP(T)                                                                                           \
    hipLaunchKernelGGL(populate_kernel<T>, dim3(grid_for(d.t[T].rows)), dim3(kBlock), 0, s, d.t[T].data, \
                       d.t[T].rows, seed)
    POP(WH); POP(DI); POP(CU); POP(HI); POP(NO); POP(OR); POP(OL); POP(ST); POP(IT);
#undef POP
    return hipGetLastError();
}

hipError_t launch_tpcc_name_index(const TpccDev &d, hipStream_t s)
{
    const uint64_t districts = d.t[DI].rows;
    if (districts == 0) return hipSuccess;
    hipLaunchKernelGGL(name_index_kernel, dim3((uint32_t)districts), dim3(kBlock), 0, s, d.t[CU].data, d.name_index);
    return hipGetLastError();
}

hipError_t launch_tpcc_relayout(const uint32_t *src, uint32_t *dst, uint64_t rows, uint32_t cols, uint32_t stride,
                                uint32_t layout, int to_device, hipStream_t s)
{
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(relayout_kernel, dim3(grid_for(rows)), dim3(kBlock), 0, s, src, dst, rows, cols, stride,
                       layout, to_device);
    return hipGetLastError();
}

hipError_t launch_tpcc_cid(const TpccDev &d, const PaymentDev &pay, bool small, hipStream_t s)
{
    if (pay.n == 0) return hipSuccess;
    if (small)
        hipLaunchKernelGGL(cid_kernel<64>, dim3((pay.n + 63) / 64), dim3(64), 0, s, d, pay);
    else
        hipLaunchKernelGGL(cid_kernel<kBlock>, dim3((pay.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d, pay);
    return hipGetLastError();
}

template <uint32_t BLK>
void execute_at(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e, bool counts,
                DeviceCounters *ctr, hipStream_t s)
{
    const uint32_t no_blocks = (no.n + BLK / 16 - 1) / (BLK / 16);
    const uint32_t blocks = std::max(1u, no_blocks + (pay.n + BLK - 1) / BLK);
    if (counts)
        hipLaunchKernelGGL((execute_kernel<true, BLK>), dim3(blocks), dim3(BLK), 0, s, d, no, pay, e.tag, e.epoch,
                           no_blocks, ctr);
    else
        hipLaunchKernelGGL((execute_kernel<false, BLK>), dim3(blocks), dim3(BLK), 0, s, d, no, pay, e.tag, e.epoch,
                           no_blocks, ctr);
}

hipError_t launch_tpcc_execute(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e,
                               bool counts, bool small, DeviceCounters *ctr, hipStream_t s)
{
    if (small) execute_at<64>(d, no, pay, e, counts, ctr, s);
    else execute_at<kBlock>(d, no, pay, e, counts, ctr, s);
    return hipGetLastError();
}

hipError_t launch_tpcc_readonly(const TpccDev &d, const ReadOnlyDev &ro, uint32_t base, const EpochTag &e,
                                DeviceCounters *ctr, hipStream_t s)
{
    const uint32_t os_waves = waves64(ro.n_os), waves = os_waves + waves64(ro.n_sl);
    if (waves == 0) return hipSuccess;
    hipLaunchKernelGGL(readonly_kernel, dim3((waves + kWaves - 1) / kWaves), dim3(kBlock), 0, s, d, ro, base, e.tag,
                       e.epoch, os_waves, ctr);
    return hipGetLastError();
}

template <uint32_t BLK>
void check_at(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e, bool counts, bool ro,
              bool dl, DeviceCounters *ctr, hipStream_t s)
{
    const uint32_t no_waves = waves64(no.n), waves = no_waves + waves64(pay.n);
    const uint32_t blocks = (waves + BLK / 64 - 1) / (BLK / 64);
    if (counts)
        hipLaunchKernelGGL((check_kernel<true, BLK>), dim3(blocks), dim3(BLK), 0, s, d, no, pay, e.tag, e.epoch,
                           no_waves, ro, dl, ctr);
    else
        hipLaunchKernelGGL((check_kernel<false, BLK>), dim3(blocks), dim3(BLK), 0, s, d, no, pay, e.tag, e.epoch,
                           no_waves, ro, dl, ctr);
}

hipError_t launch_tpcc_check(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e,
                             bool counts, bool ro, bool dl, bool small, DeviceCounters *ctr, hipStream_t s)
{
    if (waves64(no.n) + waves64(pay.n) == 0) return hipSuccess;
    if (small) check_at<64>(d, no, pay, e, counts, ro, dl, ctr, s);
    else check_at<kBlock>(d, no, pay, e, counts, ro, dl, ctr, s);
    return hipGetLastError();
}

hipError_t launch_tpcc_delivery_execute(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                        DeviceCounters *ctr, hipStream_t s)
{
    if (dl.n == 0) return hipSuccess;
    hipLaunchKernelGGL(delivery_execute_kernel, dim3((dl.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d, dl, base,
                       e.tag, e.epoch, ctr);
    return hipGetLastError();
}

hipError_t launch_tpcc_delivery_check(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                      DeviceCounters *ctr, hipStream_t s)
{
    if (dl.n == 0) return hipSuccess;
    hipLaunchKernelGGL(delivery_check_kernel, dim3((dl.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d, dl, base,
                       e.epoch, ctr);
    return hipGetLastError();
}

hipError_t launch_tpcc_delivery_writeback(const TpccDev &d, const DeliveryDev &dl, uint32_t base, const EpochTag &e,
                                          uint32_t values, hipStream_t s)
{
    if (dl.n == 0 || values != LTPG_VALUES_RMW) return hipSuccess;  // copy-back changes nothing
    hipLaunchKernelGGL(delivery_writeback_kernel, dim3((dl.n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, d, dl, base,
                       e.epoch);
    return hipGetLastError();
}

hipError_t launch_tpcc_writeback(const TpccDev &d, const NeworderDev &no, const PaymentDev &pay, const EpochTag &e,
                                 uint32_t values, bool small, hipStream_t s)
{
    const uint32_t blk = small ? 64 : kBlock;
    const uint32_t no_blocks = (no.n + blk / 16 - 1) / (blk / 16);
    const uint32_t blocks = no_blocks + (pay.n + blk - 1) / blk;
    if (blocks == 0) return hipSuccess;
    const int rmw = values == LTPG_VALUES_RMW ? 1 : 0;
    if (small)
        hipLaunchKernelGGL(writeback_kernel<64>, dim3(blocks), dim3(64), 0, s, d, no, pay, e.epoch, no_blocks, rmw);
    else
        hipLaunchKernelGGL(writeback_kernel<kBlock>, dim3(blocks), dim3(kBlock), 0, s, d, no, pay, e.epoch, no_blocks,
                           rmw);
    if (rmw && pay.n) {
        const uint32_t chunk = std::max<uint32_t>(4096, (pay.n + 255) / 256);
        hipLaunchKernelGGL(ytd_kernel, dim3((pay.n + chunk - 1) / chunk), dim3(kYtdBlock), 0, s, d, pay, chunk);
    }
    return hipGetLastError();
}

}  // namespace ltpg
