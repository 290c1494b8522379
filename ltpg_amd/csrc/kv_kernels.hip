// kv_kernels.hip - key-value execute / check / readout / write-back kernels
// for gfx950.
//
// Replaces KERNEL_EXECUTE / KERNEL_CHECK / KERNEL_WRITEBACK of
// LTPG/YCSB/0/ycsb.cu:9-58 and the per-op bodies of YCSB/0/Execute.h:11-252
// and Genericfunction.h:31-158.
//
// Mapping.  The batch is slot-major (slot s of transaction t at [s*B + t],
// the reference's LOC[B*op + NO]).  One lane owns one transaction and walks
// its slots, so every slot-level load of a wave is 64 consecutive elements
// (coalesced), per-transaction state stays in registers (no per-op
// atomicAdd on raw/war/waw[NO]), and the commit decision is fused into the
// check: a wave64 ballot of the commit predicate is one 64-bit word of the
// commit bitmap.
//
// Registers.  A row's R and W registers sit at rl.r(row) and rl.w(row)
// (RegLayout: interleaved pairs or two arrays), each a u64 priority
//     tag:8 | TID:32 | batch position:24
// where tag = 255 - epoch (epochs 1..255; the library refills every register
// with all-ones when the epoch counter wraps).  Lower value = higher
// priority, and every value of an older epoch is larger than every value of
// the current one, so a single atomicMin both registers and evicts a stale
// value, and "register < my priority" already implies "registered in this
// epoch" -- stale registers never produce a conflict.
//
// Early waw.  atomicMin returns the old value.  A write whose registration
// finds a lower priority already there has waw > 0; a write that lowers the
// register from v displaces the writer whose position is in v, which then
// has waw > 0.  Every writer except the minimum one is caught by exactly one
// of the two (the first registration that goes below a writer sees exactly
// that writer's value), so after EXECUTE the doomed[] bytes mark precisely
// the transactions with waw > 0, and CHECK only walks the others.
//
// Field copies.  A field is field_words u32 words inside a row-major record
// whose fields start at multiples of field_stride words (field_stride =
// field_words rounded up to 16 words = 64 bytes for fields of 64 bytes or
// more, so a 100-byte field touches exactly two 64-byte segments).  A wave
// moves the fields of its ops cooperatively, each op's addresses fetched from
// its owner lane by __shfl: in CHECK's read-out and in WRITEBACK, fields of
// up to 32 words two per instruction (32 lanes per field, wave_copy_pairs)
// and wider ones G lanes per field (group_copy), or, with the LANE_COPY
// shape, one lane per field with 16-byte loads (lane_copy); the counting
// form's read-out LP lanes per op (LP = next power of two >= field_words,
// at most 64).
#include "device_common.h"
#include "ltpg_internal.h"

namespace ltpg {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Copy (or read-modify-write) the fields of the active ops of a wave.
// Lane j's op: src_base[src_off + w] (+ delta) -> dst_base[dst_off + w] for
// w < fw.  Requires fw <= LP.  Ops of one call never alias each other
// (different transactions of one wave never write one row in the same
// slot), so the loads of U steps are issued before their stores.  A lane
// always handles the same word positions of the same transaction, so
// successive calls on one row by one transaction are ordered.
template <int LP, int U>
__device__ __forceinline__ void wave_copy(unsigned long long mask, uint64_t src_off,
                                          uint64_t dst_off, uint32_t delta,
                                          const uint32_t *src_base, uint32_t *dst_base,
                                          uint32_t fw)
{
    constexpr int PER = 64 / LP;  // ops per step
    const int lane = lane_id();
    const int sub = lane / LP;
    const uint32_t w0 = (uint32_t)(lane % LP);
    const bool word_ok = w0 < fw;
#pragma unroll 1
    for (int s0 = 0; s0 < LP; s0 += U) {
        uint32_t v[U];
        uint64_t d[U];
        bool a[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = (s0 + u) * PER + sub;
            const uint64_t so = __shfl(src_off, j);
            d[u] = __shfl(dst_off, j);
            const uint32_t dl = __shfl(delta, j);
            a[u] = word_ok && ((mask >> j) & 1ull);
            v[u] = a[u] ? src_base[so + w0] + dl : 0u;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (a[u]) dst_base[d[u] + w0] = v[u];
    }
}

// Fields wider than 64 words: one op at a time, 64 words per instruction.
__device__ __forceinline__ void wave_copy_wide(unsigned long long mask, uint64_t src_off,
                                               uint64_t dst_off, uint32_t delta,
                                               const uint32_t *src_base, uint32_t *dst_base,
                                               uint32_t fw)
{
    const int lane = lane_id();
    while (mask) {
        const int j = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const uint64_t so = __shfl(src_off, j);
        const uint64_t d = __shfl(dst_off, j);
        const uint32_t dl = __shfl(delta, j);
        for (uint32_t w = (uint32_t)lane; w < fw; w += 64) dst_base[d + w] = src_base[so + w] + dl;
    }
}

template <int LP>
__device__ __forceinline__ void copy_fields(unsigned long long mask, uint64_t src_off,
                                            uint64_t dst_off, uint32_t delta,
                                            const uint32_t *src_base, uint32_t *dst_base,
                                            uint32_t fw)
{
    if constexpr (LP == 0)
        wave_copy_wide(mask, src_off, dst_off, delta, src_base, dst_base, fw);
    else
        wave_copy<LP, (LP < 8 ? LP : 8)>(mask, src_off, dst_off, delta, src_base, dst_base, fw);
}

// Commit and attempt counts go to one of kCounterStripes stripes (256 bytes
// apart): thousands of blocks adding to one address serialize in the
// memory-side atomic unit (4096 same-address atomics cost ~50 us on MI355X).
__device__ __forceinline__ void add_counts(DeviceCounters *ctr, uint32_t stripe, uint32_t type,
                                           uint32_t commits, uint32_t attempts)
{
    if (commits) {
        atomicAdd(&ctr->commits[stripe][type], commits);
        atomicAdd(&ctr->totals[stripe][type], (unsigned long long)commits);
    }
    if (attempts) atomicAdd(&ctr->totals[stripe][kAttempts + type], (unsigned long long)attempts);
}

// Per-op decode shared by every phase.
struct Op {
    uint64_t row;
    uint32_t kind;
    uint32_t field;
    uint32_t len;  // rows read: 1 for a read, the clipped range for a scan, 0 otherwise
    bool valid;
};

// A slot's op from its row and meta word.  A scan (kind 3) is valid only if
// its length is at most max_scan; kernels that never look at reads pass 0.
__device__ __forceinline__ Op decode_op(uint64_t row, uint32_t m, uint64_t rows, uint32_t fields,
                                        uint32_t max_scan)
{
    Op o;
    o.row = row;
    o.kind = m & 3u;
    o.field = (m >> 8) & 0xffffu;
    const uint32_t sl = (m >> 24) + 1u;
    o.valid = o.kind != LTPG_KV_NONE && o.row < rows && o.field < fields &&
              (o.kind != LTPG_KV_SCAN || sl <= max_scan);
    const uint64_t left = o.row < rows ? rows - o.row : 0ull;
    o.len = !o.valid || o.kind == LTPG_KV_WRITE ? 0u
            : o.kind == LTPG_KV_READ            ? 1u
                                                : (uint32_t)(sl < left ? sl : left);
    return o;
}

// Loads are unconditional on a clamped index (element 0 of a non-empty batch
// when the lane has no op) and the result is selected afterwards: a load
// under a lane-dependent branch sits in its own basic block and the compiler
// waits for it before the next one, which would serialize the misses.
__device__ __forceinline__ Op load_op(const uint64_t *__restrict__ keys,
                                      const uint32_t *__restrict__ meta, uint64_t i, bool live,
                                      uint64_t rows, uint32_t fields, uint32_t max_scan = 0)
{
    const uint64_t j = live ? i : 0ull;
    const uint32_t m0 = meta[j];
    const uint64_t r0 = keys[j];
    return decode_op(live ? r0 : 0ull, live ? m0 : 0u, rows, fields, max_scan);
}

// Batch layouts.  Slot s of transaction t sits at s * B + t (slot-major,
// the reference's LOC[B*op + NO]) or, with rm, at t * S + s (request-major:
// the served form's staging, where a transaction's slots are one contiguous
// run).  Committed reads always go to the slot-major position.
__device__ __forceinline__ uint64_t op_at(uint32_t s, uint32_t t, uint32_t B, uint32_t S, bool rm)
{
    return rm ? (uint64_t)t * S + s : (uint64_t)s * B + t;
}

__device__ __forceinline__ uint32_t txn_of(uint64_t i, uint32_t B, uint32_t S, bool rm)
{
    return rm ? (uint32_t)(i / S) : (uint32_t)(i % B);
}

// Zero the per-epoch commit counts (a kernel rather than a memset so the
// epoch can be captured in a graph).
__global__ void reset_epoch_counters_kernel(DeviceCounters *ctr)
{
    for (uint32_t i = threadIdx.x; i < kCounterStripes * kCountedTypes; i += blockDim.x)
        ctr->commits[i / kCountedTypes][i % kCountedTypes] = 0;
}

// Streaming 16-byte copy, grid-stride.
__global__ void copy_u128_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void fill_u64_kernel(unsigned long long *p, uint64_t n, unsigned long long v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// u32 copy, grid-stride; the source may be pinned host memory (the served
// form's staging: a kernel on the epoch's own stream costs one launch, where
// a copy-engine transfer costs a copy call plus a cross-stream event).
__global__ void copy_u32_kernel(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Publish an epoch's commit bitmaps to pinned host memory, then its sequence
// number with a system-scope release: a host that reads the sequence number
// with an acquire load also reads the bitmaps.  One block.
__global__ void publish_kernel(PublishSet p, uint64_t *seq_out, uint64_t seq)
{
    for (int k = 0; k < 5; k++)
        for (uint32_t w = threadIdx.x; w < p.n[k]; w += blockDim.x) p.dst[k][w] = p.src[k][w];
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void copy_u64_kernel(uint64_t *__restrict__ dst, const uint64_t *__restrict__ src, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Random-access calibration: n random 8-byte loads, or n random probe +
// atomicMin pairs, over a[0..len).  The sink store is never taken for the
// all-ones contents (it keeps the loads from being optimised away).
__device__ __forceinline__ uint64_t calib_index(uint64_t i, uint64_t len)
{
    i ^= i >> 33;
    i *= 0xff51afd7ed558ccdull;
    i ^= i >> 33;
    i *= 0xc4ceb9fe1a85ec53ull;
    i ^= i >> 33;
    return i % len;
}

__global__ void calib_load_kernel(const unsigned long long *a, uint64_t len, uint64_t n, uint32_t per,
                                  unsigned long long *sink)
{
    const uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * per;
    unsigned long long acc = 0;
    for (uint32_t k = 0; k < per; ++k)
        if (base + k < n) acc += a[calib_index(base + k, len)];
    if (acc == 1ull) sink[0] = acc;
}

__global__ void calib_atomic_kernel(unsigned long long *a, uint64_t len, uint64_t n, uint32_t per,
                                    unsigned long long *sink)
{
    const uint64_t base = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * per;
    unsigned long long acc = 0;
    for (uint32_t k = 0; k < per; ++k) {
        if (base + k >= n) break;
        unsigned long long *p = &a[calib_index(base + k, len)];
        const unsigned long long me = base + k;
        if (probe(p) > me) acc += atomicMin(p, me);
    }
    if (acc == 1ull) sink[0] = acc;
}

// Cost of a grid-wide barrier, the step a persistent batched kernel would
// take between phases: gridDim.x workgroups (resident together: the kernel
// is launched cooperatively) pass `reps` barriers.  Lane 0 of each
// workgroup reads the generation, arrives on the counter, and the last to
// arrive resets the counter and advances the generation; the others poll
// the generation.  Every poll is bounded: past spin_limit polls the
// workgroup records an error and leaves, so the kernel always ends.
__global__ void grid_barrier_kernel(unsigned int *count, unsigned int *gen, uint32_t reps, uint32_t spin_limit,
                                    unsigned int *err)
{
    __shared__ uint32_t s_stop;
    for (uint32_t r = 0; r < reps; ++r) {
        if (threadIdx.x == 0) {
            uint32_t stop = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
            const unsigned g = __hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            if (!stop) {
                const unsigned arrived =
                    __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
                if (arrived == gridDim.x - 1) {
                    __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_fetch_add(gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    uint32_t spins = 0;
                    while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
                        if (++spins > spin_limit) {
                            __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            stop = 1;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
            }
            s_stop = stop;
        }
        __syncthreads();
        if (s_stop) return;
    }
}

// One store per workgroup: the smallest kernel, whose back-to-back launches
// time a kernel boundary.
__global__ void boundary_kernel(unsigned int *p) { if (threadIdx.x == 0) p[blockIdx.x] = blockIdx.x; }

// Device wall-clock stamp (constant-rate counter), one lane.
__global__ void stamp_kernel(unsigned long long *out) { out[threadIdx.x] = wall_clock64(); }

__global__ void fill_u8_kernel(uint8_t *p, uint64_t n, uint8_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Reference initial contents: every word of field c holds c
// (YCSB/0/Database.h:77-86 writes Y_c = c into column c); padding words 0.
__global__ void kv_fill_reference_kernel(uint32_t *table, uint64_t n_words, uint32_t fields,
                                         uint32_t fw, uint32_t fsw)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_words; i += stride) {
        const uint64_t f = i / fsw;
        const uint32_t w = (uint32_t)(i - f * fsw);
        table[i] = w < fw ? (uint32_t)(f % fields) : 0u;
    }
}

// Dense fields (field_words apart) <-> strided fields (field_stride apart).
__global__ void repack_kernel(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                              uint64_t n_fields, uint32_t fw, uint32_t fsw, int to_strided)
{
    const uint64_t n = n_fields * fw;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t f = i / fw;
        const uint64_t w = i - f * fw;
        if (to_strided)
            dst[f * fsw + w] = src[i];
        else
            dst[i] = src[f * fsw + w];
    }
}

// Column-major <-> row-major for u32 tables (rows x cols); with T = u64 also
// the request-major -> slot-major transpose of served key-value epochs.
template <typename T>
__global__ void transpose_kernel(const T *__restrict__ src, T *__restrict__ dst, uint64_t rows, uint32_t cols,
                                 int to_row_major)
{
    const uint64_t n = rows * cols;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        // i indexes the destination
        if (to_row_major) {
            const uint64_t r = i / cols, c = i % cols;
            dst[i] = src[c * rows + r];
        } else {
            const uint64_t c = i / rows, r = i % rows;
            dst[i] = src[r * cols + c];
        }
    }
}

// EXECUTE: register every valid op's priority on R (read) or W (write)
// (read_0 / write_0 -> __REGISTER, YCSB/0/Execute.h:11-53,
// Genericfunction.h:61-100) and record waw by displacement.  The load before
// the atomic skips registrations that cannot lower the register (a register
// only decreases, so even a stale cached value is >= its current value); on
// hot rows this turns a serialized chain of same-address atomics into
// cached loads.
//
// Reads register R only for the counting form.  The fast form needs war
// only for transactions without waw, i.e. for the minimum writer of each row
// they write, and a reader with a lower priority than that writer can flag
// it directly from the W register (kv_mark_reads_kernel).  Random 64-bit
// atomics run at about half the rate of random loads on this part, so
// leaving read ops to a load-only pass halves the atomics of an epoch.
__global__ __launch_bounds__(kBlock) void kv_execute_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    unsigned long long *__restrict__ regs, RegLayout rl, uint64_t hot, uint32_t tag,
    uint8_t *__restrict__ doomed, uint8_t epoch, bool reg_reads, bool zero_counts, uint32_t max_scan,
    DeviceCounters *__restrict__ ctr, bool rm)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const bool live = t < B;
    // the epoch's commit counts start at zero (CHECK adds to them)
    if (zero_counts && blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += kBlock)
            ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    const unsigned long long me = live ? priority(tag, tids[t], t) : 0ull;
    bool waw = false;
    uint32_t bad = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const Op o = load_op(keys, meta, op_at(s, t, B, S, rm), live, rows, fields, max_scan);
        bad += (o.kind != LTPG_KV_NONE && !o.valid) ? 1u : 0u;
        if (!o.valid) continue;
        if (o.kind == LTPG_KV_WRITE) {
            if (o.row != hot)  // the hot row never registers
                waw |= register_write(&regs[rl.w(o.row)], me, tag, doomed, epoch);
        } else if (reg_reads) {
            for (uint32_t k = 0; k < o.len; ++k)
                if (o.row + k != hot) register_read(&regs[rl.r(o.row + k)], me);
        }
    }
    if (waw) doomed[t] = epoch;
    if (bad) atomicAdd(&ctr->range_errors, bad);
}

// EXECUTE, one lane per op (batch index i) instead of one
// lane per transaction: S times the lanes, each with a single registration,
// for small or ragged batches that would leave the chip idle.
__global__ __launch_bounds__(kBlock) void kv_execute_ops_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    unsigned long long *__restrict__ regs, RegLayout rl, uint64_t hot, uint32_t tag,
    uint8_t *__restrict__ doomed, uint8_t epoch, bool reg_reads, bool zero_counts, uint32_t max_scan,
    DeviceCounters *__restrict__ ctr, bool rm)
{
    if (zero_counts && blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += kBlock)
            ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    const uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < (uint64_t)S * B;
    const uint32_t t = live ? txn_of(i, B, S, rm) : 0u;
    const Op o = load_op(keys, meta, live ? i : 0, live, rows, fields, max_scan);
    const bool bad = o.kind != LTPG_KV_NONE && !o.valid;
    if (o.valid) {
        const unsigned long long me = priority(tag, tids[t], t);
        if (o.kind == LTPG_KV_WRITE) {
            if (o.row != hot && register_write(&regs[rl.w(o.row)], me, tag, doomed, epoch)) doomed[t] = epoch;
        } else if (reg_reads) {
            for (uint32_t k = 0; k < o.len; ++k)
                if (o.row + k != hot) register_read(&regs[rl.r(o.row + k)], me);
        }
    }
    if (bad) atomicAdd(&ctr->range_errors, 1u);
}

// EXECUTE, fast form, one lane per op with the block's duplicate rows
// folded first.  Under skew thousands of in-flight writes hit the same hot
// register at once, all see it stale, and their atomics serialize on one
// address.  Here each block takes P * kBlock consecutive ops (P per lane),
// finds the minimum writer priority of every row among them in an LDS hash
// table, and only that writer goes to the global register; the block's
// other writers of the row already have waw (a lower writer exists).  Every
// non-minimum writer of a row is still doomed exactly once.
// P (ops per lane) is 2, or 4 with LTPG_TUNE_MORE_ITEMS: a wider fold and
// more loads in flight per lane, against half the blocks.
template <uint32_t P>
__global__ __launch_bounds__(kBlock) void kv_execute_dedup_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    unsigned long long *__restrict__ regs, RegLayout rl, uint64_t hot, uint32_t tag,
    uint8_t *__restrict__ doomed, uint8_t epoch, bool zero_counts, uint32_t max_scan,
    DeviceCounters *__restrict__ ctr, bool rm)
{
    constexpr uint32_t kDedupOps = P * kBlock;
    constexpr uint32_t kDedupSlots = 2 * kDedupOps;  // load factor <= 1/2: probing always ends
    __shared__ unsigned long long s_key[kDedupSlots];  // row + 1, 0 = empty
    __shared__ unsigned long long s_min[kDedupSlots];
    if (zero_counts && blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += kBlock)
            ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    for (uint32_t k = threadIdx.x; k < kDedupSlots; k += kBlock) {
        s_key[k] = 0ull;
        s_min[k] = ~0ull;
    }
    __syncthreads();
    const uint64_t n_ops = (uint64_t)S * B;
    const uint64_t base = (uint64_t)blockIdx.x * kDedupOps;
    uint64_t row[P];
    unsigned long long me[P];
    uint32_t tt[P], slot[P];
    bool wr[P];
    uint32_t bad = 0;
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        const uint64_t i = base + (uint64_t)k * kBlock + threadIdx.x;
        const bool live = i < n_ops;
        const uint32_t t = live ? txn_of(i, B, S, rm) : 0u;
        const Op o = load_op(keys, meta, live ? i : 0, live, rows, fields, max_scan);
        bad += (o.kind != LTPG_KV_NONE && !o.valid) ? 1u : 0u;
        wr[k] = o.valid && o.kind == LTPG_KV_WRITE && o.row != hot;
        row[k] = o.row;
        tt[k] = t;
        me[k] = wr[k] ? priority(tag, tids[t], t) : 0ull;
        slot[k] = 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        if (!wr[k]) continue;
        const unsigned long long key = row[k] + 1ull;
        uint32_t h = (uint32_t)((row[k] * 0x9E3779B97F4A7C15ull) >> 40) & (kDedupSlots - 1);
        while (true) {
            const unsigned long long prev = atomicCAS(&s_key[h], 0ull, key);
            if (prev == 0ull || prev == key) break;
            h = (h + 1) & (kDedupSlots - 1);
        }
        slot[k] = h;
        atomicMin(&s_min[h], me[k]);
    }
    __syncthreads();
    // register_write for the block minima, split into phases so that the
    // four probes, then the four atomics, are in flight together
    bool lead[P], waw[P];
    unsigned long long cur[P], old[P];
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        waw[k] = wr[k] && s_min[slot[k]] < me[k];
        lead[k] = wr[k] && !waw[k];
        cur[k] = lead[k] ? probe(&regs[rl.w(row[k])]) : ~0ull;
    }
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        old[k] = ~0ull;
        if (lead[k] && cur[k] > me[k]) old[k] = atomicMin(&regs[rl.w(row[k])], me[k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < P; ++k) {
        if (!lead[k]) continue;
        if (cur[k] < me[k] || old[k] < me[k]) {
            waw[k] = true;
        } else if (cur[k] > me[k] && old[k] > me[k] && old[k] != ~0ull &&
                   (uint32_t)(old[k] >> kTagShift) == tag) {
            doomed[old[k] & kPosMask] = epoch;  // displaced: that writer has waw
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < P; ++k)
        if (waw[k]) doomed[tt[k]] = epoch;
    if (bad) atomicAdd(&ctr->range_errors, bad);
}

// ---- row-partitioned EXECUTE (LTPG_TUNE_PARTITIONED, fast form) -----------
//
// The register phase touches one random register line per operation; the
// lines of a 10M-row table overflow the L2s, so each 8-byte probe or atomic
// moves a 128-byte line.  Here the epoch's register operations (every write
// op, every row a read or scan op reads; hot-row ops never register) are
// first binned by row range into buckets of 2^kPartShift rows:
//   count    streaming pass over the ops: per-bucket sizes (LDS histogram,
//            one global add per bucket and block), range errors;
//   scan     one workgroup: bucket starts, cursors, and slices of at most
//            kPartSlice operations (a skewed bucket gets several);
//   scatter  second pass over the ops: each operation's (priority, row in
//            bucket, is-write) to its bucket, a block reserving one run per
//            bucket;
//   process  one workgroup per slice: the W minima of the bucket's rows in
//            LDS (probe, then a 64-bit LDS atomicMin), then the flags of the
//            fast form from them -- a writer behind the minimum is doomed; a
//            reader behind it has raw; a reader ahead of it gives the minimum
//            writer war.  A bucket split over several slices folds each
//            slice's minima into the global W registers instead;
//   flags    one workgroup per slice of such a bucket: the same flags from
//            the global W registers.
// The flags equal those of kv_execute_kernel + kv_mark_reads_kernel: both
// compare every operation with the minimum W priority of its row.
constexpr uint32_t kPartBlock = 1024;
constexpr uint32_t kPartPer = 16;  // operations per thread in the binning passes
constexpr uint32_t kPartChunk = kPartBlock * kPartPer;
constexpr uint32_t kPartRows = 1u << kPartShift;

struct PartArgs {
    const uint64_t *keys;
    const uint32_t *meta;
    const uint32_t *tids;
    uint32_t B, S;
    uint64_t rows;
    uint32_t fields, max_scan;
    uint64_t hot;
    uint32_t nb, cap_slices;
    uint32_t *count, *start, *cursor, *slices;
    unsigned long long *ent_me;
    uint32_t *ent_key;
    unsigned long long *regs;
    RegLayout rl;
    uint32_t tag;
    uint8_t epoch;
    bool rm;
    uint8_t *doomed, *raw_flag, *war_flag;
    DeviceCounters *ctr;
};

// kPartGroup consecutive op slots of a thread (stride kPartBlock), all
// loads issued before any is used.
constexpr uint32_t kPartGroup = 4;

template <uint32_t G>
__device__ __forceinline__ void part_load(const PartArgs &a, uint64_t first, Op (&o)[G])
{
    const uint64_t n_ops = (uint64_t)a.S * a.B;
    uint64_t r[G];
    uint32_t m[G];
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) {
        const uint64_t i = first + (uint64_t)u * kPartBlock;
        const uint64_t j = i < n_ops ? i : 0ull;
        r[u] = a.keys[j];
        m[u] = a.meta[j];
    }
#pragma unroll
    for (uint32_t u = 0; u < G; ++u) {
        const bool live = first + (uint64_t)u * kPartBlock < n_ops;
        o[u] = decode_op(live ? r[u] : 0ull, live ? m[u] : 0u, a.rows, a.fields, a.max_scan);
    }
}

__global__ __launch_bounds__(kPartBlock) void kv_part_count_kernel(PartArgs a, bool zero_counts)
{
    __shared__ uint32_t s_hist[kPartMaxBuckets];
    if (zero_counts && blockIdx.x == 0)
        for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += kPartBlock)
            a.ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    for (uint32_t b = threadIdx.x; b < a.nb; b += kPartBlock) s_hist[b] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kPartChunk + threadIdx.x;
    uint32_t bad = 0;
#pragma unroll 1
    for (uint32_t g = 0; g < kPartPer; g += kPartGroup) {
        Op o[kPartGroup];
        part_load(a, base + (uint64_t)g * kPartBlock, o);
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u) {
            bad += (o[u].kind != LTPG_KV_NONE && !o[u].valid) ? 1u : 0u;
            if (!o[u].valid) continue;
            if (o[u].kind == LTPG_KV_WRITE) {
                if (o[u].row != a.hot) atomicAdd(&s_hist[o[u].row >> kPartShift], 1u);
            } else {
                for (uint32_t j = 0; j < o[u].len; ++j)
                    if (o[u].row + j != a.hot) atomicAdd(&s_hist[(o[u].row + j) >> kPartShift], 1u);
            }
        }
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < a.nb; b += kPartBlock) {
        const uint32_t c = s_hist[b];
        if (c) atomicAdd(&a.count[b], c);
    }
    if (bad) atomicAdd(&a.ctr->range_errors, bad);
}

// Exclusive scan of the bucket sizes into starts, cursors and slices (one
// workgroup; each thread owns a run of consecutive buckets).  Clears the
// sizes for the next epoch.
__global__ __launch_bounds__(kPartBlock) void kv_part_scan_kernel(PartArgs a)
{
    __shared__ uint32_t s_n[kPartBlock], s_sl[kPartBlock];
    const uint32_t per = (a.nb + kPartBlock - 1) / kPartBlock;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t n = 0, ns = 0;
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t b = b0 + j;
        if (b < a.nb) {
            const uint32_t c = a.count[b];
            n += c;
            ns += (c + kPartSlice - 1) / kPartSlice;
        }
    }
    s_n[threadIdx.x] = n;
    s_sl[threadIdx.x] = ns;
    __syncthreads();
    // inclusive scans (Hillis-Steele over the block)
    for (uint32_t d = 1; d < kPartBlock; d <<= 1) {
        const uint32_t x = threadIdx.x >= d ? s_n[threadIdx.x - d] : 0u;
        const uint32_t y = threadIdx.x >= d ? s_sl[threadIdx.x - d] : 0u;
        __syncthreads();
        s_n[threadIdx.x] += x;
        s_sl[threadIdx.x] += y;
        __syncthreads();
    }
    uint32_t at = s_n[threadIdx.x] - n, sl = s_sl[threadIdx.x] - ns;
    for (uint32_t j = 0; j < per; ++j) {
        const uint32_t b = b0 + j;
        if (b >= a.nb) break;
        const uint32_t c = a.count[b];
        a.start[b] = at;
        a.cursor[b] = at;
        for (uint32_t k = 0; k * kPartSlice < c; ++k, ++sl) {
            if (sl >= a.cap_slices) continue;
            a.slices[3 * sl] = b;
            a.slices[3 * sl + 1] = at + k * kPartSlice;
            a.slices[3 * sl + 2] = at + min(c, (k + 1) * kPartSlice);
        }
        at += c;
        a.count[b] = 0;
    }
    if (threadIdx.x == kPartBlock - 1) {
        a.start[a.nb] = s_n[kPartBlock - 1];
        a.slices[3 * a.cap_slices] = min(s_sl[kPartBlock - 1], a.cap_slices);
    }
}

__global__ __launch_bounds__(kPartBlock) void kv_part_scatter_kernel(PartArgs a)
{
    __shared__ uint32_t s_at[kPartMaxBuckets];
    for (uint32_t b = threadIdx.x; b < a.nb; b += kPartBlock) s_at[b] = 0;
    __syncthreads();
    const uint64_t base = (uint64_t)blockIdx.x * kPartChunk + threadIdx.x;
#pragma unroll 1
    for (uint32_t g = 0; g < kPartPer; g += kPartGroup) {
        Op o[kPartGroup];
        part_load(a, base + (uint64_t)g * kPartBlock, o);
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u) {
            if (!o[u].valid) continue;
            if (o[u].kind == LTPG_KV_WRITE) {
                if (o[u].row != a.hot) atomicAdd(&s_at[o[u].row >> kPartShift], 1u);
            } else {
                for (uint32_t j = 0; j < o[u].len; ++j)
                    if (o[u].row + j != a.hot) atomicAdd(&s_at[(o[u].row + j) >> kPartShift], 1u);
            }
        }
    }
    __syncthreads();
    // one run per bucket: the block's count becomes its run's start
    for (uint32_t b = threadIdx.x; b < a.nb; b += kPartBlock) {
        const uint32_t c = s_at[b];
        if (c) s_at[b] = atomicAdd(&a.cursor[b], c);
    }
    __syncthreads();
    const uint64_t n_ops = (uint64_t)a.S * a.B;
    constexpr uint32_t G = kPartGroup;
#pragma unroll 1
    for (uint32_t g = 0; g < kPartPer; g += G) {
        Op o[G];
        unsigned long long me[G];
        part_load(a, base + (uint64_t)g * kPartBlock, o);
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) {
            // op indices stay below 2^32 (kv_part_prepare): 32-bit division
            const uint64_t i = base + (uint64_t)(g + u) * kPartBlock;
            const uint32_t i32 = i < n_ops ? (uint32_t)i : 0u;
            const uint32_t t = a.rm ? i32 / a.S : i32 % a.B;
            me[u] = priority(a.tag, a.tids[t], t);
        }
#pragma unroll
        for (uint32_t u = 0; u < G; ++u) {
            if (!o[u].valid) continue;
            if (o[u].kind == LTPG_KV_WRITE) {
                if (o[u].row == a.hot) continue;
                const uint32_t at = atomicAdd(&s_at[o[u].row >> kPartShift], 1u);
                a.ent_me[at] = me[u];
                a.ent_key[at] = ((uint32_t)(o[u].row & (kPartRows - 1)) << 1) | 1u;
            } else {
                for (uint32_t j = 0; j < o[u].len; ++j) {
                    const uint64_t r = o[u].row + j;
                    if (r == a.hot) continue;
                    const uint32_t at = atomicAdd(&s_at[r >> kPartShift], 1u);
                    a.ent_me[at] = me[u];
                    a.ent_key[at] = (uint32_t)(r & (kPartRows - 1)) << 1;
                }
            }
        }
    }
}

// Binning of point operations (no scans) through LDS: a block takes
// P ops per thread, ranks each in its bucket with an LDS counter, turns the
// counts into block-local starts, reserves one run per bucket in the global
// bucket with one atomic, places the entries bucket by bucket in LDS and
// writes them out in that order -- consecutive lanes write consecutive
// entries of a run, so the stores coalesce instead of scattering 8- and
// 4-byte writes over every bucket.  MaxB bounds the buckets.  Runs are as
// long as the block's share of a bucket, so a block takes as many ops as
// its LDS holds: 8 per thread (122 KB) with at most kPartFewBuckets
// buckets, 4 per thread beside the 64 KB of counters for 16384.
constexpr uint32_t kPartFewBuckets = 2048;

// Exclusive prefix sum over the block of one value per thread; *total gets
// the sum.  Wave scans by shuffles, then the 16 wave totals through LDS.
__device__ __forceinline__ uint32_t block_exclusive_sum(uint32_t v, uint32_t *s_wt, uint32_t *total)
{
    constexpr uint32_t kWaves = kPartBlock / 64;
    const uint32_t lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    uint32_t inc = v;
#pragma unroll
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t x = __shfl_up(inc, d);
        if (lane >= d) inc += x;
    }
    if (lane == 63) s_wt[wave] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t w = 0; w < kWaves; ++w) {
        const uint32_t t = s_wt[w];
        before += w < wave ? t : 0u;
        all += t;
    }
    *total = all;
    return before + inc - v;
}

template <uint32_t MaxB, uint32_t P>
__global__ __launch_bounds__(kPartBlock) void kv_part_scatter_staged_kernel(PartArgs a)
{
    constexpr uint32_t kPer = (MaxB + kPartBlock - 1) / kPartBlock;  // buckets per thread at most
    __shared__ uint32_t s_at[MaxB];              // counts, then local starts, then global - local
    __shared__ unsigned long long st_me[P * kPartBlock];
    __shared__ uint32_t st_key[P * kPartBlock];
    __shared__ uint16_t st_b[P * kPartBlock];
    __shared__ uint32_t s_wt[kPartBlock / 64];
    for (uint32_t b = threadIdx.x; b < a.nb; b += kPartBlock) s_at[b] = 0;
    __syncthreads();
    const uint64_t n_ops = (uint64_t)a.S * a.B;
    const uint64_t base = (uint64_t)blockIdx.x * (P * kPartBlock) + threadIdx.x;
    Op o[P];
    part_load(a, base, o);
    uint32_t bk[P], rank[P];
    bool use[P];
    unsigned long long me[P];
#pragma unroll
    for (uint32_t u = 0; u < P; ++u) {
        use[u] = o[u].valid && o[u].len <= 1 && o[u].row != a.hot;
        bk[u] = (uint32_t)(o[u].row >> kPartShift);
        const uint64_t i = base + (uint64_t)u * kPartBlock;
        const uint32_t i32 = i < n_ops ? (uint32_t)i : 0u;
        const uint32_t t = a.rm ? i32 / a.S : i32 % a.B;
        me[u] = priority(a.tag, a.tids[t], t);
    }
#pragma unroll
    for (uint32_t u = 0; u < P; ++u)
        rank[u] = use[u] ? atomicAdd(&s_at[bk[u]], 1u) : 0u;
    __syncthreads();
    // thread j owns buckets [j * per, (j + 1) * per): their counts, then
    // their local starts from the block-wide exclusive sum
    const uint32_t per = (a.nb + kPartBlock - 1) / kPartBlock;
    const uint32_t b0 = threadIdx.x * per;
    uint32_t cnt[kPer], st[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        cnt[j] = (j < per && b0 + j < a.nb) ? s_at[b0 + j] : 0u;
        sum += cnt[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_sum(sum, s_wt, &total);
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) {
        st[j] = run;
        run += cnt[j];
    }
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j)
        if (cnt[j]) s_at[b0 + j] = st[j];
    __syncthreads();
    // this thread's entries go to their bucket's local run
    uint32_t slot[P];
#pragma unroll
    for (uint32_t u = 0; u < P; ++u) slot[u] = use[u] ? s_at[bk[u]] + rank[u] : 0u;
    // one global run per non-empty bucket; s_at[b] becomes global - local
    uint32_t delta[kPer];
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j) delta[j] = cnt[j] ? atomicAdd(&a.cursor[b0 + j], cnt[j]) - st[j] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < P; ++u) {
        if (!use[u]) continue;
        st_me[slot[u]] = me[u];
        st_key[slot[u]] = ((uint32_t)(o[u].row & (kPartRows - 1)) << 1) | (o[u].kind == LTPG_KV_WRITE ? 1u : 0u);
        st_b[slot[u]] = (uint16_t)bk[u];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kPer; ++j)
        if (cnt[j]) s_at[b0 + j] = delta[j];
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < total; j += kPartBlock) {
        const uint32_t at = j + s_at[st_b[j]];
        a.ent_me[at] = st_me[j];
        a.ent_key[at] = st_key[j];
    }
}

// The fast form's flags of one operation against its row's minimum W
// priority w (~0: no writer this epoch).
__device__ __forceinline__ void part_flag(const PartArgs &a, uint32_t key, unsigned long long me,
                                          unsigned long long w)
{
    const uint32_t pos = (uint32_t)(me & kPosMask);
    if (key & 1u) {
        if (w < me) a.doomed[pos] = a.epoch;
    } else if (w != ~0ull) {
        if (w < me)
            a.raw_flag[pos] = a.epoch;
        else if (w > me)
            a.war_flag[w & kPosMask] = a.epoch;
    }
}

__global__ __launch_bounds__(kPartBlock) void kv_part_process_kernel(PartArgs a)
{
    __shared__ unsigned long long s_w[kPartRows];
    if (blockIdx.x >= a.slices[3 * a.cap_slices]) return;
    const uint32_t b = a.slices[3 * blockIdx.x], beg = a.slices[3 * blockIdx.x + 1], end = a.slices[3 * blockIdx.x + 2];
    const bool multi = a.start[b + 1] - a.start[b] > kPartSlice;
    for (uint32_t r = threadIdx.x; r < kPartRows; r += kPartBlock) s_w[r] = ~0ull;
    __syncthreads();
    for (uint32_t e0 = beg + threadIdx.x; e0 < end; e0 += kPartGroup * kPartBlock) {
        uint32_t key[kPartGroup];
        unsigned long long me[kPartGroup];
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u) {
            const uint32_t e = e0 + u * kPartBlock;
            key[u] = e < end ? a.ent_key[e] : 0u;
            me[u] = e < end ? a.ent_me[e] : ~0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u)
            if ((key[u] & 1u) && s_w[key[u] >> 1] > me[u]) atomicMin(&s_w[key[u] >> 1], me[u]);
    }
    __syncthreads();
    if (multi) {
        const uint64_t row0 = (uint64_t)b << kPartShift;
        for (uint32_t r = threadIdx.x; r < kPartRows; r += kPartBlock) {
            const unsigned long long v = s_w[r];
            if (v == ~0ull) continue;
            unsigned long long *p = &a.regs[a.rl.w(row0 + r)];
            if (probe(p) > v) atomicMin(p, v);
        }
        return;
    }
    for (uint32_t e0 = beg + threadIdx.x; e0 < end; e0 += kPartGroup * kPartBlock) {
        uint32_t key[kPartGroup];
        unsigned long long me[kPartGroup];
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u) {
            const uint32_t e = e0 + u * kPartBlock;
            key[u] = e < end ? a.ent_key[e] : 0u;
            me[u] = e < end ? a.ent_me[e] : ~0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u)
            if (e0 + u * kPartBlock < end) part_flag(a, key[u], me[u], s_w[key[u] >> 1]);
    }
}

// The bucket's final W minima are read from the registers once, in row
// order, into LDS; each operation then finds its row's minimum there.
__global__ __launch_bounds__(kPartBlock) void kv_part_flags_kernel(PartArgs a)
{
    __shared__ unsigned long long s_w[kPartRows];
    if (blockIdx.x >= a.slices[3 * a.cap_slices]) return;
    const uint32_t b = a.slices[3 * blockIdx.x], beg = a.slices[3 * blockIdx.x + 1], end = a.slices[3 * blockIdx.x + 2];
    if (a.start[b + 1] - a.start[b] <= kPartSlice) return;  // decided in kv_part_process_kernel
    const uint64_t row0 = (uint64_t)b << kPartShift;
    for (uint32_t r = threadIdx.x; r < kPartRows; r += kPartBlock) {
        unsigned long long w = ~0ull;
        if (row0 + r < a.rows) {
            w = a.regs[a.rl.w(row0 + r)];
            if ((uint32_t)(w >> kTagShift) != a.tag) w = ~0ull;  // no writer this epoch
        }
        s_w[r] = w;
    }
    __syncthreads();
    for (uint32_t e0 = beg + threadIdx.x; e0 < end; e0 += kPartGroup * kPartBlock) {
        uint32_t key[kPartGroup];
        unsigned long long me[kPartGroup];
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u) {
            const uint32_t e = e0 + u * kPartBlock;
            key[u] = e < end ? a.ent_key[e] : 0u;
            me[u] = e < end ? a.ent_me[e] : ~0ull;
        }
#pragma unroll
        for (uint32_t u = 0; u < kPartGroup; ++u)
            if (e0 + u * kPartBlock < end) part_flag(a, key[u], me[u], s_w[key[u] >> 1]);
    }
}

// Read marking (fast form), one lane per op after every write has
// registered.  A read of row x by t against w = W[x] of this epoch:
//   w < me: t reads a row a higher-priority transaction writes -> raw(t);
//   w > me: the row's minimum writer (position bits of w) writes a row a
//           higher-priority transaction reads -> war for that writer.
// A writer with waw aborts anyway, so flagging only the minimum writer
// gives every transaction that can still commit its exact war bit; raw(t)
// is exact for every t.  Flags hold the epoch number, so equal stores race
// harmlessly and nothing is cleared between epochs.
template <uint32_t P>
__global__ __launch_bounds__(kBlock) void kv_mark_reads_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    const unsigned long long *__restrict__ regs, RegLayout rl, uint64_t hot, uint32_t tag,
    uint8_t *__restrict__ raw_flag, uint8_t *__restrict__ war_flag, uint8_t epoch, uint32_t max_scan, bool rm)
{
    // P ops per lane (1, or 2 with LTPG_TUNE_MORE_ITEMS), stride kBlock; all
    // their loads issued before any is used
    const uint64_t base = (uint64_t)blockIdx.x * (P * kBlock) + threadIdx.x;
    const uint64_t n_ops = (uint64_t)S * B;
    Op o[P];
    uint32_t t[P];
#pragma unroll
    for (uint32_t u = 0; u < P; ++u) {
        const uint64_t i = base + (uint64_t)u * kBlock;
        const bool live = i < n_ops;
        t[u] = live ? txn_of(i, B, S, rm) : 0u;
        o[u] = load_op(keys, meta, live ? i : 0, live, rows, fields, max_scan);
    }
#pragma unroll
    for (uint32_t u = 0; u < P; ++u) {
        if (o[u].len == 0) continue;  // writes, empty and invalid slots
        const unsigned long long me = priority(tag, tids[t[u]], t[u]);
        bool raw = false;
        for (uint32_t k = 0; k < o[u].len; ++k) {  // a scan: every row of its range
            if (o[u].row + k == hot) continue;
            const unsigned long long w = regs[rl.w(o[u].row + k)];
            if ((uint32_t)(w >> kTagShift) != tag) continue;  // no writer this epoch
            if (w < me)
                raw = true;
            else if (w > me)
                war_flag[w & kPosMask] = epoch;
        }
        if (raw) raw_flag[t[u]] = epoch;
    }
}

// CHECK + COMMIT, counting form (check_r / check_w_0 / check_w_1 and
// __CHECK, YCSB/0/Execute.h:109-175, Genericfunction.h:101-144; writeback_0,
// Execute.h:184-191): every transaction counts raw (reads whose W register
// holds a higher priority), war (writes vs R) and waw (writes vs W) over all
// its ops -- the reference's counters, kept for parity checks of them --
// and commits iff waw == 0 && (raw == 0 || war == 0).
__global__ __launch_bounds__(kBlock) void kv_check_counting_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    const unsigned long long *__restrict__ regs, RegLayout rl, uint64_t hot, uint32_t tag,
    uint64_t *__restrict__ bitmap, uint32_t *__restrict__ conflicts, uint32_t max_scan,
    DeviceCounters *__restrict__ ctr, bool rm)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const int lane = lane_id();
    const bool live = t < B;
    const unsigned long long me = live ? priority(tag, tids[t], t) : 0ull;
    uint32_t raw = 0, war = 0, waw = 0;
    for (uint32_t s = 0; s < S; ++s) {
        const Op o = load_op(keys, meta, op_at(s, t, B, S, rm), live, rows, fields, max_scan);
        if (!o.valid) continue;
        if (o.kind != LTPG_KV_WRITE) {
            for (uint32_t k = 0; k < o.len; ++k)  // the hot row never conflicts
                raw += (o.row + k != hot && regs[rl.w(o.row + k)] < me) ? 1u : 0u;
        } else if (o.row != hot) {
            war += regs[rl.r(o.row)] < me ? 1u : 0u;
            waw += regs[rl.w(o.row)] < me ? 1u : 0u;
        }
    }
    const bool commit = live && waw == 0 && (raw == 0 || war == 0);
    if (live) conflicts[t] = raw | (war << 8) | (waw << 16);
    const unsigned long long word = __ballot(commit);
    const uint32_t wave_base = t - (uint32_t)lane;
    if (lane == 0 && wave_base < B) {
        bitmap[wave_base / 64] = word;
        const uint32_t n = (uint32_t)__popcll(word);
        const uint32_t live_n = (B - wave_base) < 64u ? (B - wave_base) : 64u;
        add_counts(ctr, (blockIdx.x * (kBlock / 64) + threadIdx.x / 64) % kCounterStripes, LTPG_TYPE_KV, n, live_n);
    }
}

// READOUT: committed transactions' read fields, gathered from the snapshot
// before any write-back of the epoch (read_1 / __READ, YCSB/0/Execute.h:
// 27-37, Genericfunction.h:31-45).  Reads of aborted transactions are never
// delivered, so they are not gathered.
template <int LP>
__global__ __launch_bounds__(kBlock) void kv_readout_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta, uint32_t B, uint32_t S,
    uint64_t rows, uint32_t fields, uint32_t fw, uint32_t fsw, const uint32_t *__restrict__ table,
    const uint64_t *__restrict__ bitmap, uint32_t *__restrict__ read_values, uint32_t max_scan, uint32_t L,
    bool rm)
{
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const int lane = lane_id();
    const uint32_t wave_base = t - (uint32_t)lane;
    if (wave_base >= B) return;
    const uint64_t word = bitmap[wave_base / 64];
    if (word == 0) return;
    const bool committed = t < B && ((word >> lane) & 1ull);
    for (uint32_t s = 0; s < S; ++s) {
        const uint64_t i = (uint64_t)s * B + t;  // slot-major: where its reads go
        const Op o = load_op(keys, meta, op_at(s, t, B, S, rm), committed, rows, fields, max_scan);
        for (uint32_t k = 0;; ++k) {  // row k of each read / scan
            const unsigned long long mask = __ballot(k < o.len);
            if (!mask) break;
            copy_fields<LP>(mask, ((o.row + k) * fields + o.field) * fsw, (i * L + k) * fw, 0u, table,
                            read_values, fw);
        }
    }
}

// Block-local compaction.  A chunk is CH consecutive transactions (CH =
// 256, or 64 as a launch shape); the transactions that still have work
// (committed ones, for the read-out of CHECK and for WRITEBACK) are gathered
// into an LDS list, then processed G lanes per transaction (one lane per
// slot), so lanes stay busy whatever the abort rate and every op load of a
// transaction is issued at once.  A wave walks the list in rounds of 64 / G
// transactions; for small batches gridDim.y blocks share one chunk's list
// (each builds the same list and takes every gridDim.y-th round), so that a
// small epoch runs as many short waves rather than a few long ones.
constexpr uint32_t kChunk = kBlock;

// Gathers the picked transactions (one per thread) into s_list in thread
// order, so that every block of a chunk builds the same list and can take
// its own share of it by index.  Every thread of the block calls it (it
// holds two barriers); s_cnt has one word per wave.  Returns the length.
__device__ __forceinline__ uint32_t list_build(bool pick, uint32_t t, uint32_t *s_list, uint32_t *s_cnt)
{
    const int lane = lane_id();
    const uint32_t wave = threadIdx.x / 64;
    const unsigned long long m = __ballot(pick);
    if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0, n = 0;
#pragma unroll
    for (uint32_t w = 0; w < kBlock / 64; ++w) {
        const uint32_t c = s_cnt[w];
        base += w < wave ? c : 0u;
        n += c;
    }
    if (pick) s_list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = t;
    __syncthreads();
    return n;
}

// Position of the j-th set bit of m (j < popcount(m)), by halving.
__device__ __forceinline__ uint32_t nth_set_bit(unsigned long long m, uint32_t j)
{
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t width = 32; width >= 1; width >>= 1) {
        const unsigned long long lo = m & ((1ull << width) - 1ull);
        const uint32_t c = (uint32_t)__popcll(lo);
        if (j >= c) {
            j -= c;
            m >>= width;
            pos += width;
        } else {
            m = lo;
        }
    }
    return pos;
}

// Field copies of the wave's flagged lanes, fields of at most 32 words, two
// fields per instruction: lanes 0-31 copy the field of one flagged lane and
// lanes 32-63 the next, one word per lane, so a 100-byte field is a single
// coalesced access.  Flagged lane k: src[src_off + w] (+ delta) ->
// dst[dst_off + w], w < fw.  Up to 32 fields' loads are issued before their
// stores, which requires that no two flagged fields of one call write the
// same word (callers guarantee it).
__device__ __forceinline__ void wave_copy_pairs(unsigned long long mask, uint64_t src_off, uint64_t dst_off,
                                                uint32_t delta, const uint32_t *src, uint32_t *dst, uint32_t fw)
{
    const uint32_t n = (uint32_t)__popcll(mask);
    const uint32_t half = (uint32_t)lane_id() >> 5, wl = (uint32_t)lane_id() & 31u;
    for (uint32_t b = 0; b < n; b += 32) {
        uint32_t v[16], dl[16];
        uint64_t d[16];
        bool ok[16];
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u) {
            const uint32_t j = b + 2 * u + half;
            const bool has = j < n;
            const int owner = has ? (int)nth_set_bit(mask, j) : 0;
            const uint64_t so = __shfl(src_off, owner);
            d[u] = __shfl(dst_off, owner);
            dl[u] = __shfl(delta, owner);
            ok[u] = has && wl < fw;
            // unconditional load (word 0 when idle), so all 16 are in flight
            v[u] = src[ok[u] ? so + wl : 0ull];
        }
#pragma unroll
        for (uint32_t u = 0; u < 16; ++u)
            if (ok[u]) dst[d[u] + wl] = v[u] + dl[u];
    }
}

// Random-gather ceiling: n fields of fw words, each from a random slot
// (stride_w words apart) of src, copied to dst[i * fw ..] -- the access
// pattern of CHECK's read-out -- with wave_copy_pairs, 64 fields per wave
// step.
__global__ __launch_bounds__(kBlock) void calib_gather_kernel(const uint32_t *__restrict__ src, uint64_t slots,
                                                             uint32_t stride_w, uint32_t fw,
                                                             uint32_t *__restrict__ dst, uint64_t n, uint32_t per)
{
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    for (uint32_t k = 0; k < per; ++k) {
        const uint64_t i = (wave * per + k) * 64 + (uint64_t)lane_id();
        const bool has = i < n;
        const uint64_t so = has ? calib_index(i, slots) * stride_w : 0ull;
        wave_copy_pairs(__ballot(has), so, has ? i * fw : 0ull, 0u, src, dst, fw);
    }
}

// Per-wave staging of field copies.  A wave walks its transactions in
// rounds of 64 / G; a round of a low-contention batch flags only a few
// fields, and copying them right away leaves a wave with one round's fields
// in flight.  Instead each round appends its flagged fields (source,
// destination, delta) to the wave's own LDS list, and the list is copied
// once it holds 64 or more (or at the end): wave_copy_pairs then has up to 32
// fields in flight per step, and rounds of op loads run back to back.  The
// list holds a round's worth more than the flush threshold.  Only the wave
// that owns a list touches it, so a wavefront-scope fence orders its lanes'
// LDS accesses.
constexpr uint32_t kStage = 128;

struct StageList {
    unsigned long long src[kStage];
    unsigned long long dst[kStage];
    uint32_t delta[kStage];
};

__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// Append the flagged lanes' copies; returns the new (wave-uniform) length.
__device__ __forceinline__ uint32_t stage_push(StageList &l, uint32_t n, bool flag, uint64_t src_off,
                                               uint64_t dst_off, uint32_t delta)
{
    const unsigned long long m = __ballot(flag);
    if (flag) {
        const uint32_t at = n + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull));
        l.src[at] = src_off;
        l.dst[at] = dst_off;
        l.delta[at] = delta;
    }
    return n + (uint32_t)__popcll(m);
}

// Copy the wave's n staged fields (fw <= 32) and empty the list.  The
// staged fields never overlap: callers stage distinct destinations.
__device__ __forceinline__ void stage_flush(StageList &l, uint32_t n, const uint32_t *src, uint32_t *dst,
                                            uint32_t fw)
{
    wave_lds_fence();
    for (uint32_t b = 0; b < n; b += 64) {
        const uint32_t e = b + (uint32_t)lane_id();
        const bool has = e < n;
        const uint64_t so = has ? l.src[e] : 0ull, d = has ? l.dst[e] : 0ull;
        const uint32_t dl = has ? l.delta[e] : 0u;
        wave_copy_pairs(__ballot(has), so, d, dl, src, dst, fw);
    }
    wave_lds_fence();
}

// One lane copies one field of fw <= 32 words (src_off a multiple of 4
// words): every 16-byte load of the field is issued before the first store,
// so a wave has all its fields in flight in one round.  Stores are 16 bytes
// wide where dst_off is a multiple of 4 words and the vector lies inside the
// field, else word by word, so nothing past word fw is written.
__device__ __forceinline__ void lane_copy(bool active, uint64_t src_off, uint64_t dst_off,
                                          uint32_t delta, const uint32_t *src, uint32_t *dst,
                                          uint32_t fw)
{
    if (!active) return;
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src + src_off);
    uint4 v[8];
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u)
        if (4 * u < fw) v[u] = s4[u];
    const bool vec = (dst_off & 3ull) == 0;
#pragma unroll
    for (uint32_t u = 0; u < 8; ++u) {
        if (4 * u >= fw) continue;
        const uint32_t w[4] = {v[u].x + delta, v[u].y + delta, v[u].z + delta, v[u].w + delta};
        if (vec && 4 * u + 4 <= fw) {
            *reinterpret_cast<uint4 *>(dst + dst_off + 4 * u) = make_uint4(w[0], w[1], w[2], w[3]);
        } else {
#pragma unroll
            for (uint32_t r = 0; r < 4; ++r)
                if (4 * u + r < fw) dst[dst_off + 4 * u + r] = w[r];
        }
    }
}

// Field copies of a group's flagged slots.  The op of slot k lives in lane
// grp * G + k; its field is src[src_off + w] -> dst[dst_off + w] (+ delta),
// w < fw, and lane gl of the group handles words gl, gl + G, ...  Loads of
// KB slots are issued before their stores, which requires that no two
// flagged slots of one call write the same word (callers guarantee it).
template <int G, int KB>
__device__ __forceinline__ void group_copy(unsigned long long flags, uint64_t src_off,
                                           uint64_t dst_off, uint32_t delta, const uint32_t *src,
                                           uint32_t *dst, uint32_t fw, uint32_t S, uint32_t grp,
                                           uint32_t gl)
{
    for (uint32_t k0 = 0; k0 < S; k0 += KB) {
        uint64_t so[KB], dd[KB];
        uint32_t dl[KB];
        bool a[KB];
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const uint32_t k = k0 + u;
            const int from = (int)(grp * G + (k < (uint32_t)G ? k : 0u));
            so[u] = __shfl(src_off, from);
            dd[u] = __shfl(dst_off, from);
            dl[u] = __shfl(delta, from);
            a[u] = k < S && ((flags >> from) & 1ull);
        }
        bool any = false;
#pragma unroll
        for (int u = 0; u < KB; ++u) any |= a[u];
        if (!__ballot(any)) continue;
        for (uint32_t w = gl; w < fw; w += G) {
            uint32_t v[KB];
            // unconditional loads (inactive slots read word 0) so that all KB
            // misses are in flight together; only active slots store
#pragma unroll
            for (int u = 0; u < KB; ++u) v[u] = src[a[u] ? so[u] + w : 0ull];
#pragma unroll
            for (int u = 0; u < KB; ++u)
                if (a[u]) dst[dd[u] + w] = v[u] + dl[u];
        }
    }
}

// CHECK + COMMIT + read-out, fast form (rules as kv_check_counting_kernel).
// EXECUTE and the read marking have decided everything: a transaction
// commits iff it has no waw (doomed) and not both raw and war.  A block
// compacts its committed transactions into an LDS list and gathers their
// read fields right away, G lanes per transaction: no write-back has run
// yet, so the table is the snapshot (read_1 / __READ, YCSB/0/Execute.h:
// 27-37).  Commit bits collect in LDS and leave as whole bitmap words.
template <int G, int CH>
__global__ __launch_bounds__(kBlock) void kv_check_fast_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta, uint32_t B, uint32_t S,
    uint64_t rows, uint32_t fields, uint32_t fw, uint32_t fsw, const uint8_t *__restrict__ doomed,
    const uint8_t *__restrict__ raw_flag, const uint8_t *__restrict__ war_flag, uint8_t epoch,
    const uint32_t *__restrict__ table, uint32_t *__restrict__ read_values, uint64_t *__restrict__ bitmap,
    uint32_t max_scan, uint32_t L, uint32_t per_lane, DeviceCounters *__restrict__ ctr, bool rm)
{
    __shared__ uint32_t s_list[CH];
    __shared__ unsigned long long s_bits[CH / 64];
    __shared__ uint32_t s_cnt[kBlock / 64];
    __shared__ StageList s_stage[kBlock / 64];
    const int lane = lane_id();
    const uint32_t chunk0 = blockIdx.x * CH;
    const uint32_t t0 = chunk0 + threadIdx.x;
    const bool commit = threadIdx.x < CH && t0 < B && doomed[t0] != epoch &&
                        !(raw_flag[t0] == epoch && war_flag[t0] == epoch);
    {
        const unsigned long long m = __ballot(commit);
        if (threadIdx.x < CH && lane == 0) s_bits[threadIdx.x / 64] = m;
    }
    const uint32_t n = list_build(commit, t0, s_list, s_cnt);
    constexpr uint32_t PER = 64 / G;  // transactions per wave per round
    const uint32_t grp = (uint32_t)lane / G, gl = (uint32_t)lane % G;
    const bool staged = !per_lane && fw <= 32 && L == 1;
    uint32_t pending = 0;
    for (uint32_t j0 = (blockIdx.y * (kBlock / 64) + threadIdx.x / 64) * PER; j0 < n;
         j0 += gridDim.y * (kBlock / 64) * PER) {
        const uint32_t j = j0 + grp;
        const bool has = j < n;
        const uint32_t t = has ? s_list[j] : 0u;
        const uint64_t i = (uint64_t)gl * B + t;  // slot-major: where its reads go
        const Op o = load_op(keys, meta, op_at(gl, t, B, S, rm), has && gl < S, rows, fields, max_scan);
        if (staged) {
            // every read slot has its own destination
            pending = stage_push(s_stage[threadIdx.x / 64], pending, o.len != 0, (o.row * fields + o.field) * fsw,
                                 i * fw, 0u);
            if (pending >= 64) {
                stage_flush(s_stage[threadIdx.x / 64], pending, table, read_values, fw);
                pending = 0;
            }
        } else if (per_lane) {
            // one lane per read op: all fields of a round in flight at once
            for (uint32_t k = 0; k < o.len; ++k)
                lane_copy(true, ((o.row + k) * fields + o.field) * fsw, (i * L + k) * fw, 0u, table,
                          read_values, fw);
        } else if (fw <= 2 && L > 1) {
            // narrow fields of scans: each lane copies its own rows, four
            // loads in flight at a time
            const uint64_t src0 = (o.row * fields + o.field) * fsw, step = (uint64_t)fields * fsw;
            for (uint32_t k0 = 0; k0 < o.len; k0 += 4) {
                uint32_t v[4][2];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
                    for (uint32_t w = 0; w < 2; ++w)
                        v[u][w] = table[(k0 + u < o.len && w < fw) ? src0 + (k0 + u) * step + w : 0ull];
#pragma unroll
                for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
                    for (uint32_t w = 0; w < 2; ++w)
                        if (k0 + u < o.len && w < fw) read_values[(i * L + k0 + u) * fw + w] = v[u][w];
            }
        } else {
            for (uint32_t k = 0;; ++k) {  // row k of each read / scan
                const unsigned long long rd = __ballot(k < o.len);
                if (!rd) break;
                if (fw <= 32)
                    wave_copy_pairs(rd, ((o.row + k) * fields + o.field) * fsw, (i * L + k) * fw, 0u, table,
                                    read_values, fw);
                else
                    group_copy<G, 8>(rd, ((o.row + k) * fields + o.field) * fsw, (i * L + k) * fw, 0u, table,
                                     read_values, fw, S, grp, gl);
            }
        }
    }
    if (pending) stage_flush(s_stage[threadIdx.x / 64], pending, table, read_values, fw);
    if (blockIdx.y != 0) return;  // the chunk's other blocks only read out
    const uint32_t nwords = (B + 63) / 64;
    if (threadIdx.x < CH / 64 && chunk0 / 64 + threadIdx.x < nwords)
        bitmap[chunk0 / 64 + threadIdx.x] = s_bits[threadIdx.x];
    if (threadIdx.x == 0) {
        const uint32_t live_n = (B - chunk0) < (uint32_t)CH ? (B - chunk0) : (uint32_t)CH;
        add_counts(ctr, blockIdx.x % kCounterStripes, LTPG_TYPE_KV, n, live_n);
    }
}

// WRITEBACK: committed transactions apply their write slots (writeback_w /
// __WRITEBACK, YCSB/0/Execute.h:193-206, Genericfunction.h:145-158).
// Value: word + delta, delta = TID | 1 for RMW and 0 for copy-back, applied
// per write op in slot order.  The committed writer of a row is unique (the
// minimum W registrant), so reading the table here reads the snapshot.
// Several write ops of one transaction on one field are folded: the first
// in slot order writes word + count * delta, the others nothing -- the same
// result as applying them one by one, and it lets every store of a round go
// out without ordering between them.
template <int G, int CH>
__global__ __launch_bounds__(kBlock) void kv_writeback_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields,
    uint32_t fw, uint32_t fsw, uint32_t *__restrict__ table, const uint64_t *__restrict__ bitmap,
    uint32_t values, uint64_t hot, uint32_t per_lane, bool rm)
{
    __shared__ uint32_t s_list[CH];
    __shared__ uint32_t s_cnt[kBlock / 64];
    __shared__ StageList s_stage[kBlock / 64];
    const int lane = lane_id();
    const uint32_t chunk0 = blockIdx.x * CH;
    const uint32_t t0 = chunk0 + threadIdx.x;
    // t0 < B also guards a bitmap word left by an earlier, larger batch
    // when this phase runs without a fresh CHECK
    const uint32_t n = list_build(threadIdx.x < CH && t0 < B && ((bitmap[t0 / 64] >> (t0 % 64)) & 1ull), t0,
                                  s_list, s_cnt);
    constexpr uint32_t PER = 64 / G;
    const uint32_t grp = (uint32_t)lane / G, gl = (uint32_t)lane % G;
    uint32_t pending = 0;
    for (uint32_t j0 = (blockIdx.y * (kBlock / 64) + threadIdx.x / 64) * PER; j0 < n;
         j0 += gridDim.y * (kBlock / 64) * PER) {
        const uint32_t j = j0 + grp;
        const bool has = j < n;
        const uint32_t t = has ? s_list[j] : 0u;
        const uint32_t step = (has && values == LTPG_VALUES_RMW) ? (tids[t] | 1u) : 0u;
        const Op o = load_op(keys, meta, op_at(gl, t, B, S, rm), has && gl < S, rows, fields);
        // writes to the hot row go through kv_hot_sum_kernel
        const bool wr = o.valid && o.kind == LTPG_KV_WRITE && o.row != hot;
        const unsigned long long wm = __ballot(wr);
        if (!wm) continue;
        const uint64_t off = (o.row * fields + o.field) * fsw;
        // fold duplicates: count this field among the group's write slots
        uint32_t count = 0;
        bool first = true;
        for (uint32_t k = 0; k < S && k < (uint32_t)G; ++k) {
            const int from = (int)(grp * G + k);
            const uint64_t ok = __shfl(off, from);
            if (((wm >> from) & 1ull) && ok == off) {
                ++count;
                first &= k >= gl;
            }
        }
        if (per_lane) {
            lane_copy(wr && first, off, off, step * count, table, table, fw);
            continue;
        }
        if (fw <= 32) {
            // a row's committed writer is unique and its duplicate writes
            // are folded: staged fields never overlap
            pending = stage_push(s_stage[threadIdx.x / 64], pending, wr && first, off, off, step * count);
            if (pending >= 64) {
                stage_flush(s_stage[threadIdx.x / 64], pending, table, table, fw);
                pending = 0;
            }
            continue;
        }
        group_copy<G, 8>(__ballot(wr && first), off, off, step * count, table, table, fw, S, grp, gl);
    }
    if (pending) stage_flush(s_stage[threadIdx.x / 64], pending, table, table, fw);
}

// One small epoch in one workgroup (LTPG_TUNE_SMALL_EPOCH, and the served
// form's epochs): EXECUTE, read marking, CHECK with the read-out, and
// WRITEBACK back to back, separated by workgroup barriers instead of kernel
// boundaries.  The rules are those of the fast form (kv_execute_dedup_kernel,
// kv_mark_reads_kernel, kv_check_fast_kernel, kv_writeback_kernel) with the
// whole epoch in one workgroup: the ops, TIDs and conflict flags sit in LDS,
// and the W registrations go to an LDS hash of the written rows (row + 1 ->
// minimum TID << 24 | position) instead of the per-row registers in HBM,
// which this path never touches.  The batch may sit in pinned host memory
// (read once).  With seq_out the commit bitmap is also written to host_bits
// and then seq is stored there with a system-scope release, as
// publish_kernel does after a multi-kernel epoch.  A hot row (at most
// kSmallHotFields fields) never registers or conflicts, and its committed
// deltas are summed per field in LDS and added to the row with the other
// writes; rows < 2^32 - 1.
constexpr uint32_t kSmallBlock = 1024;
constexpr uint32_t kSmallOps = LTPG_SMALL_EPOCH_OPS;
constexpr uint32_t kSmallPer = kSmallOps / kSmallBlock;  // op slots per thread
constexpr uint32_t kSmallSlots = 2 * kSmallOps;          // hash load <= 1/2: probing always ends
constexpr uint32_t kSmallUnroll = 4;                     // field words in flight per thread
constexpr uint32_t kSmallMerge = 8;                      // items per thread of the one-round copy
constexpr uint32_t kSmallHotFields = 1024;               // fields of a table with a hot row

__device__ __forceinline__ uint32_t small_hash(uint64_t row)
{
    return (uint32_t)((row * 0x9E3779B97F4A7C15ull) >> 40) & (kSmallSlots - 1);
}

// Slot of a written row in the hash, or -1 if no op of the epoch writes it.
__device__ __forceinline__ int small_find(const uint32_t *s_key, uint64_t row)
{
    const uint32_t key = (uint32_t)row + 1u;
    for (uint32_t h = small_hash(row);; h = (h + 1) & (kSmallSlots - 1)) {
        const uint32_t k = s_key[h];
        if (k == key) return (int)h;
        if (k == 0u) return -1;
    }
}

__global__ __launch_bounds__(kSmallBlock) void kv_small_epoch_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta, const uint32_t *__restrict__ tids,
    uint32_t B, uint32_t S, bool rm, uint64_t rows, uint32_t fields, uint32_t fw, uint32_t fsw, uint32_t max_scan,
    uint32_t L, uint32_t values, uint64_t hot, uint32_t *__restrict__ table, uint32_t *__restrict__ read_values,
    uint64_t *__restrict__ bitmap, uint64_t *__restrict__ host_bits, uint64_t *seq_out, uint64_t seq,
    DeviceCounters *__restrict__ ctr)
{
    __shared__ uint64_t s_row[kSmallOps];
    __shared__ uint32_t s_meta[kSmallOps];
    __shared__ uint32_t s_tid[kSmallOps];
    __shared__ uint32_t s_key[kSmallSlots];            // row + 1 of a written row, 0 = empty
    __shared__ unsigned long long s_min[kSmallSlots];  // its writers' minimum TID << 24 | position
    __shared__ uint8_t s_waw[kSmallOps], s_raw[kSmallOps], s_war[kSmallOps];
    __shared__ uint32_t s_list[kSmallOps];             // committed transactions
    __shared__ uint32_t s_fold[kSmallOps];             // per committed (txn, slot): folded write count
    __shared__ unsigned long long s_bits[kSmallOps / 64];  // the commit bitmap
    __shared__ uint32_t s_hot[kSmallHotFields];        // committed hot-row deltas per field
    __shared__ uint32_t s_n, s_bad;
    const uint32_t n = B * S;
    // the batch: every load issued before the first use
    uint64_t r[kSmallPer];
    uint32_t m[kSmallPer], tv[kSmallPer];
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) {
        const uint32_t i = k * kSmallBlock + threadIdx.x;
        r[k] = keys[i < n ? i : 0u];
        m[k] = meta[i < n ? i : 0u];
        tv[k] = tids[i < B ? i : 0u];
    }
    for (uint32_t k = threadIdx.x; k < kSmallSlots; k += kSmallBlock) {
        s_key[k] = 0u;
        s_min[k] = ~0ull;
    }
    const bool has_hot = hot < rows;
    if (has_hot)
        for (uint32_t f = threadIdx.x; f < fields; f += kSmallBlock) s_hot[f] = 0u;
    // the epoch's commit counts start at zero (thread 0 adds to them last)
    for (uint32_t c = threadIdx.x; c < kCounterStripes * kCountedTypes; c += kSmallBlock)
        ctr->commits[c / kCountedTypes][c % kCountedTypes] = 0;
    if (threadIdx.x == 0) {
        s_n = 0;
        s_bad = 0;
    }
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) {
        const uint32_t i = k * kSmallBlock + threadIdx.x;
        if (i < n) {
            s_row[i] = r[k];
            s_meta[i] = m[k];
        }
        if (i < B) {
            s_tid[i] = tv[k];
            s_waw[i] = 0;
            s_raw[i] = 0;
            s_war[i] = 0;
        }
    }
    // (the counter stores and thread 0's later atomics meet in one L2; the
    // barriers order them)
    __syncthreads();
    // EXECUTE: W registrations into the hash
    uint32_t bad = 0;
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) {
        const uint32_t i = k * kSmallBlock + threadIdx.x;
        if (i >= n) continue;
        const Op o = decode_op(r[k], m[k], rows, fields, max_scan);
        bad += (o.kind != LTPG_KV_NONE && !o.valid) ? 1u : 0u;
        if (!o.valid || o.kind != LTPG_KV_WRITE || o.row == hot) continue;  // hot-row writes never register
        const uint32_t t = txn_of(i, B, S, rm);
        const uint32_t key = (uint32_t)o.row + 1u;
        uint32_t h = small_hash(o.row);
        while (true) {
            const uint32_t prev = atomicCAS(&s_key[h], 0u, key);
            if (prev == 0u || prev == key) break;
            h = (h + 1) & (kSmallSlots - 1);
        }
        atomicMin(&s_min[h], ((unsigned long long)s_tid[t] << kTidShift) | t);
    }
    if (bad) atomicAdd(&s_bad, bad);
    __syncthreads();
    // waw of every write; raw of every read, war of the row's minimum writer
    // when a higher-priority transaction reads it
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) {
        const uint32_t i = k * kSmallBlock + threadIdx.x;
        if (i >= n) continue;
        const Op o = decode_op(r[k], m[k], rows, fields, max_scan);
        if (!o.valid) continue;
        const uint32_t t = txn_of(i, B, S, rm);
        const unsigned long long me = ((unsigned long long)s_tid[t] << kTidShift) | t;
        if (o.kind == LTPG_KV_WRITE) {
            if (o.row != hot && s_min[small_find(s_key, o.row)] < me) s_waw[t] = 1;
            continue;
        }
        bool raw = false;
        for (uint32_t q = 0; q < o.len; ++q) {  // a scan: every row of its range
            if (o.row + q == hot) continue;  // the hot row never conflicts
            const int h = small_find(s_key, o.row + q);
            if (h < 0) continue;
            const unsigned long long w = s_min[h];
            if (w < me)
                raw = true;
            else if (w > me)
                s_war[w & kPosMask] = 1;
        }
        if (raw) s_raw[t] = 1;
    }
    __syncthreads();
    // CHECK: commit bits, the bitmap and the list of committed transactions
    const uint32_t nw = (B + 63) / 64;
    for (uint32_t t0 = 0; t0 < B; t0 += kSmallBlock) {
        const uint32_t t = t0 + threadIdx.x;
        const bool c = t < B && !s_waw[t] && !(s_raw[t] && s_war[t]);
        const unsigned long long bits = __ballot(c);
        if (c) s_list[atomicAdd(&s_n, 1u)] = t;
        if (lane_id() == 0 && t / 64 < nw) {
            bitmap[t / 64] = bits;
            s_bits[t / 64] = bits;
        }
    }
    __syncthreads();
    const uint32_t nc = s_n;
    // write folding, one wave per committed transaction (lane = slot): the
    // first write slot of a transaction on a field writes word + count *
    // delta, the others nothing.  Hot-row writes add their delta to the
    // field's sum instead (kv_hot_sum_kernel's reduce).
    for (uint32_t j = threadIdx.x / 64; j < nc; j += kSmallBlock / 64) {
        const uint32_t t = s_list[j], sl = lane_id();
        const uint32_t i = sl < S ? (rm ? t * S + sl : sl * B + t) : 0u;
        const Op o = decode_op(s_row[i], s_meta[i], rows, fields, max_scan);
        const bool any_wr = sl < S && o.valid && o.kind == LTPG_KV_WRITE;
        if (any_wr && o.row == hot && values == LTPG_VALUES_RMW) atomicAdd(&s_hot[o.field], s_tid[t] | 1u);
        const bool wr = any_wr && o.row != hot;
        const unsigned long long wm = __ballot(wr);
        const uint64_t key = o.row * fields + o.field;
        uint32_t count = 0;
        bool first = true;
        for (uint32_t k = 0; k < S; ++k) {
            const uint64_t kk = __shfl(key, (int)k);
            if (((wm >> k) & 1ull) && kk == key) {
                ++count;
                first &= k >= sl;
            }
        }
        if (sl < S) s_fold[j * S + sl] = wr && first ? count : 0u;
    }
    __syncthreads();
    // read-out of the committed reads and write-back of the committed
    // writes.  Reads see the snapshot: when all items fit one round, every
    // load is issued before a barrier and every store after it; otherwise
    // the read-out runs in rounds, then a barrier, then the write-back.
    const uint32_t per = L * fw, total_r = nc * S * per;
    const uint32_t total_w = values == LTPG_VALUES_RMW ? nc * S * fw : 0u;
    const uint32_t total_h = values == LTPG_VALUES_RMW && has_hot ? fields * fw : 0u;
    const auto read_item = [&](uint32_t x, uint64_t &src, uint64_t &dst) -> bool {
        const uint32_t j = x / per, q = (x % per) / fw, w = x % fw;
        const uint32_t t = s_list[j / S], sl = j % S;
        const uint32_t i = rm ? t * S + sl : sl * B + t;
        const Op o = decode_op(s_row[i], s_meta[i], rows, fields, max_scan);
        const bool ok = q < o.len;
        src = ok ? ((o.row + q) * fields + o.field) * fsw + w : 0ull;
        dst = ((((uint64_t)sl * B + t) * L) + q) * fw + w;
        return ok;
    };
    const auto write_item = [&](uint32_t x, uint64_t &off, uint32_t &add) -> bool {
        if (x >= total_w) {  // a word of the hot row: its field's sum
            const uint32_t f = (x - total_w) / fw, w = (x - total_w) % fw;
            add = s_hot[f];
            off = add ? (hot * fields + f) * fsw + w : 0ull;
            return add != 0u;
        }
        const uint32_t j = x / fw, w = x % fw;
        const uint32_t count = s_fold[j];
        const uint32_t t = s_list[j / S], sl = j % S;
        const uint32_t i = rm ? t * S + sl : sl * B + t;
        const Op o = decode_op(s_row[i], s_meta[i], rows, fields, max_scan);
        add = count * (s_tid[t] | 1u);
        off = count ? (o.row * fields + o.field) * fsw + w : 0ull;
        return count != 0u;
    };
    if (total_r + total_w + total_h <= kSmallMerge * kSmallBlock) {
        uint64_t src[kSmallMerge], dst[kSmallMerge];
        uint32_t v[kSmallMerge], add[kSmallMerge];
        bool ok[kSmallMerge], rd[kSmallMerge];
#pragma unroll
        for (uint32_t u = 0; u < kSmallMerge; ++u) {
            const uint32_t x = u * kSmallBlock + threadIdx.x;
            rd[u] = x < total_r;
            add[u] = 0u;
            if (rd[u])
                ok[u] = read_item(x, src[u], dst[u]);
            else if (x < total_r + total_w + total_h)
                ok[u] = write_item(x - total_r, src[u], add[u]);
            else
                ok[u] = false;
            if (!ok[u]) src[u] = 0ull;
            if (!rd[u]) dst[u] = src[u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kSmallMerge; ++u) v[u] = table[src[u]];
        __syncthreads();
#pragma unroll
        for (uint32_t u = 0; u < kSmallMerge; ++u) {
            if (!ok[u]) continue;
            if (rd[u])
                read_values[dst[u]] = v[u];
            else
                table[dst[u]] = v[u] + add[u];
        }
    } else {
        for (uint32_t b0 = 0; b0 < total_r; b0 += kSmallUnroll * kSmallBlock) {
            uint64_t src[kSmallUnroll], dst[kSmallUnroll];
            bool ok[kSmallUnroll];
            uint32_t v[kSmallUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u) {
                const uint32_t x = b0 + u * kSmallBlock + threadIdx.x;
                ok[u] = x < total_r && read_item(x, src[u], dst[u]);
                if (!ok[u]) src[u] = 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u) v[u] = table[src[u]];
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u)
                if (ok[u]) read_values[dst[u]] = v[u];
        }
        __syncthreads();
        for (uint32_t b0 = 0; b0 < total_w + total_h; b0 += kSmallUnroll * kSmallBlock) {
            uint64_t off[kSmallUnroll];
            uint32_t add[kSmallUnroll], v[kSmallUnroll];
            bool ok[kSmallUnroll];
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u) {
                const uint32_t x = b0 + u * kSmallBlock + threadIdx.x;
                ok[u] = x < total_w + total_h && write_item(x, off[u], add[u]);
                if (!ok[u]) off[u] = 0ull;
            }
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u) v[u] = table[off[u]];
#pragma unroll
            for (uint32_t u = 0; u < kSmallUnroll; ++u)
                if (ok[u]) table[off[u]] = v[u] + add[u];
        }
    }
    if (threadIdx.x == 0) {
        add_counts(ctr, 0, LTPG_TYPE_KV, nc, B);
        if (s_bad) atomicAdd(&ctr->range_errors, s_bad);
    }
    if (seq_out) {
        // one wave copies the bitmap to host memory, then releases seq at
        // system scope: a host that reads seq with an acquire load reads the
        // bitmap too
        __syncthreads();
        if (threadIdx.x < 64) {
            for (uint32_t w = threadIdx.x; w < nw; w += 64) host_bits[w] = s_bits[w];
            __threadfence_system();
            if (threadIdx.x == 0) __hip_atomic_store(seq_out, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

// Hot row: committed writes to it add their delta (TID | 1 per word) in any
// order.  A block sums its transactions' deltas per field in LDS and adds
// each non-zero sum to one of kCounterStripes accumulator rows; then
// kv_hot_apply_kernel folds the stripes into the row and clears them.
constexpr uint32_t kHotLdsFields = 1024;

__global__ __launch_bounds__(kBlock) void kv_hot_sum_kernel(
    const uint64_t *__restrict__ keys, const uint32_t *__restrict__ meta,
    const uint32_t *__restrict__ tids, uint32_t B, uint32_t S, uint64_t rows, uint32_t fields, uint64_t hot,
    const uint64_t *__restrict__ bitmap, uint32_t *__restrict__ acc, bool rm)
{
    __shared__ uint32_t bins[kHotLdsFields];
    const bool lds = fields <= kHotLdsFields;
    if (lds)
        for (uint32_t f = threadIdx.x; f < fields; f += kBlock) bins[f] = 0;
    __syncthreads();
    uint32_t *stripe = acc + (uint64_t)(blockIdx.x % kCounterStripes) * fields;
    const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
    const bool committed = t < B && ((bitmap[t / 64] >> (t % 64)) & 1ull);
    if (committed) {
        const uint32_t delta = tids[t] | 1u;
        for (uint32_t s = 0; s < S; ++s) {
            const Op o = load_op(keys, meta, op_at(s, t, B, S, rm), true, rows, fields);
            if (!o.valid || o.kind != LTPG_KV_WRITE || o.row != hot) continue;
            if (lds) atomicAdd(&bins[o.field], delta);
            else atomicAdd(&stripe[o.field], delta);
        }
    }
    __syncthreads();
    if (lds)
        for (uint32_t f = threadIdx.x; f < fields; f += kBlock)
            if (bins[f]) atomicAdd(&stripe[f], bins[f]);
}

__global__ __launch_bounds__(kBlock) void kv_hot_apply_kernel(uint32_t *__restrict__ table, uint64_t hot,
                                                             uint32_t fields, uint32_t fw, uint32_t fsw,
                                                             uint32_t *__restrict__ acc)
{
    const uint32_t f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= fields) return;
    uint32_t total = 0;
    for (int k = 0; k < kCounterStripes; ++k) {
        total += acc[(uint64_t)k * fields + f];
        acc[(uint64_t)k * fields + f] = 0;
    }
    if (total == 0) return;
    uint32_t *w = table + (hot * fields + f) * (uint64_t)fsw;
    for (uint32_t k = 0; k < fw; ++k) w[k] += total;
}

// Rows `idx` of the table, every field, packed (row i of the output holds
// fields * fw words); one thread per output word.  The caller has checked
// every index against the table.
__global__ __launch_bounds__(kBlock) void kv_gather_rows_kernel(const uint32_t *__restrict__ table,
                                                                const uint64_t *__restrict__ idx, uint64_t n,
                                                                uint32_t fields, uint32_t fw, uint32_t fsw,
                                                                uint32_t *__restrict__ out)
{
    const uint64_t per = (uint64_t)fields * fw;
    for (uint64_t x = (uint64_t)blockIdx.x * kBlock + threadIdx.x; x < n * per; x += (uint64_t)gridDim.x * kBlock) {
        const uint64_t i = x / per, r = x % per, f = r / fw, w = r % fw;
        out[x] = table[(idx[i] * fields + f) * fsw + w];
    }
}

// ---- served key-value epochs in one persistent workgroup ------------------
//
// The device loop's op decode and write hash in 32-bit row arithmetic (its
// tables have fewer than 2^32 - 1 rows): the same fields, validity and
// scan length as decode_op, with a multiplicative hash on 32 bits.
struct Op32 {
    uint32_t row, kind, field, len;
    bool valid;
};

__device__ __forceinline__ Op32 decode_op32(uint64_t key, uint32_t m, uint32_t rows, uint32_t fields,
                                            uint32_t max_scan)
{
    Op32 o;
    o.kind = m & 3u;
    o.field = (m >> 8) & 0xffffu;
    const uint32_t sl = (m >> 24) + 1u;
    o.row = (uint32_t)key;
    o.valid = o.kind != LTPG_KV_NONE && (key >> 32) == 0ull && o.row < rows && o.field < fields &&
              (o.kind != LTPG_KV_SCAN || sl <= max_scan);
    const uint32_t left = o.row < rows ? rows - o.row : 0u;
    o.len = !o.valid || o.kind == LTPG_KV_WRITE ? 0u : o.kind == LTPG_KV_READ ? 1u : (sl < left ? sl : left);
    return o;
}

// The loop keeps an op as (row, meta) in 32 bits each, validated once when
// the request is admitted: meta bit 2 (never set in an op's meta by its
// layout: kind in bits 0-1, field in 8-23, scan length in 24-31) records that
// the op is valid.
constexpr uint32_t kLoopValid = 4u;

__device__ __forceinline__ uint32_t loop_pack_meta(uint64_t key, uint32_t m, uint32_t rows, uint32_t fields,
                                                   uint32_t max_scan)
{
    const Op32 o = decode_op32(key, m, rows, fields, max_scan);
    return (m & ~kLoopValid) | (o.valid ? kLoopValid : 0u);
}

__device__ __forceinline__ Op32 loop_decode(uint32_t row, uint32_t m, uint32_t rows)
{
    Op32 o;
    o.kind = m & 3u;
    o.field = (m >> 8) & 0xffffu;
    o.row = row;
    o.valid = (m & kLoopValid) != 0u;
    const uint32_t sl = (m >> 24) + 1u, left = rows - row;  // (row < rows when valid)
    o.len = !o.valid || o.kind == LTPG_KV_WRITE ? 0u : o.kind == LTPG_KV_READ ? 1u : (sl < left ? sl : left);
    return o;
}

static_assert(kSmallSlots == 4096, "loop_hash yields 12 bits");
__device__ __forceinline__ uint32_t loop_hash(uint32_t row) { return (row * 0x9E3779B1u) >> 20; }

__device__ __forceinline__ int loop_find(const uint32_t *s_key, uint32_t row)
{
    const uint32_t key = row + 1u;
    for (uint32_t h = loop_hash(row);; h = (h + 1) & (kSmallSlots - 1)) {
        const uint32_t k = s_key[h];
        if (k == key) return (int)h;
        if (k == 0u) return -1;
    }
}

// kv_serve_loop_kernel runs the served form's key-value epochs back to back
// in one 1024-thread workgroup, without a kernel launch or a host round trip
// between them:
//
//   admission  the requests aborted by the last epoch (kept in LDS, in
//              priority order) head the next epoch, followed by arrivals
//              taken from the request ring in arrival order, up to the epoch
//              cap.  Priority inside an epoch is batch position, which is
//              arrival order: every carried request arrived before every new
//              one, and carried requests keep their relative order.  The
//              next epoch is decided and admitted right after an epoch's
//              CHECK, from arrivals prefetched during the epoch, into the
//              other halves of the double-buffered LDS arrays, so one
//              barrier separates two epochs; after an idle or stalled turn
//              (or an exit test) the loop's top admits from the ring;
//   phases     those of kv_small_epoch_kernel: W registrations in an LDS
//              hash of the written rows, read marking, Aria's commit rule,
//              read-out of committed reads (when the host keeps them) and
//              write-back of committed writes, the RMW delta being the low
//              32 bits of the arrival sequence number | 1;
//   results    one record per epoch in the done ring (pinned host memory),
//              written by wave 0 right after CHECK (wave 0 takes no part in
//              the write-back) and published by a system-scope release of
//              ctl->done_tail after it;
//   requeue    the aborted requests are compacted, in position order, to the
//              front of the other half of the double-buffered LDS op and
//              sequence arrays and head the next epoch.  Without kept reads
//              no barrier separates the write-back, the publish and the
//              requeue: wave 0 publishes while waves 1.. issue write-back
//              atomics.
//
// Every wait is bounded: the loop ends when the host sets ctl->stop, after
// idle_ticks without work, after stall_ticks with the done ring full and
// after slice_ticks in all; the kernel then saves the carried requests to
// device memory and clears ctl->running.  Control decisions are taken from
// values thread 0 broadcasts through LDS, so every wave leaves the loop at
// the same iteration.
constexpr uint32_t kLoopMaxT = kSmallOps;  // transactions per epoch at one op slot each
// s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15): wait for the wave's vector memory
// accesses only (gfx9 encoding: vmcnt in bits 3:0 and 15:14)
constexpr int kWaitVmcnt0 = 0x0F70;

__device__ __forceinline__ uint64_t sys_load(const uint64_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ __forceinline__ void sys_store(uint64_t *p, uint64_t v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A request-ring entry, written by the host: a system-scope load, which
// bypasses the GPU caches, so an entry of a ring slot the host has reused is
// never served from a line cached on an earlier lap.
template <typename T>
__device__ __forceinline__ T ring_load(const T *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Workgroup barrier for LDS only: it orders LDS accesses like __syncthreads
// but does not wait for the wave's global loads, which stay in flight across
// it (the served loop keeps PCIe reads in flight over several phases).
__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive prefix count of the flags of items k * kSmallBlock + threadIdx.x
// (k < kSmallPer, in item order): per-wave ballots, then one pass over the
// wave totals in LDS.  Returns the total; s_w holds 16 * kSmallPer + 1 words.
__device__ __forceinline__ uint32_t block_count_before(const bool (&f)[kSmallPer], uint32_t (&before)[kSmallPer],
                                                       uint32_t *s_w)
{
    constexpr uint32_t kWaves = kSmallBlock / 64;
    const uint32_t lane = (uint32_t)lane_id(), wave = threadIdx.x / 64;
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t pre[kSmallPer];
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) {
        const unsigned long long m = __ballot(f[k]);
        pre[k] = (uint32_t)__popcll(m & lt);
        if (lane == 0) s_w[k * kWaves + wave] = (uint32_t)__popcll(m);
    }
    lds_sync();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (uint32_t j = 0; j < kSmallPer * kWaves; ++j) {
            const uint32_t c = s_w[j];
            s_w[j] = run;
            run += c;
        }
        s_w[kSmallPer * kWaves] = run;
    }
    lds_sync();
#pragma unroll
    for (uint32_t k = 0; k < kSmallPer; ++k) before[k] = s_w[k * kWaves + wave] + pre[k];
    const uint32_t total = s_w[kSmallPer * kWaves];
    lds_sync();
    return total;
}

// Unsigned division by a divisor fixed for the kernel's life (ops per
// request, field words): q = high 64 bits of n * M with M = floor(2^64 / d)
// + 1 (exactly 2^64 / d for a power of two), exact for every 32-bit n because
// n * d < 2^64 (Granlund and Montgomery, "Division by invariant integers
// using multiplication", 1994).  Two 32x32->64 multiplies instead of the
// ~25-instruction division sequence.
struct FastDiv {
    uint64_t m;
    uint32_t d;
    __device__ explicit FastDiv(uint32_t d_) : m(d_ > 1 ? ~0ull / d_ + 1ull : 0ull), d(d_) {}
    __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        if (d == 1) return n;
        const uint64_t hi = (uint64_t)n * (uint32_t)(m >> 32);
        const uint64_t lo = ((uint64_t)n * (uint32_t)m) >> 32;
        return (uint32_t)((hi + lo) >> 32);
    }
};

__global__ __launch_bounds__(kSmallBlock) void kv_serve_loop_kernel(ServeLoopArgs a)
{
    // op arrays, double-buffered: the requeue writes the carried requests
    // into the half the next epoch reads while this epoch's half is read
    // (rows and validated metas: loop_pack_meta)
    __shared__ uint32_t s_rowb[2][kSmallOps];
    __shared__ uint32_t s_metab[2][kSmallOps];
    __shared__ uint64_t s_seqb[2][kLoopMaxT];        // arrival sequence number by position, double-buffered
    __shared__ uint32_t s_key[kSmallSlots];          // row + 1 of a written row, 0 = empty
    __shared__ uint32_t s_min[kSmallSlots];          // its writers' minimum position
    // conflict flags by position, double-buffered: the next epoch's are
    // cleared while waves may still read this epoch's
    __shared__ uint8_t s_wawb[2][kLoopMaxT], s_rawb[2][kLoopMaxT], s_warb[2][kLoopMaxT];
    __shared__ uint32_t s_list[kLoopMaxT];           // committed transactions, in position order
    __shared__ uint32_t s_dst[kLoopMaxT];            // an aborted position's place in the next epoch, ~0 if committed
    // the same, one copy per wave, for epochs of at most 64 transactions
    // (each wave counts the commits itself)
    __shared__ uint32_t s_listw[kSmallBlock / 64][64], s_dstw[kSmallBlock / 64][64];
    __shared__ uint32_t s_w[16 * kSmallPer + 1];
    __shared__ uint64_t s_busy;  // wall-clock ticks inside epochs
    // control words, double-buffered: thread 0 fills the half the loop is
    // not reading, and every wave switches halves after the next barrier,
    // so reading them needs no barrier of its own
    __shared__ uint64_t s_ctl[2][4];
    __shared__ uint32_t s_bad;

    // kernel arguments as locals: the lambdas below capture by reference,
    // and a reference to the argument struct would force it into scratch
    const uint32_t S = a.S, fields = a.fields, fw = a.fw, fsw = a.fsw, L = a.L, max_scan = a.max_scan;
    const uint32_t values = a.values;
    const FastDiv dS(S), dfw(fw), dper(L * fw);
    const uint64_t rows = a.rows, hot = a.hot, done_mask = a.done_mask;
    // rows < 2^32 - 1 here (kv_serve_loop_ok): 32-bit rows; a hot row
    // outside the table becomes a value no row takes
    const uint32_t rows32 = (uint32_t)rows, hot32 = hot < rows ? (uint32_t)hot : 0xFFFFFFFFu;
    ServeCtl *ctl = a.ctl;
    uint32_t *table = a.table;
    uint32_t *reads = a.reads;
    uint64_t *done = a.done;
    // where the previous launch stopped (fields only the kernel writes)
    // (the start time too comes from thread 0: every exit test below must
    // give the same answer in every wave)
    __shared__ uint64_t s_start;
    if (threadIdx.x == 0) {
        s_ctl[0][0] = sys_load(&ctl->head);
        s_ctl[0][1] = sys_load(&ctl->done_tail);
        s_ctl[0][2] = sys_load(&ctl->epochs);
        s_ctl[0][3] = (sys_load(&ctl->n_carry) & 0xffffffffull) | (sys_load(&ctl->cap) << 32);
        s_start = wall_clock64();
    }
    __syncthreads();
    uint64_t head = s_ctl[0][0], done_tail = s_ctl[0][1], epoch = s_ctl[0][2];
    uint32_t n_carry = (uint32_t)(s_ctl[0][3] & 0xffffffffull), cap = (uint32_t)(s_ctl[0][3] >> 32);
    uint32_t cp = 0;  // the half of s_ctl being read
    uint32_t op = 0;  // the half of s_rowb / s_metab being read
    const uint64_t t_start = s_start;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n_carry * S; i += kSmallBlock) {
        s_rowb[0][i] = (uint32_t)a.carry_keys[i];  // saved by the last launch, already validated
        s_metab[0][i] = a.carry_meta[i];
    }
    for (uint32_t t = threadIdx.x; t < n_carry; t += kSmallBlock) s_seqb[0][t] = a.carry_seq[t];
    // the write hash starts empty; each epoch empties the slots its writes
    // took once its marking is done
    for (uint32_t h = threadIdx.x; h < kSmallSlots; h += kSmallBlock) {
        s_key[h] = 0u;
        s_min[h] = ~0u;
    }
    uint64_t t_work = t_start, code = 0;
    // profile: thread 0 stamps the end of each part (after its barrier)
    // (sums in LDS: a register array would be spilled)
    const bool prof = a.profile != 0;
    __shared__ uint64_t s_prof[kServeProfParts + 2];  // sums, then the last stamp and the epoch count
    if (prof && threadIdx.x == 0) {
        for (int k = 0; k < kServeProfParts; ++k) s_prof[k] = 0;
        s_prof[kServeProfParts] = t_start;
        s_prof[kServeProfParts + 1] = 0;
    }
    const auto stamp = [&](int part) {
        if (prof && threadIdx.x == 0) {
            const uint64_t t = wall_clock64();
            s_prof[part] += t - s_prof[kServeProfParts];
            s_prof[kServeProfParts] = t;
        }
    };
    // The host's control words (tail, stop, done_head) cross PCIe: during an
    // epoch thread 0 issues their loads after admission and stores the
    // values at the end of the epoch, so the next admission decides on
    // values at most one epoch old without waiting for them; only an idle or
    // stalled loop waits for a fresh read.  Ring entries are read only below
    // the tail so read (the host writes them before it advances the tail),
    // with ring_load, after the tail's load has returned.  poll() fills the
    // other half of s_ctl; the caller switches halves after a barrier.
    const auto poll = [&]() {
        if (threadIdx.x == 0) {
            uint64_t *c = s_ctl[cp ^ 1u];
            c[0] = __hip_atomic_load(&ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            c[1] = __hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            c[2] = __hip_atomic_load(&ctl->done_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            c[3] = wall_clock64();
        }
    };
    // Releases the records written so far: every lane of wave 0 (the lanes
    // that store records) issues a system-scope release fence, then lane 0
    // stores done_tail (with kept reads, every wave fenced its stores before
    // the barrier that precedes this).  A release-only fence: the acquire
    // half of __threadfence_system would also invalidate the L2.
    const auto publish = [&](uint64_t tail_value) {
        if (threadIdx.x < 64) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            if (threadIdx.x == 0)
                __hip_atomic_store(&ctl->done_tail, tail_value, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    };
    poll();
    if (threadIdx.x == 0) {
        s_busy = 0;
        s_bad = 0;
    }
    __syncthreads();
    cp ^= 1u;
    // The epoch about to run, uniform in every wave: admitted (its ops,
    // sequence numbers and zeroed flags already in half op, admitted by the
    // last epoch right after its CHECK), B transactions of which n_new new
    // arrivals, now its start time and tail_seen the ring tail it knows.
    bool admitted = false;
    uint32_t B = 0, n_new = 0;
    uint64_t now = 0, tail_seen = 0;
    while (true) {
        uint32_t *const s_row = s_rowb[op];
        uint32_t *const s_meta = s_metab[op];
        uint64_t *const s_seq = s_seqb[op];
        uint8_t *const s_waw = s_wawb[op], *const s_raw = s_rawb[op], *const s_war = s_warb[op];
        if (!admitted) {
        const uint64_t tail = s_ctl[cp][0], stop = s_ctl[cp][1], done_head = s_ctl[cp][2];
        now = s_ctl[cp][3];
        tail_seen = tail;
        if (stop) {
            code = kServeExitStop;
            break;
        }
        if (now - t_start > a.slice_ticks) {
            code = kServeExitSlice;
            break;
        }
        const uint64_t avail = tail > head ? tail - head : 0ull;
        const uint32_t room = cap > n_carry ? cap - n_carry : 0u;
        n_new = (uint32_t)(avail < room ? avail : (uint64_t)room);
        B = n_carry + n_new;
        if (B == 0) {
            if (now - t_work > a.idle_ticks) {
                code = kServeExitIdle;
                break;
            }
            __builtin_amdgcn_s_sleep(32);
            poll();
            lds_sync();
            cp ^= 1u;
            stamp(9);
            continue;
        }
        if (done_tail + 1 + B - done_head > done_mask + 1) {
            if (now - t_work > a.stall_ticks) {
                code = kServeExitStall;
                break;
            }
            __builtin_amdgcn_s_sleep(32);
            poll();
            lds_sync();
            cp ^= 1u;
            stamp(9);
            continue;
        }
        stamp(0);
        // admission after an idle turn, an exit test or a full done ring:
        // arrivals head .. head + n_new - 1 behind the carried requests, read
        // from the ring (every ring load issued before the first LDS store)
        {
            const uint32_t n = B * S, c0 = n_carry * S;
            uint64_t r[kSmallPer] = {};
            uint32_t m[kSmallPer] = {};
#pragma unroll
            for (uint32_t k = 0; k < kSmallPer; ++k) {
                const uint32_t i = k * kSmallBlock + threadIdx.x;
                const bool live = i >= c0 && i < n;
                const uint32_t jn = live ? dS.div(i - c0) : 0u;
                const uint64_t at = live ? ((head + jn) & a.ring_mask) * S + (i - c0 - jn * S) : 0ull;
                if (n_new) {  // uniform: no PCIe read for an epoch of carried requests
                    r[k] = ring_load(&a.ring_keys[at]);
                    m[k] = ring_load(&a.ring_meta[at]);
                }
            }
#pragma unroll
            for (uint32_t k = 0; k < kSmallPer; ++k) {
                const uint32_t i = k * kSmallBlock + threadIdx.x;
                if (i >= c0 && i < n) {
                    s_row[i] = (uint32_t)r[k];
                    s_meta[i] = loop_pack_meta(r[k], m[k], rows32, fields, max_scan);
                }
            }
        }
        for (uint32_t t = n_carry + threadIdx.x; t < B; t += kSmallBlock) s_seq[t] = head + (t - n_carry);
        for (uint32_t t = threadIdx.x; t < B; t += kSmallBlock) {
            s_waw[t] = 0;
            s_raw[t] = 0;
            s_war[t] = 0;
        }
        head += n_new;
        lds_sync();
        stamp(1);
        }  // !admitted
        const uint32_t n = B * S;
        // the arrivals this epoch took are in LDS (in every wave, before the
        // barrier above): their ring slots may be reused by the host
        if (threadIdx.x == 0 && n_new) __hip_atomic_store(&ctl->head, head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        // the next admission's control words: loads in flight during the epoch
        uint64_t c_tail = 0, c_stop = 0, c_dh = 0;
        if (threadIdx.x == 0) {
            c_tail = __hip_atomic_load(&ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            c_stop = __hip_atomic_load(&ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            c_dh = __hip_atomic_load(&ctl->done_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        // and its arrivals: the published requests after this admission, up
        // to a full epoch, loaded now and admitted into the next epoch right
        // after this one's CHECK
        const uint32_t pf_n = (uint32_t)(tail_seen - head < a.cap_max ? tail_seen - head : (uint64_t)a.cap_max);
        uint64_t pf_r[kSmallPer] = {};
        uint32_t pf_m[kSmallPer] = {};
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            const uint32_t i = k * kSmallBlock + threadIdx.x;
            const uint32_t q = dS.div(i);
            if (i < pf_n * S) {
                const uint64_t at = ((head + q) & a.ring_mask) * S + (i - q * S);
                pf_r[k] = ring_load(&a.ring_keys[at]);
                pf_m[k] = ring_load(&a.ring_meta[at]);
            }
        }
        uint32_t r[kSmallPer];
        uint32_t m[kSmallPer];
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            const uint32_t i = k * kSmallBlock + threadIdx.x;
            r[k] = i < n ? s_row[i] : 0u;
            m[k] = i < n ? s_meta[i] : 0u;
        }
        // EXECUTE: W registrations into the hash (priority = position).  A
        // single-row read takes a slot for its row as well (its minimum
        // stays ~0 unless a write registers there), so that the marking
        // below reads the row's minimum without probing; every slot taken is
        // kept for the marking.
        uint32_t bad = 0;
        uint32_t wslot[kSmallPer];  // ~0: no slot
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            wslot[k] = ~0u;
            const uint32_t i = k * kSmallBlock + threadIdx.x;
            if (i >= n) continue;
            const Op32 o = loop_decode(r[k], m[k], rows32);
            bad += (o.kind != LTPG_KV_NONE && !o.valid) ? 1u : 0u;
            if (!o.valid || o.kind == LTPG_KV_SCAN || o.row == hot32) continue;
            const uint32_t key = o.row + 1u;
            uint32_t h = loop_hash(o.row);
            while (true) {
                const uint32_t prev = atomicCAS(&s_key[h], 0u, key);
                if (prev == 0u || prev == key) break;
                h = (h + 1) & (kSmallSlots - 1);
            }
            wslot[k] = h;
            if (o.kind == LTPG_KV_WRITE) atomicMin(&s_min[h], dS.div(i));
        }
        if (bad) atomicAdd(&s_bad, bad);
        lds_sync();
        stamp(2);
        // waw of every write; raw of every read, war of the row's minimum
        // writer when a higher-priority transaction reads it
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            const uint32_t i = k * kSmallBlock + threadIdx.x;
            if (i >= n) continue;
            const Op32 o = loop_decode(r[k], m[k], rows32);
            if (!o.valid) continue;
            const uint32_t t = dS.div(i);
            if (o.kind == LTPG_KV_WRITE) {
                if (o.row != hot32 && s_min[wslot[k]] < t) s_waw[t] = 1;
                continue;
            }
            // (a minimum of ~0: no write registered the row)
            bool raw = false;
            if (o.kind == LTPG_KV_READ) {
                if (wslot[k] != ~0u) {
                    const uint32_t w = s_min[wslot[k]];
                    if (w < t)
                        raw = true;
                    else if (w > t && w != ~0u)
                        s_war[w] = 1;
                }
            } else {
                for (uint32_t q = 0; q < o.len; ++q) {
                    if (o.row + q == hot32) continue;
                    const int h = loop_find(s_key, o.row + q);
                    if (h < 0) continue;
                    const uint32_t w = s_min[h];
                    if (w < t)
                        raw = true;
                    else if (w > t && w != ~0u)
                        s_war[w] = 1;
                }
            }
            if (raw) s_raw[t] = 1;
        }
        // the control words loaded during the epoch, and the time, to LDS
        // before the barrier: every wave decides on them after CHECK
        if (threadIdx.x == 0) {
            uint64_t *cn = s_ctl[cp ^ 1u];
            cn[0] = c_tail;
            cn[1] = c_stop;
            cn[2] = c_dh;
            cn[3] = wall_clock64();
            s_busy += cn[3] - now;
        }
        lds_sync();
        stamp(3);
        // CHECK: commit bits and the committed list in position order.  An
        // epoch of at most 64 transactions (the common case under the
        // commit-based cap) is counted with one ballot: by every wave into
        // its own copies of the arrays, with no barrier after it, when the
        // host keeps no reads; else by wave 0 alone.
        uint32_t nc;
        const uint32_t *dstv = s_dst, *listv = s_list;  // the arrays this wave reads below
        if (B <= 64 && !reads) {
            const uint32_t t = (uint32_t)lane_id(), wave = threadIdx.x / 64;
            const bool ct = t < B && !s_waw[t] && !(s_raw[t] && s_war[t]);
            const unsigned long long m = __ballot(ct);
            const uint32_t bf = (uint32_t)__popcll(m & (t ? (~0ull >> (64 - t)) : 0ull));
            uint32_t *const dw = s_dstw[wave], *const lw = s_listw[wave];
            if (t < B) dw[t] = ct ? ~0u : t - bf;
            if (ct) lw[bf] = t;
            nc = (uint32_t)__popcll(m);
            dstv = dw;
            listv = lw;
            // the wave's lanes read each other's entries below: LDS accesses
            // of one wave complete in order; keep the compiler's order too
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            stamp(4);
        } else {
        if (B <= 64) {
            if (threadIdx.x < 64) {
                const uint32_t t = threadIdx.x;
                const bool ct = t < B && !s_waw[t] && !(s_raw[t] && s_war[t]);
                const unsigned long long m = __ballot(ct);
                const uint32_t bf = (uint32_t)__popcll(m & (t ? (~0ull >> (64 - t)) : 0ull));
                if (t < B) s_dst[t] = ct ? ~0u : t - bf;
                if (ct) s_list[bf] = t;
                if (t == 0) s_w[0] = (uint32_t)__popcll(m);
            }
        } else {
            bool c[kSmallPer];
            uint32_t before[kSmallPer];
#pragma unroll
            for (uint32_t k = 0; k < kSmallPer; ++k) {
                const uint32_t t = k * kSmallBlock + threadIdx.x;
                c[k] = t < B && !s_waw[t] && !(s_raw[t] && s_war[t]);
            }
            const uint32_t total = block_count_before(c, before, s_w);
#pragma unroll
            for (uint32_t k = 0; k < kSmallPer; ++k) {
                const uint32_t t = k * kSmallBlock + threadIdx.x;
                if (t < B) s_dst[t] = c[k] ? ~0u : t - before[k];
                if (c[k]) s_list[before[k]] = t;
            }
            if (threadIdx.x == 0) s_w[0] = total;
        }
        // With kept reads the read-out below loads the table, so the last
        // epoch's write-back atomics must be ordered before it: a barrier
        // with workgroup-scope release / acquire of global memory.  Without,
        // nothing in the loop reads the table and an LDS barrier suffices.
        if (reads)
            __syncthreads();
        else
            lds_sync();
        nc = s_w[0];
        stamp(4);
        }
        // The next epoch, decided now in every wave alike: the aborted
        // requests head it (the requeue below moves them), followed by the
        // prefetched arrivals up to its cap.  It is admitted now, into the
        // other halves, unless the host asked to stop, the slice is over, it
        // would be empty or the done ring has no room for its record; the
        // loop's top then decides again on fresh values.  The arrivals go to
        // LDS before the write-back: their loads returned during the phases
        // above, and no load result is then left pending behind the
        // write-back atomics (its first use would wait for those as well).
        const uint32_t carry_n = B - nc;
        uint32_t cap_n = a.cap_max;
        if (a.growth) {
            const uint64_t g = (uint64_t)a.growth * nc;
            cap_n = (uint32_t)(g < a.floor ? a.floor : (g > a.cap_max ? a.cap_max : g));
        }
        const uint64_t n_tail = s_ctl[cp ^ 1u][0], n_stop = s_ctl[cp ^ 1u][1], n_dh = s_ctl[cp ^ 1u][2],
                       n_now = s_ctl[cp ^ 1u][3];
        uint32_t new_n;
        {
            const uint64_t avail = n_tail > head ? n_tail - head : 0ull;
            const uint32_t room = cap_n > carry_n ? cap_n - carry_n : 0u;
            new_n = (uint32_t)(avail < room ? avail : (uint64_t)room);
            new_n = new_n < pf_n ? new_n : pf_n;
        }
        const uint32_t B_n = carry_n + new_n;
        const bool go = !n_stop && n_now - t_start <= a.slice_ticks && B_n > 0 &&
                        done_tail + 1 + B + 1 + B_n - n_dh <= done_mask + 1;
        if (go) {
            uint32_t *const nx_row = s_rowb[op ^ 1u] + (uint64_t)carry_n * S;
            uint32_t *const nx_meta = s_metab[op ^ 1u] + (uint64_t)carry_n * S;
#pragma unroll
            for (uint32_t k = 0; k < kSmallPer; ++k) {
                const uint32_t i = k * kSmallBlock + threadIdx.x;
                if (i < new_n * S) {
                    nx_row[i] = (uint32_t)pf_r[k];
                    nx_meta[i] = loop_pack_meta(pf_r[k], pf_m[k], rows32, fields, max_scan);
                }
            }
            for (uint32_t t = threadIdx.x; t < new_n; t += kSmallBlock) s_seqb[op ^ 1u][carry_n + t] = head + t;
            for (uint32_t t = threadIdx.x; t < B_n; t += kSmallBlock) {
                s_wawb[op ^ 1u][t] = 0;
                s_rawb[op ^ 1u][t] = 0;
                s_warb[op ^ 1u][t] = 0;
            }
        }
        // the marking is done: empty the hash slots this epoch's writes took
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k)
            if (wslot[k] != ~0u) {
                s_key[wslot[k]] = 0u;
                s_min[wslot[k]] = ~0u;
            }
        // the epoch's record, stored now by wave 0 (which takes no part in
        // the write-back below) so that its PCIe writes are done by the
        // time done_tail is released
        if (threadIdx.x < 64) {
            if (threadIdx.x == 0)
                done[done_tail & done_mask] = (uint64_t)B | ((uint64_t)nc << 16) | (epoch << 32);
            for (uint32_t t = threadIdx.x; t < B; t += 64)
                done[(done_tail + 1 + t) & done_mask] = s_seq[t] | (dstv[t] == ~0u ? kDoneCommitted : 0ull);
        }
        // read-out (when kept) and write-back; reads see the snapshot, so
        // every read completes before the first write-back store
        const uint32_t per = L * fw;
        const uint32_t total_r = reads ? nc * S * per : 0u;
        const uint32_t total_w = values == LTPG_VALUES_RMW ? nc * S * fw : 0u;
        const uint64_t entry_words = (uint64_t)S * per;
        const auto read_item = [&](uint32_t x, uint64_t &src, uint64_t &dst) -> bool {
            const uint32_t j = dper.div(x), rem = x - j * per, q = dfw.div(rem), w = rem - q * fw;
            const uint32_t jt = dS.div(j), t = listv[jt], sl = j - jt * S;
            const uint32_t i = t * S + sl;
            const Op32 o = loop_decode(s_row[i], s_meta[i], rows32);
            const bool ok = q < o.len;
            src = ok ? ((uint64_t)(o.row + q) * fields + o.field) * fsw + w : 0ull;
            dst = ((done_tail + 1 + t) & done_mask) * entry_words + ((uint64_t)sl * L + q) * fw + w;
            return ok;
        };
        // a committed write adds its delta to every word of its field with
        // an atomic add: several writes of one transaction to one field (and
        // the writes to the hot row, which never conflict) then need no
        // folding, and the sum is the sequential one (modular addition)
        const auto write_item = [&](uint32_t x, uint64_t &off, uint32_t &add) -> bool {
            const uint32_t j = dfw.div(x), w = x - j * fw;
            const uint32_t jt = dS.div(j), t = listv[jt], sl = j - jt * S;
            const uint32_t i = t * S + sl;
            const Op32 o = loop_decode(s_row[i], s_meta[i], rows32);
            const bool wr = o.valid && o.kind == LTPG_KV_WRITE;
            add = (uint32_t)s_seq[t] | 1u;
            off = wr ? ((uint64_t)o.row * fields + o.field) * fsw + w : 0ull;
            return wr;
        };
        // the read-out (kept reads only) in rounds, then a barrier, then the
        // write-back; without kept reads an epoch's writes are one round.
        // (The barrier closes every path through the read-out, so no table
        // load is left pending after it.)
        if (total_r) {
            for (uint32_t b0 = 0; b0 < total_r; b0 += kSmallUnroll * kSmallBlock) {
                uint64_t src[kSmallUnroll], dst[kSmallUnroll];
                bool ok[kSmallUnroll];
                uint32_t v[kSmallUnroll];
#pragma unroll
                for (uint32_t u = 0; u < kSmallUnroll; ++u) {
                    const uint32_t x = b0 + u * kSmallBlock + threadIdx.x;
                    ok[u] = x < total_r && read_item(x, src[u], dst[u]);
                    if (!ok[u]) src[u] = 0ull;
                }
#pragma unroll
                for (uint32_t u = 0; u < kSmallUnroll; ++u)
                    v[u] = __hip_atomic_load(&table[src[u]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                for (uint32_t u = 0; u < kSmallUnroll; ++u)
                    if (ok[u]) reads[dst[u]] = v[u];
            }
            // every table load has returned (a load whose value is not kept
            // is otherwise never waited for, and the first later write of
            // its register would wait for every access issued after it,
            // the write-back atomics included), then every read of the
            // snapshot is done before the first write
            __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
            __syncthreads();
        }
        {
            // waves 1.. issue the write-back atomics and never wait for
            // them inside the loop (no later load result is pending behind
            // them); the next epoch's read-out is ordered after them by its
            // barrier (kept reads), and the launch ends after them
            if (threadIdx.x >= 64)
                for (uint32_t x = threadIdx.x - 64; x < total_w; x += kSmallBlock - 64) {
                    uint64_t off;
                    uint32_t add;
                    if (write_item(x, off, add)) atomicAdd(&table[off], add);
                }
        }
        // committed reads went to host memory from every wave
        if (reads && total_r) __threadfence_system();
        if (threadIdx.x == 64) {
            add_counts(a.ctr, (uint32_t)(epoch % kCounterStripes), LTPG_TYPE_KV, nc, B);
            const uint32_t bad_ops = s_bad;
            s_bad = 0;  // (the next EXECUTE adds to it after the barrier below)
            if (bad_ops) atomicAdd(&a.ctr->range_errors, bad_ops);
        }
        // Kept reads: every wave fenced its stores above, and the barrier
        // orders those fences before the release below.  Without, the record
        // is wave 0's alone and wave 0 publishes it while waves 1.. are still
        // issuing the write-back (nothing after this point reads what they
        // write: the requeue reads this epoch's LDS halves and writes the
        // other ones).
        if (reads) lds_sync();
        stamp(6);
        publish(done_tail + 1 + B);
        done_tail += 1 + B;
        stamp(7);
        // requeue: the aborted requests move to the front, in position order
        // (everything is re-read from LDS: nothing of the phases above stays
        // in registers across the read-out and write-back)
        uint64_t sq[kSmallPer];
        uint32_t ro[kSmallPer], mo[kSmallPer], dst_t[kSmallPer], dst_i[kSmallPer];
        bool ab[kSmallPer], abop[kSmallPer];
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            const uint32_t t = k * kSmallBlock + threadIdx.x;
            dst_t[k] = t < B ? dstv[t] : ~0u;
            ab[k] = dst_t[k] != ~0u;
            sq[k] = ab[k] ? s_seq[t] : 0ull;
            const uint32_t i = t;
            const uint32_t ti = i < n ? dS.div(i) : 0u;
            const uint32_t di = i < n ? dstv[ti] : ~0u;
            abop[k] = di != ~0u;
            dst_i[k] = di * S + (i - ti * S);
            ro[k] = abop[k] ? s_row[i] : 0u;
            mo[k] = abop[k] ? s_meta[i] : 0u;
        }
        // everything goes to the other halves, which no wave reads in this
        // epoch: no barrier between the reads above and the stores below
#pragma unroll
        for (uint32_t k = 0; k < kSmallPer; ++k) {
            if (ab[k]) s_seqb[op ^ 1u][dst_t[k]] = sq[k];
            if (abop[k]) {
                s_rowb[op ^ 1u][dst_i[k]] = ro[k];
                s_metab[op ^ 1u][dst_i[k]] = mo[k];
            }
        }
        n_carry = carry_n;
        cap = cap_n;
        ++epoch;
        t_work = now;
        admitted = go;
        if (go) {
            B = B_n;
            n_new = new_n;
            head += new_n;
            now = n_now;
            tail_seen = n_tail;
        }
        lds_sync();
        cp ^= 1u;
        op ^= 1u;
        stamp(8);
        if (prof && threadIdx.x == 0) s_prof[kServeProfParts + 1] += 1;
    }
    // save the carried requests for the next launch (same stream: ordered)
    for (uint32_t i = threadIdx.x; i < n_carry * S; i += kSmallBlock) {
        a.carry_keys[i] = s_rowb[op][i];  // (the packed form; only this kernel reads it back)
        a.carry_meta[i] = s_metab[op][i];
    }
    for (uint32_t t = threadIdx.x; t < n_carry; t += kSmallBlock) a.carry_seq[t] = s_seqb[op][t];
    __syncthreads();
    if (threadIdx.x == 0) {
        if (prof) {
            for (int k = 0; k < kServeProfParts; ++k) ctl->prof[k] += s_prof[k];
            ctl->prof_epochs += s_prof[kServeProfParts + 1];
        }
        ctl->busy_ticks += s_busy;
        sys_store(&ctl->n_carry, n_carry);
        sys_store(&ctl->cap, cap);
        sys_store(&ctl->epochs, epoch);
        sys_store(&ctl->head, head);
        sys_store(&ctl->exit_code, code);
        sys_store(&ctl->running, 0ull);
    }
}

// Latencies of the steps an epoch of the device loop chains (the terms of
// its dependent-chain roofline), in wall-clock ticks per step, measured by
// thread 0 of one 1024-thread workgroup while every lane runs the same chain:
//   [0] an LDS-only workgroup barrier (lds_sync) over 16 waves;
//   [1] a dependent LDS load (each lane chases its own index chain);
//   [2] a dependent LDS atomic with return (a compare-and-swap that never
//       swaps, along the chain);
//   [3] a dependent system-scope load of pinned host memory (a PCIe round
//       trip), thread 0 alone;
//   [4] a publish -- a system-scope release fence, then a relaxed
//       system-scope store to pinned host memory -- back to back, thread 0
//       alone (each fence also waits for the previous store).
// Every index stays inside its array: LDS chain values are masked to 4096,
// host chain values are masked to chain_words (a power of two).
__global__ __launch_bounds__(kSmallBlock) void latency_probe_kernel(uint64_t *host, uint32_t chain_words,
                                                                   unsigned long long *out, uint32_t reps,
                                                                   uint32_t host_reps, uint32_t mask,
                                                                   uint64_t sentinel)
{
    // mask (4095) and sentinel (never reached) come from the host, so the
    // compiler can neither drop the chains nor the waits for their ends
    __shared__ uint32_t s_chain[4096];
    for (uint32_t i = threadIdx.x; i < 4096; i += kSmallBlock) s_chain[i] = (i * 2053u + 1u) & 4095u;
    __syncthreads();
    unsigned long long t0 = 0;
    // [0] barriers
    if (threadIdx.x == 0) t0 = wall_clock64();
    for (uint32_t r = 0; r < reps; ++r) lds_sync();
    if (threadIdx.x == 0) out[0] = wall_clock64() - t0;
    __syncthreads();
    // [1] dependent LDS loads
    uint32_t idx = (threadIdx.x * 4u) & mask;
    if (threadIdx.x == 0) t0 = wall_clock64();
    for (uint32_t r = 0; r < reps; ++r) idx = s_chain[idx] & mask;
    if (idx == (uint32_t)sentinel) out[5] = idx;  // waits for the last load
    if (threadIdx.x == 0) out[1] = wall_clock64() - t0;
    __syncthreads();
    // [2] dependent LDS atomics with return
    idx = (threadIdx.x * 4u) & mask;
    if (threadIdx.x == 0) t0 = wall_clock64();
    // a compare-and-swap whose compare value is never stored: it returns the
    // slot unchanged (an add of 0 would be turned into a plain load)
    for (uint32_t r = 0; r < reps; ++r) idx = atomicCAS(&s_chain[idx], (uint32_t)sentinel, 0u) & mask;
    if (idx == (uint32_t)sentinel) out[5] = idx;
    if (threadIdx.x == 0) out[2] = wall_clock64() - t0;
    __syncthreads();
    if (threadIdx.x == 0) {
        // [3] PCIe round trips
        uint64_t j = 0;
        t0 = wall_clock64();
        for (uint32_t r = 0; r < host_reps; ++r) j = ring_load(&host[j]) & (chain_words - 1);
        if (j == sentinel) out[5] = j;
        out[3] = wall_clock64() - t0;
        // [4] publishes
        t0 = wall_clock64();
        for (uint32_t r = 0; r < host_reps; ++r) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            __hip_atomic_store(&host[chain_words], (uint64_t)r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        out[4] = wall_clock64() - t0;
    }
}

int lanes_for(uint32_t fw)
{
    if (fw > 64) return 0;
    int lp = 1;
    while ((uint32_t)lp < fw) lp <<= 1;
    return lp;
}

inline dim3 grid_for(uint64_t n_threads)
{
    return dim3((unsigned)((n_threads + kBlock - 1) / kBlock));
}

inline dim3 grid_stride(uint64_t n)
{
    uint64_t g = (n + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g == 0) g = 1;
    return dim3((unsigned)g);
}

#define LTPG_LP_SWITCH(lp, CALL) \
    switch (lp) {                \
    case 1: CALL(1); break;      \
    case 2: CALL(2); break;      \
    case 4: CALL(4); break;      \
    case 8: CALL(8); break;      \
    case 16: CALL(16); break;    \
    case 32: CALL(32); break;    \
    case 64: CALL(64); break;    \
    default: CALL(0); break;     \
    }

}  // namespace

hipError_t launch_reset_epoch_counters(DeviceCounters *ctr, hipStream_t s)
{
    hipLaunchKernelGGL(reset_epoch_counters_kernel, dim3(1), dim3(kBlock), 0, s, ctr);
    return hipGetLastError();
}

hipError_t launch_copy_u128(void *dst, const void *src, uint64_t n16, hipStream_t s)
{
    uint64_t g = (n16 + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g == 0) g = 1;
    hipLaunchKernelGGL(copy_u128_kernel, dim3((unsigned)g), dim3(kBlock), 0, s, static_cast<uint4 *>(dst),
                       static_cast<const uint4 *>(src), n16);
    return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long *p, uint64_t n, unsigned long long v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u64_kernel, grid_stride(n), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_copy_u32(uint32_t *dst, const uint32_t *src, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_u32_kernel, grid_stride(n), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_publish(const PublishSet &p, uint64_t *seq_out, uint64_t seq, hipStream_t s)
{
    hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(kBlock), 0, s, p, seq_out, seq);
    return hipGetLastError();
}

hipError_t launch_copy_u64(uint64_t *dst, const uint64_t *src, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(copy_u64_kernel, grid_stride(n), dim3(kBlock), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_calibration(unsigned long long *a, uint64_t len, uint64_t n, int atomics,
                              unsigned long long *sink, hipStream_t s)
{
    constexpr uint32_t per = 16;
    const uint64_t threads = (n + per - 1) / per;
    const dim3 grid((unsigned)((threads + kBlock - 1) / kBlock));
    if (atomics)
        hipLaunchKernelGGL(calib_atomic_kernel, grid, dim3(kBlock), 0, s, a, len, n, per, sink);
    else
        hipLaunchKernelGGL(calib_load_kernel, grid, dim3(kBlock), 0, s, a, len, n, per, sink);
    return hipGetLastError();
}

hipError_t launch_gather_calibration(const uint32_t *src, uint64_t slots, uint32_t stride_w, uint32_t fw,
                                     uint32_t *dst, uint64_t n, hipStream_t s)
{
    constexpr uint32_t per = 4;
    const uint64_t waves = (n + 64ull * per - 1) / (64ull * per);
    const dim3 grid((unsigned)((waves * 64 + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(calib_gather_kernel, grid, dim3(kBlock), 0, s, src, slots, stride_w, fw, dst, n, per);
    return hipGetLastError();
}

hipError_t launch_grid_barrier(uint32_t blocks, unsigned int *count, unsigned int *gen, uint32_t reps,
                               uint32_t spin_limit, unsigned int *err, hipStream_t s)
{
    void *args[] = {&count, &gen, &reps, &spin_limit, &err};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(grid_barrier_kernel), dim3(blocks), dim3(kBlock),
                                      args, 0, s);
}

hipError_t launch_boundary(uint32_t blocks, unsigned int *p, hipStream_t s)
{
    hipLaunchKernelGGL(boundary_kernel, dim3(blocks), dim3(kBlock), 0, s, p);
    return hipGetLastError();
}

hipError_t launch_stamp(unsigned long long *out, hipStream_t s)
{
    hipLaunchKernelGGL(stamp_kernel, dim3(1), dim3(1), 0, s, out);
    return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t *p, uint64_t n, uint8_t v, hipStream_t s)
{
    hipLaunchKernelGGL(fill_u8_kernel, grid_stride(n), dim3(kBlock), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_kv_fill_reference(const KvState &kv, hipStream_t s)
{
    const uint64_t n = kv.rows * kv.fields * kv.field_stride;
    hipLaunchKernelGGL(kv_fill_reference_kernel, grid_stride(n), dim3(kBlock), 0, s, kv.table, n,
                       kv.fields, kv.field_words, kv.field_stride);
    return hipGetLastError();
}

hipError_t launch_repack(const uint32_t *src, uint32_t *dst, uint64_t n_fields, uint32_t fw,
                         uint32_t fsw, int to_strided, hipStream_t s)
{
    hipLaunchKernelGGL(repack_kernel, grid_stride(n_fields * fw), dim3(kBlock), 0, s, src, dst, n_fields, fw,
                       fsw, to_strided);
    return hipGetLastError();
}

hipError_t launch_kv_transpose(const uint32_t *src, uint32_t *dst, uint64_t rows, uint32_t cols,
                               int to_row_major, hipStream_t s)
{
    hipLaunchKernelGGL(transpose_kernel<uint32_t>, grid_stride(rows * cols), dim3(kBlock), 0, s, src, dst, rows,
                       cols, to_row_major);
    return hipGetLastError();
}

RegLayout reg_layout(const KvState &kv)
{
    RegLayout rl;
    rl.row_mul = kv.split_regs ? 1 : 2;
    rl.w_off = kv.split_regs ? kv.rows : 1;
    return rl;
}

hipError_t launch_kv_execute(const KvState &kv, const KvBatchDev &b, const EpochTag &e, bool counts,
                             DeviceCounters *ctr, hipStream_t s, bool zero_counts)
{
    // an empty epoch still starts its commit counts at zero
    if (b.n_txns == 0) return zero_counts ? launch_reset_epoch_counters(ctr, s) : hipSuccess;
    const uint64_t ops = (uint64_t)b.n_txns * b.n_slots;
    const dim3 op_grid((unsigned)((ops + kBlock - 1) / kBlock));
    if (kv.partitioned && !counts && kv.part.ent_me) {
        PartArgs a;
        a.keys = b.keys;
        a.meta = b.meta;
        a.tids = b.tids;
        a.B = b.n_txns;
        a.S = b.n_slots;
        a.rows = kv.rows;
        a.fields = kv.fields;
        a.max_scan = kv.max_scan;
        a.hot = kv.hot_row;
        a.nb = kv.part.n_buckets;
        a.cap_slices = kv.part.cap_slices;
        a.count = kv.part.count;
        a.start = kv.part.start;
        a.cursor = kv.part.cursor;
        a.slices = kv.part.slices;
        a.ent_me = kv.part.ent_me;
        a.ent_key = kv.part.ent_key;
        a.regs = kv.regs;
        a.rl = reg_layout(kv);
        a.tag = e.tag;
        a.epoch = e.epoch;
        a.rm = b.request_major;
        a.doomed = kv.doomed;
        a.raw_flag = kv.raw_flag;
        a.war_flag = kv.war_flag;
        a.ctr = ctr;
        const dim3 chunks((unsigned)((ops + kPartChunk - 1) / kPartChunk));
        // slices: one per non-empty bucket, plus one per kPartSlice operations
        const uint64_t entries = ops * (kv.max_scan > 1 ? kv.max_scan : 1);
        uint64_t slices = a.nb + entries / kPartSlice + 1;
        if (slices > kv.part.cap_slices) slices = kv.part.cap_slices;
        hipLaunchKernelGGL(kv_part_count_kernel, chunks, dim3(kPartBlock), 0, s, a, zero_counts);
        hipLaunchKernelGGL(kv_part_scan_kernel, dim3(1), dim3(kPartBlock), 0, s, a);
        if (kv.max_scan > 1)
            hipLaunchKernelGGL(kv_part_scatter_kernel, chunks, dim3(kPartBlock), 0, s, a);
        else if (a.nb <= kPartFewBuckets)
            hipLaunchKernelGGL((kv_part_scatter_staged_kernel<kPartFewBuckets, 8>),
                               dim3((unsigned)((ops + 8 * kPartBlock - 1) / (8 * kPartBlock))), dim3(kPartBlock), 0, s,
                               a);
        else
            hipLaunchKernelGGL((kv_part_scatter_staged_kernel<kPartMaxBuckets, 4>),
                               dim3((unsigned)((ops + 4 * kPartBlock - 1) / (4 * kPartBlock))), dim3(kPartBlock), 0, s,
                               a);
        hipLaunchKernelGGL(kv_part_process_kernel, dim3((unsigned)slices), dim3(kPartBlock), 0, s, a);
        hipLaunchKernelGGL(kv_part_flags_kernel, dim3((unsigned)slices), dim3(kPartBlock), 0, s, a);
        return hipGetLastError();
    }
    if (kv.exec_per_op && !counts) {
        if (kv.more_items)
            hipLaunchKernelGGL(kv_execute_dedup_kernel<4>, dim3((unsigned)((ops + 4 * kBlock - 1) / (4 * kBlock))),
                               dim3(kBlock), 0, s, b.keys, b.meta, b.tids, b.n_txns, b.n_slots, kv.rows, kv.fields,
                               kv.regs, reg_layout(kv), kv.hot_row, e.tag, kv.doomed, e.epoch, zero_counts,
                               kv.max_scan, ctr, b.request_major);
        else
            hipLaunchKernelGGL(kv_execute_dedup_kernel<2>, dim3((unsigned)((ops + 2 * kBlock - 1) / (2 * kBlock))),
                               dim3(kBlock), 0, s, b.keys, b.meta, b.tids, b.n_txns, b.n_slots, kv.rows, kv.fields,
                               kv.regs, reg_layout(kv), kv.hot_row, e.tag, kv.doomed, e.epoch, zero_counts,
                               kv.max_scan, ctr, b.request_major);
    } else if (kv.exec_per_op) {
        hipLaunchKernelGGL(kv_execute_ops_kernel, op_grid, dim3(kBlock), 0, s, b.keys, b.meta, b.tids, b.n_txns,
                           b.n_slots, kv.rows, kv.fields, kv.regs, reg_layout(kv), kv.hot_row, e.tag, kv.doomed,
                           e.epoch, counts, zero_counts, kv.max_scan, ctr, b.request_major);
    } else {
        hipLaunchKernelGGL(kv_execute_kernel, grid_for(b.n_txns), dim3(kBlock), 0, s, b.keys, b.meta, b.tids,
                           b.n_txns, b.n_slots, kv.rows, kv.fields, kv.regs, reg_layout(kv), kv.hot_row, e.tag,
                           kv.doomed, e.epoch, counts, zero_counts, kv.max_scan, ctr, b.request_major);
    }
    if (!counts) {
        if (kv.more_items)
            hipLaunchKernelGGL(kv_mark_reads_kernel<2>, dim3((unsigned)((ops + 2 * kBlock - 1) / (2 * kBlock))),
                               dim3(kBlock), 0, s, b.keys, b.meta, b.tids, b.n_txns, b.n_slots, kv.rows, kv.fields,
                               kv.regs, reg_layout(kv), kv.hot_row, e.tag, kv.raw_flag, kv.war_flag, e.epoch,
                               kv.max_scan, b.request_major);
        else
            hipLaunchKernelGGL(kv_mark_reads_kernel<1>, op_grid, dim3(kBlock), 0, s, b.keys, b.meta, b.tids,
                               b.n_txns, b.n_slots, kv.rows, kv.fields, kv.regs, reg_layout(kv), kv.hot_row, e.tag,
                               kv.raw_flag, kv.war_flag, e.epoch, kv.max_scan, b.request_major);
    }
    return hipGetLastError();
}

int group_for(uint32_t S)
{
    int g = 1;
    while ((uint32_t)g < S) g <<= 1;
    return g;
}

#define LTPG_G_SWITCH(g, CALL)   \
    switch (g) {                \
    case 1: CALL(1); break;     \
    case 2: CALL(2); break;     \
    case 4: CALL(4); break;     \
    case 8: CALL(8); break;     \
    case 16: CALL(16); break;   \
    case 32: CALL(32); break;   \
    default: CALL(64); break;   \
    }

// A chunk's committed transactions are shared out over gridDim.y blocks,
// each of which rebuilds the chunk's list (CH flag reads), until the grid
// has about 2048 blocks or every wave of the chunk copies a single round of
// transactions: a small epoch's read-out and write-back then run as many
// short waves instead of a few long ones.
inline dim3 grid_chunks(uint32_t B, uint32_t ch, uint32_t g)
{
    const uint32_t chunks = (B + ch - 1) / ch;
    const uint32_t per_block = (kBlock / 64) * (64 / g);  // transactions per block and round
    const uint32_t max_split = (ch + per_block - 1) / per_block;
    uint32_t split = (2048 + chunks - 1) / chunks;
    if (split > max_split) split = max_split;
    return dim3(chunks, split < 1 ? 1 : split);
}

// The per-lane copy needs fields that start on 16-byte boundaries and fit
// eight 16-byte loads.
static uint32_t lane_copy_ok(const KvState &kv)
{
    return kv.lane_copy && kv.field_stride % 4 == 0 && kv.field_words <= 32 ? 1u : 0u;
}

hipError_t launch_kv_check(const KvState &kv, const KvBatchDev &b, const EpochTag &e, bool counts,
                           DeviceCounters *ctr, hipStream_t s)
{
    if (b.n_txns == 0) return hipSuccess;
    if (counts) {
        hipLaunchKernelGGL(kv_check_counting_kernel, grid_for(b.n_txns), dim3(kBlock), 0, s, b.keys,
                           b.meta, b.tids, b.n_txns, b.n_slots, kv.rows, kv.fields, kv.regs, reg_layout(kv),
                           kv.hot_row, e.tag, kv.bitmap, kv.conflicts, kv.max_scan, ctr, b.request_major);
        return hipGetLastError();
    }
#define CALL1(GV, CHV)                                                                               \
    hipLaunchKernelGGL((kv_check_fast_kernel<GV, CHV>), grid_chunks(b.n_txns, CHV, GV), dim3(kBlock), 0, s, b.keys, \
                       b.meta, b.n_txns, b.n_slots, kv.rows, kv.fields, kv.field_words, kv.field_stride,   \
                       kv.doomed, kv.raw_flag, kv.war_flag, e.epoch, kv.table, kv.read_values, kv.bitmap,  \
                       kv.max_scan, kv.read_slot, lane_copy_ok(kv), ctr, b.request_major)
#define CALL(GV)                    \
    if (kv.small_chunks)            \
        CALL1(GV, 64);              \
    else                            \
        CALL1(GV, kChunk)
    LTPG_G_SWITCH(group_for(b.n_slots), CALL)
#undef CALL
#undef CALL1
    return hipGetLastError();
}

hipError_t launch_kv_readout(const KvState &kv, const KvBatchDev &b, hipStream_t s)
{
    if (b.n_txns == 0) return hipSuccess;
    const int lp = lanes_for(kv.field_words);
#define CALL(LPV)                                                                          \
    hipLaunchKernelGGL(kv_readout_kernel<LPV>, grid_for(b.n_txns), dim3(kBlock), 0, s,    \
                       b.keys, b.meta, b.n_txns, b.n_slots, kv.rows, kv.fields,           \
                       kv.field_words, kv.field_stride, kv.table, kv.bitmap, kv.read_values, kv.max_scan, \
                       kv.read_slot, b.request_major)
    LTPG_LP_SWITCH(lp, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_kv_writeback(const KvState &kv, const KvBatchDev &b, uint32_t values,
                               hipStream_t s)
{
    if (b.n_txns == 0) return hipSuccess;
#define CALL1(GV, CHV)                                                                                      \
    hipLaunchKernelGGL((kv_writeback_kernel<GV, CHV>), grid_chunks(b.n_txns, CHV, GV), dim3(kBlock), 0, s, b.keys, \
                       b.meta, b.tids, b.n_txns, b.n_slots, kv.rows, kv.fields, kv.field_words,             \
                       kv.field_stride, kv.table, kv.bitmap, values, kv.hot_row, lane_copy_ok(kv), b.request_major)
#define CALL(GV)         \
    if (kv.small_chunks) \
        CALL1(GV, 64);   \
    else                 \
        CALL1(GV, kChunk)
    LTPG_G_SWITCH(group_for(b.n_slots), CALL)
#undef CALL
#undef CALL1
    if (kv.hot_row != LTPG_NO_HOT_ROW && values == LTPG_VALUES_RMW) {
        hipLaunchKernelGGL(kv_hot_sum_kernel, grid_for(b.n_txns), dim3(kBlock), 0, s, b.keys, b.meta, b.tids,
                           b.n_txns, b.n_slots, kv.rows, kv.fields, kv.hot_row, kv.bitmap, kv.hot_acc,
                           b.request_major);
        hipLaunchKernelGGL(kv_hot_apply_kernel, grid_for(kv.fields), dim3(kBlock), 0, s, kv.table, kv.hot_row,
                           kv.fields, kv.field_words, kv.field_stride, kv.hot_acc);
    }
    return hipGetLastError();
}

bool kv_small_epoch_ok(const KvState &kv, const KvBatchDev &b)
{
    return b.n_txns > 0 && b.n_slots > 0 && (uint64_t)b.n_txns * b.n_slots <= kSmallOps &&
           (kv.hot_row == LTPG_NO_HOT_ROW || kv.fields <= kSmallHotFields) && kv.rows < 0xFFFFFFFFull;
}

hipError_t launch_kv_small_epoch(const KvState &kv, const KvBatchDev &b, uint32_t values, DeviceCounters *ctr,
                                 uint64_t *host_bits, uint64_t *seq_out, uint64_t seq, hipStream_t s)
{
    if (!kv_small_epoch_ok(kv, b)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(kv_small_epoch_kernel, dim3(1), dim3(kSmallBlock), 0, s, b.keys, b.meta, b.tids, b.n_txns,
                       b.n_slots, b.request_major, kv.rows, kv.fields, kv.field_words, kv.field_stride, kv.max_scan,
                       kv.read_slot, values, kv.hot_row, kv.table, kv.read_values, kv.bitmap, host_bits, seq_out, seq,
                       ctr);
    return hipGetLastError();
}

hipError_t launch_kv_gather_rows(const KvState &kv, const uint64_t *idx, uint64_t n, uint32_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(kv_gather_rows_kernel, grid_stride(n * kv.fields * kv.field_words), dim3(kBlock), 0, s,
                       kv.table, idx, n, kv.fields, kv.field_words, kv.field_stride, out);
    return hipGetLastError();
}

bool kv_serve_loop_ok(const KvState &kv, uint32_t S)
{
    return S >= 1 && S <= 64 && S <= kSmallOps && (kv.hot_row == LTPG_NO_HOT_ROW || kv.fields <= kSmallHotFields) &&
           kv.rows < 0xFFFFFFFFull;
}

hipError_t launch_latency_probe(uint64_t *host, uint32_t chain_words, unsigned long long *out, uint32_t reps,
                                uint32_t host_reps, hipStream_t s)
{
    if (chain_words == 0 || (chain_words & (chain_words - 1))) return hipErrorInvalidValue;
    hipLaunchKernelGGL(latency_probe_kernel, dim3(1), dim3(kSmallBlock), 0, s, host, chain_words, out, reps, host_reps,
                       4095u, ~0ull);
    return hipGetLastError();
}

hipError_t launch_kv_serve_loop(const ServeLoopArgs &a, hipStream_t s)
{
    if (a.S < 1 || a.S > 64 || (uint64_t)a.cap_max * a.S > kSmallOps || a.cap_max == 0 || a.floor == 0)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(kv_serve_loop_kernel, dim3(1), dim3(kSmallBlock), 0, s, a);
    return hipGetLastError();
}

}  // namespace ltpg
