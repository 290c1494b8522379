// device_common.h - registration primitives shared by the key-value and
// order-entry kernels (the reference's __REGISTER / __CHECK semantics,
// LTPG/TPCC-allopt-final/Genericfunction.h:108-414).
//
// Register value: tag:8 | TID:32 | batch position:24, tag = 255 - epoch.
// Lower value = higher priority; every value of an older epoch is larger
// than every value of the current one, so "register < my priority" implies
// "registered in this epoch by a higher-priority transaction".
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ltpg {

constexpr int kTagShift = 56;
constexpr int kTidShift = 24;
constexpr unsigned long long kPosMask = 0xFFFFFFull;

__device__ __forceinline__ unsigned long long priority(uint32_t tag, uint32_t tid, uint32_t pos)
{
    return ((unsigned long long)tag << kTagShift) | ((unsigned long long)tid << kTidShift) | pos;
}

// Probe of a register before the atomic: a system-scope relaxed load, which
// reads the coherent copy the atomics update.  Any value would be safe (a
// register only ever decreases, so every value it ever held is >= its
// current one), but on hot rows a stale cached copy sends far more lanes
// into the contended atomic; measured on MI355X at Zipf 0.99 the coherent
// probe is clearly the fastest.
__device__ __forceinline__ unsigned long long probe(const unsigned long long *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Register a read (atomicMin on R); skipped when it cannot lower R.
__device__ __forceinline__ void register_read(unsigned long long *r, unsigned long long me)
{
    if (probe(r) > me) atomicMin(r, me);
}

// Register a write (atomicMin on W) and detect waw early.  Returns true when
// a higher-priority writer is registered (this op has waw).  When this write
// lowers the register from a current-epoch value v, the writer whose
// position is in v now has waw: its doomed byte is set.  Every writer but the
// minimum one is caught by exactly one of the two: the first registration
// that goes below a writer sees exactly that writer's value.
__device__ __forceinline__ bool finish_write(unsigned long long *r, unsigned long long cur, unsigned long long me,
                                             uint32_t tag, uint8_t *doomed, uint8_t epoch);

__device__ __forceinline__ bool register_write(unsigned long long *r, unsigned long long me,
                                               uint32_t tag, uint8_t *doomed, uint8_t epoch)
{
    return finish_write(r, probe(r), me, tag, doomed, epoch);
}

// register_write in two steps, so that a thread with several registers
// issues all their probes before the first atomic: cur is probe(r).
__device__ __forceinline__ bool finish_write(unsigned long long *r, unsigned long long cur, unsigned long long me,
                                             uint32_t tag, uint8_t *doomed, uint8_t epoch)
{
    if (cur < me) return true;
    if (cur > me) {
        const unsigned long long old = atomicMin(r, me);
        if (old < me) return true;
        if (old > me && (uint32_t)(old >> kTagShift) == tag) doomed[old & kPosMask] = epoch;
    }
    return false;
}

// Where a row's R and W registers live: interleaved (row_mul 2, w_off 1: one
// 16-byte pair per row) or split (row_mul 1, w_off rows: two arrays).  Either
// layout may be used in any epoch without clearing: a stale value at any
// index carries an older epoch's tag and compares greater than every value
// of the current epoch.
struct RegLayout {
    uint64_t row_mul;
    uint64_t w_off;
    __host__ __device__ uint64_t r(uint64_t row) const { return row * row_mul; }
    __host__ __device__ uint64_t w(uint64_t row) const { return row * row_mul + w_off; }
};

}  // namespace ltpg
