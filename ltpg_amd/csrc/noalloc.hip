// noalloc.hip - a key-value epoch that allocates nothing.
//
// ltpg_kv_epoch_noalloc runs the same EXECUTE / CHECK / WRITEBACK kernels as
// ltpg_epoch_run on a key-value context, but over a table and batch the
// caller owns, with every per-row register, per-transaction flag and counter
// in `scratch`: a file-scope device array of fixed size that each GPU holds
// from module load to process end.  For callers that cannot call hipMalloc
// (code under graph capture, tight serving loops).
//
// Host code never passes `scratch` itself to a kernel (in host code the name
// is a host-side stand-in, not a device address): the first call on each
// device looks its address up with hipGetSymbolAddress, with that device
// current, and every later call on the device uses that address.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "ltpg_internal.h"

namespace ltpg {
namespace {

constexpr uint64_t kRegsBytes = 2ull * LTPG_NOALLOC_MAX_ROWS * 8;  // interleaved R / W per row
constexpr uint64_t kFlagsAt = kRegsBytes;                         // doomed, raw, war: one byte per txn each
constexpr uint64_t kCountersAt = kFlagsAt + 3ull * LTPG_NOALLOC_MAX_TXNS;
constexpr uint64_t kScratchBytes = kCountersAt + (sizeof(DeviceCounters) + 255) / 256 * 256;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// The epoch's commit count and range errors (result[0], result[1]).
__global__ void noalloc_result_kernel(const DeviceCounters *ctr, uint32_t *result)
{
    if (threadIdx.x != 0) return;
    uint32_t c = 0;
    for (int k = 0; k < kCounterStripes; ++k) c += ctr->commits[k][LTPG_TYPE_KV];
    result[0] = c;
    result[1] = ctr->range_errors;
}

// scratch's device address, per device (looked up once each)
std::mutex g_mu;
std::vector<uint8_t *> g_scratch;

int scratch_of(int device, uint8_t **out)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if ((size_t)device >= g_scratch.size()) g_scratch.resize(device + 1, nullptr);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            LTPG_FAIL(LTPG_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = static_cast<uint8_t *>(p);
    }
    *out = g_scratch[device];
    return LTPG_OK;
}

}  // namespace
}  // namespace ltpg

using namespace ltpg;

extern "C" int ltpg_kv_epoch_noalloc(uint32_t *table, uint64_t rows, uint32_t fields, uint32_t field_bytes,
                                     const ltpg_kv_batch *batch, uint32_t values, uint64_t *bitmap,
                                     uint32_t *read_values, uint32_t *result, void *hip_stream)
{
    if (!table || !batch) LTPG_FAIL(LTPG_ERR_INVALID, "null table or batch");
    if (rows == 0 || rows > LTPG_NOALLOC_MAX_ROWS)
        LTPG_FAIL(LTPG_ERR_INVALID, "rows must be in [1, LTPG_NOALLOC_MAX_ROWS]");
    if (fields == 0 || field_bytes == 0 || field_bytes % 4 || field_bytes > 1024)
        LTPG_FAIL(LTPG_ERR_INVALID, "fields >= 1 and field_bytes a multiple of 4 in [4, 1024]");
    if (values != LTPG_VALUES_COPYBACK && values != LTPG_VALUES_RMW) LTPG_FAIL(LTPG_ERR_INVALID, "unknown values");
    const uint32_t B = batch->n_txns, S = batch->n_slots;
    if (B > LTPG_NOALLOC_MAX_TXNS) LTPG_FAIL(LTPG_ERR_INVALID, "n_txns exceeds LTPG_NOALLOC_MAX_TXNS");
    if (B && (S == 0 || S > 64)) LTPG_FAIL(LTPG_ERR_INVALID, "n_slots must be in [1, 64]");
    if (B && (!batch->keys || !batch->meta || !batch->tids || !bitmap || !read_values))
        LTPG_FAIL(LTPG_ERR_INVALID, "null batch array, bitmap or read values");
    int device = 0;
    LTPG_HIP_TRY(hipGetDevice(&device));
    uint8_t *p = nullptr;
    int rc = scratch_of(device, &p);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(hip_stream);
    KvState kv;
    kv.rows = rows;
    kv.fields = fields;
    kv.field_words = field_bytes / 4;
    kv.field_stride = kv.field_words;  // the caller's table is packed
    kv.max_slots = S;
    kv.table = table;
    kv.regs = reinterpret_cast<unsigned long long *>(p);
    kv.doomed = p + kFlagsAt;
    kv.raw_flag = kv.doomed + LTPG_NOALLOC_MAX_TXNS;
    kv.war_flag = kv.raw_flag + LTPG_NOALLOC_MAX_TXNS;
    kv.read_values = read_values;
    kv.bitmap = bitmap;
    DeviceCounters *ctr = reinterpret_cast<DeviceCounters *>(p + kCountersAt);
    KvBatchDev b;
    b.n_txns = B;
    b.n_slots = S;
    b.keys = batch->keys;
    b.meta = batch->meta;
    b.tids = batch->tids;
    // the last call's contents are still there: registers back to "no
    // registrant", this batch's flags and the counters to zero (fill kernels,
    // so that a caller may capture the call into a graph)
    LTPG_HIP_TRY(launch_fill_u64(kv.regs, 2 * rows, ~0ull, s));
    LTPG_HIP_TRY(launch_fill_u8(p + kCountersAt, sizeof(DeviceCounters), 0, s));
    if (B) {
        LTPG_HIP_TRY(launch_fill_u8(kv.doomed, B, 0, s));
        LTPG_HIP_TRY(launch_fill_u8(kv.raw_flag, B, 0, s));
        LTPG_HIP_TRY(launch_fill_u8(kv.war_flag, B, 0, s));
        EpochTag e;
        e.epoch = 1;
        e.tag = 254;
        LTPG_HIP_TRY(launch_kv_execute(kv, b, e, false, ctr, s));
        LTPG_HIP_TRY(launch_kv_check(kv, b, e, false, ctr, s));
        LTPG_HIP_TRY(launch_kv_writeback(kv, b, values, s));
    }
    if (result) {
        hipLaunchKernelGGL(noalloc_result_kernel, dim3(1), dim3(64), 0, s, ctr, result);
        LTPG_HIP_TRY(hipGetLastError());
    }
    return LTPG_OK;
}
