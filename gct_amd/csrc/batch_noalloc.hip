// A batch execute that allocates nothing: gct_batch_execute_noalloc.
//
// No reference counterpart.  It is for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  gct_batch_execute keeps each batch's work table and
// work counters in buffers the batch allocates; here they live in `scratch`,
// a fixed array of this library, one per device, and the table reaches it
// through kernel arguments (no host staging buffer, no copy from host
// memory).  The caller's plans, projections and volumes are the only other
// memory touched.
#include "plan_internal.hpp"

#include <mutex>

namespace gct {
namespace {

constexpr uint32_t kCounterBytes = 256;                        // one counter per kernel group
constexpr uint32_t kMaxGroups = kCounterBytes / sizeof(uint32_t);
constexpr size_t kScratchBytes = kCounterBytes + size_t(GCT_NOALLOC_MAX_SCANS) * sizeof(BatchScan);
static_assert(sizeof(BatchScan) % 16 == 0, "table entries keep 16-byte alignment");

// Counters (kCounterBytes), then the work table.  Each device holds its own
// copy; calls on one device share it and keep the last call's contents.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device `dev`, found once
// per device (the array's host-side name is not a device address).
char *scratch_on(int dev)
{
    if (dev < 0 || dev >= kMaxDevices) fail(GCT_ERR_UNSUPPORTED, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch[dev] = static_cast<char *>(p);
    }
    return g_scratch[dev];
}

// Table entries passed by value to the kernel that stores them: 16 entries
// (2.3 KB) per launch keep the arguments well inside the 4 KB kernel
// argument limit.
constexpr uint32_t kChunk = 16;
struct TableChunk {
    uint32_t first, count, pad0, pad1;
    BatchScan scans[kChunk];
};
static_assert(sizeof(TableChunk) <= 4096);

__global__ __launch_bounds__(256) void put_table_kernel(TableChunk c, BatchScan *table)
{
    const uint32_t words = c.count * static_cast<uint32_t>(sizeof(BatchScan) / 4u);
    const uint32_t *src = reinterpret_cast<const uint32_t *>(c.scans);
    uint32_t *dst = reinterpret_cast<uint32_t *>(table + c.first);
    for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

void put_chunk(const TableChunk &c, BatchScan *table, hipStream_t s)
{
    hipLaunchKernelGGL(put_table_kernel, dim3(1), dim3(256), 0, s, c, table);
    hip_check(hipGetLastError(), "store batch table");
}

}  // namespace
}  // namespace gct

using namespace gct;

extern "C" gct_status gct_batch_execute_noalloc(gct_plan *const *plans, uint32_t n, const float *const *proj,
                                                float *const *vol, void *stream)
{
    return guarded([&] {
        if (!proj || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        check_batch_plans(plans, n);
        if (n > GCT_NOALLOC_MAX_SCANS)
            fail(GCT_ERR_INVALID_ARGUMENT, "gct_batch_execute_noalloc takes at most " +
                                               std::to_string(GCT_NOALLOC_MAX_SCANS) + " scans, got " +
                                               std::to_string(n));
        DeviceGuard dg(plans[0]->device);
        char *base = scratch_on(plans[0]->device);
        uint32_t *counters = reinterpret_cast<uint32_t *>(base);
        BatchScan *table = reinterpret_cast<BatchScan *>(base + kCounterBytes);
        hipStream_t s = static_cast<hipStream_t>(stream);

        uint32_t order[GCT_NOALLOC_MAX_SCANS];
        BatchGroup groups[kMaxGroups];
        batch_order(plans, n, order);
        const uint32_t ng = batch_groups(plans, order, n, groups, kMaxGroups);
        for (uint32_t k = 0; k < n; ++k)
            if (!proj[order[k]] || !vol[order[k]]) fail(GCT_ERR_INVALID_ARGUMENT, "null projections or volume");

        // The table, in work order, chunk by chunk (tile numbers restart per group).
        TableChunk chunk{};
        uint32_t tile = 0, gi = 0;
        for (uint32_t k = 0; k < n; ++k) {
            if (groups[gi].first_scan == k) tile = 0;
            const uint32_t oi = order[k];
            const BatchScan e = batch_entry(*plans[oi], vol[oi], tile);
            tile += e.tiles_x * e.tiles_y * e.tiles_z;
            groups[gi].total_tiles = tile;
            if (k + 1 == groups[gi].first_scan + groups[gi].num_scans) ++gi;
            if (chunk.count == 0) chunk.first = k;
            chunk.scans[chunk.count++] = e;
            if (chunk.count == kChunk || k + 1 == n) {
                put_chunk(chunk, table, s);
                chunk.count = 0;
            }
        }
        batch_filters(plans, order, n, proj, s);
        launch_zero_counters(counters, ng, s);   // a fill kernel: counters start at 0 every call
        // groups one after another on the caller's stream (no side streams or events to create)
        for (uint32_t g = 0; g < ng; ++g) launch_backproject_batch_group(table, groups[g], counters + g, s);
    });
}
