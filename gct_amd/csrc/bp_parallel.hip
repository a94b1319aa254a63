// Backprojection kernels for 2-D parallel and cone-parallel detectors.
// One translation unit per geometry family keeps the template instantiations
// apart so they compile in parallel.
#include "bp_device.hpp"

namespace gct {

void launch_backproject_parallel(const BpLaunch &a, hipStream_t s)
{
    switch (a.kind) {
    case GCT_PARALLEL_2D: backproject_kind<GCT_PARALLEL_2D>(a, s); break;
    default: backproject_kind<GCT_CONE_PARALLEL>(a, s); break;
    }
}

void launch_batch_group_parallel(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    switch (g.kind) {
    case GCT_PARALLEL_2D: batch_group_kind<GCT_PARALLEL_2D>(scans, g, counter, grid, s); break;
    default: batch_group_kind<GCT_CONE_PARALLEL>(scans, g, counter, grid, s); break;
    }
}

}  // namespace gct
