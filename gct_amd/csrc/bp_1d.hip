// Backprojection kernels for 1-D detectors (parallel, fan arc, fan line): 2-D images.
// One translation unit per geometry family keeps the template instantiations
// apart so they compile in parallel.
#include "bp_device.hpp"

namespace gct {

void launch_backproject_1d(const BpLaunch &a, hipStream_t s)
{
    switch (a.kind) {
    case GCT_PARALLEL_1D: backproject_kind<GCT_PARALLEL_1D>(a, s); break;
    case GCT_FAN_ARC_1D: backproject_kind<GCT_FAN_ARC_1D>(a, s); break;
    default: backproject_kind<GCT_FAN_LINE_1D>(a, s); break;
    }
}

void launch_batch_group_1d(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    switch (g.kind) {
    case GCT_PARALLEL_1D: batch_group_kind<GCT_PARALLEL_1D>(scans, g, counter, grid, s); break;
    case GCT_FAN_ARC_1D: batch_group_kind<GCT_FAN_ARC_1D>(scans, g, counter, grid, s); break;
    default: batch_group_kind<GCT_FAN_LINE_1D>(scans, g, counter, grid, s); break;
    }
}

}  // namespace gct
