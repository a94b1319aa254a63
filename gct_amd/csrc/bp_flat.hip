// Backprojection kernels for flat cone-beam detectors.
// One translation unit per geometry family keeps the template instantiations
// apart so they compile in parallel.
#include "bp_device.hpp"

namespace gct {

void launch_backproject_flat(const BpLaunch &a, hipStream_t s)
{
    backproject_kind<GCT_FLAT>(a, s);
}

void launch_batch_group_flat(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    batch_group_kind<GCT_FLAT>(scans, g, counter, grid, s);
}

}  // namespace gct
