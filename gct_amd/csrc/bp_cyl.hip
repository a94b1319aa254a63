// Backprojection kernels for cylindrical detectors.
// One translation unit per geometry family keeps the template instantiations
// apart so they compile in parallel.
#include "bp_device.hpp"

namespace gct {

void launch_backproject_cyl(const BpLaunch &a, hipStream_t s)
{
    backproject_kind<GCT_CYLINDRICAL>(a, s);
}

void launch_batch_group_cyl(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    batch_group_kind<GCT_CYLINDRICAL>(scans, g, counter, grid, s);
}

}  // namespace gct
