// Internals shared by the host sources of libgct_amd: device buffers, the
// plan object and the launch-argument builders.  Not part of the C ABI.
#pragma once

#include "common.hpp"

#include <memory>

namespace gct {

// Device allocation; with a stream it is stream-ordered (hipMallocAsync /
// hipFreeAsync from the device's memory pool), so a serving request neither
// allocates nor frees with a device-wide synchronisation.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    hipStream_t stream = nullptr;
    bool ordered = false;
    DevBuf() = default;
    explicit DevBuf(size_t n) : bytes(n)
    {
        if (n) hip_check(hipMalloc(&p, n), "hipMalloc");
    }
    DevBuf(size_t n, hipStream_t s, bool stream_ordered) : bytes(n), stream(s), ordered(stream_ordered)
    {
        if (!n) return;
        if (ordered) hip_check(hipMallocAsync(&p, n, s), "hipMallocAsync");
        else hip_check(hipMalloc(&p, n), "hipMalloc");
    }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes), stream(o.stream), ordered(o.ordered)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes; stream = o.stream; ordered = o.ordered;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            if (ordered) (void)hipFreeAsync(p, stream);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

struct DeviceGuard {
    int prev = 0;
    explicit DeviceGuard(int dev)
    {
        hip_check(hipGetDevice(&prev), "hipGetDevice");
        if (dev != prev) hip_check(hipSetDevice(dev), "hipSetDevice");
    }
    ~DeviceGuard() { (void)hipSetDevice(prev); }
};

}  // namespace gct

// ---- the plan ---------------------------------------------------------------

struct gct_plan {
    using DevBuf = gct::DevBuf;
    int device = 0;
    gct_ct_geom geom{};
    std::vector<gct_proj_geom> views;
    gct_image_geom img{};
    gct_recon_opts opts{};
    gct::RotationInfo rot;
    uint32_t nu = 0, nv = 0, V = 0, L = 0;
    uint32_t pitch = 0, rows_pv = 0;
    uint32_t filt_views = 0;   // views the filtered buffer holds
    bool circular = true;
    bool rows_inside = false;   // every voxel projects inside the detector rows (the LDS kernel drops its row clamp)
    bool narrow_fan = false;    // arc kinds: the kernels' fan angle needs no octant reduction (narrow_fan_ok)
    bool box_inner = false;     // the LDS kernel may drop the row clamp for boxes the detector did not cut (box_inner_ok)
    bool setup_f64 = false;
    float inv_dgamma = 0.0f;
    float scale = 1.0f;
    uint32_t z_unroll = 16;
    uint32_t view_block = 0;   // LDS kernel shape as requested (0: default of z_unroll; see kLdsShapes)
    uint32_t lds_rows = 0, lds_pairs = 0;
    bool round_robin = false;   // LDS kernel: tiles dealt round robin to the XCDs (autotuned)
    int variant = GCT_BP_LDS;
    hipStream_t stream = nullptr;     // stream-ordered plans (serving) allocate and free on it
    bool ordered = false;
    DevBuf d_brec, d_brec64, d_frec, d_red, d_H, d_tw, d_filt, d_views;
    DevBuf d_vol_pad;   // forward projection: zero-padded volume copies, allocated on first use
    bool spiral = false;   // spiral FDK: no redundancy weighting, per-slice view windows
    DevBuf d_zwin;         // spiral: [first, last) per z slice
};


namespace gct {

std::unique_ptr<gct_plan> make_plan(const gct_ct_geom &geom, const gct_image_geom &img, const gct_recon_opts &o,
                                    uint32_t filt_views, hipStream_t stream = nullptr, bool ordered = false,
                                    const std::vector<uint32_t> *spiral_win = nullptr);
FilterLaunch filter_args(gct_plan &p, const float *proj, uint32_t proj_view_offset, uint32_t first, uint32_t n,
                         uint32_t filt_view_offset);
BpLaunch bp_args(gct_plan &p, float *vol, uint32_t first, uint32_t n, bool accumulate, bool finish,
                 uint32_t filt_view_offset);
FpLaunch fp_args(gct_plan &p, const float *vol, float *proj, float *pad);
uint64_t plan_device_bytes(const gct_ct_geom &g, const gct_image_geom &img, const gct_recon_opts &o, uint32_t fviews);
uint64_t volume_bytes(const gct_image_geom &img);
size_t storage_bytes(int storage);
gct_recon_opts opts_or_default(const gct_recon_opts *o);

// Elements of a plan's filtered buffer beyond its last view.  The LDS staging
// loads of a view start at a padded row r <= nv + 2 (the view's last) and
// column c <= nu + 1, and add a lane offset of at most 63 elements, or one
// row and 31 elements in the half-row form; with pitch >= nu + 3 they reach at
// most pitch + 30 elements past the view's end (62 in the full-row form).  The
// buffer's range check does not cover the row offset (a scalar operand), so
// the buffer itself must: two spare zero rows and 64 elements do, with a row
// to spare.
inline uint64_t filt_spare_elems(uint32_t pitch) { return 2u * uint64_t(pitch) + 64u; }

// ---- batches (batch.cpp) -----------------------------------------------------
// Allocation-free pieces shared by gct_batch_* and gct_batch_execute_noalloc.

// Rejects an empty list, null plans, plans on several devices and plans with a
// partial filtered buffer.
void check_batch_plans(gct_plan *const *plans, uint32_t n);
// Work order of the scans: grouped by (kind, storage, FP64 set-up, tile type),
// larger scans first within a group, creation order among equals.
void batch_order(gct_plan *const *plans, uint32_t n, uint32_t *order);
// Groups of consecutive scans in `order` that share a persistent kernel; returns
// how many (GCT_ERR_UNSUPPORTED beyond max_groups).
uint32_t batch_groups(gct_plan *const *plans, const uint32_t *order, uint32_t n, BatchGroup *groups,
                      uint32_t max_groups);
// One scan's work-table entry, its tiles numbered from tile_begin.
BatchScan batch_entry(gct_plan &p, float *vol, uint32_t tile_begin);
// The work table in work order (tile numbers restart per group; sets the
// groups' total_tiles).
void batch_table(gct_plan *const *plans, const uint32_t *order, uint32_t n, float *const *vol, BatchGroup *groups,
                 uint32_t ngroups, BatchScan *table);
// Fused weighting + filter launches of every scan, in work order.
void batch_filters(gct_plan *const *plans, const uint32_t *order, uint32_t n, const float *const *proj, hipStream_t s);

}  // namespace gct
