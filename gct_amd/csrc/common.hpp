// Internal declarations shared by the host sources and the HIP kernels of
// libgct_amd.  Nothing here crosses the C ABI.
#pragma once

#include "gct_amd.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace gct {

// ---- errors -------------------------------------------------------------

struct Error : std::runtime_error {
    gct_status code;
    Error(gct_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

[[noreturn]] inline void fail(gct_status c, const std::string &m) { throw Error(c, m); }

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        gct_status c = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GCT_ERR_OUT_OF_MEMORY
                                                                                    : GCT_ERR_RUNTIME;
        throw Error(c, std::string(what) + ": " + hipGetErrorString(e));
    }
}

// Runs f, converting exceptions into a status and the thread's last error.
template <class F>
gct_status guarded(F &&f)
{
    try {
        f();
        return GCT_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host allocation failed");
        return GCT_ERR_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return GCT_ERR_RUNTIME;
    }
}

// ---- geometry helpers (geometry.cpp) --------------------------------------

__host__ __device__ inline bool is_1d(int kind) { return kind <= GCT_FAN_LINE_1D; }
__host__ __device__ inline bool is_parallel_family(int kind)
{
    return kind == GCT_PARALLEL_1D || kind == GCT_PARALLEL_2D || kind == GCT_CONE_PARALLEL;
}
__host__ __device__ inline bool is_arc_row(int kind) { return kind == GCT_FAN_ARC_1D || kind == GCT_CYLINDRICAL; }

struct RotationInfo {
    int dir = 1;
    float incr = 0.0f;
    float range = 0.0f;
    uint32_t nppr = 0;
    std::vector<float> rot_ang;
    std::vector<float> rel_ang;
};

// getRotationGeom; throws GCT_ERR_RUNTIME for an invalid trajectory.
RotationInfo rotation_info(const gct_ct_geom &g);
void validate_geom(const gct_ct_geom &g);
void validate_image(const gct_ct_geom &g, const gct_image_geom &img);
// Per-slice view windows [first, last) of the spiral FDK (empty: slice skipped).
std::vector<uint32_t> spiral_windows(const gct_ct_geom &g, const gct_image_geom &img);
std::vector<double> ramp_taps_f64(int kind, uint32_t n, float width, float cutoff, int window,
                                  std::vector<float> *taps32);
uint32_t next_pow2(uint32_t n);
// FFT length of the ramp filter for rows of n samples: the shortest length
// 2^k (at least 8) or 3 x 2^k (at least 48) whose circular convolution with
// the n-tap kernel, placed at indices -floor(n/2) .. n-1-floor(n/2) mod L,
// wraps nothing into outputs 0 .. n-1: L >= n + floor(n/2).  The reference uses nextPow2(2 n)
// (ramp_filtering_fft.cuh:36-48); the kept outputs are the same linear
// convolution either way.
uint32_t fft_len(uint32_t n);
// The widest detector row fft_len allows at the largest FFT (4096 points).
constexpr uint32_t kMaxDetCols = 2730;

// ---- device-side records --------------------------------------------------

// Per-view backprojection record: 16 floats, read with one scalar load.
//   u[4], v[4], w[4]: affine forms in voxel indices (1, ix, iy, iz)
//   k[4]: kind-specific constants (see plan.cpp: build_records)
// Detector coordinates derived from a record are padded ones (column + 1,
// row + 1), the indices of the filtered views' zero-bordered layout.
struct alignas(64) BpRec {
    float u[4];
    float v[4];
    float w[4];
    float k[4];
};

// The same record in FP64 for the FP64 setup path (volumes >= 512^3).
struct alignas(64) BpRec64 {
    double u[4];
    double v[4];
    double w[4];
    double k[4];
};

// Per-view filter record: cosine-weight constants.
struct alignas(16) FiltRec {
    float cu, cv, sdd, pad;
};

// ---- kernel launchers (kernels.hip) ---------------------------------------

struct FilterLaunch {
    int kind;
    uint32_t nu, nv, L;
    uint32_t first_view, num_views;
    float du, dv;
    const float *proj;            // [V][Nv][Nu] (view index relative to proj base)
    uint32_t proj_view_offset;    // view index of proj's first view
    const FiltRec *frec;          // per view (absolute)
    const float *red;             // [V][Nu] (absolute)
    const float2 *H;              // [L]
    const float2 *tw;             // [L]
    void *filt;                   // padded output buffer (absolute)
    uint32_t pitch, rows_pv;
    int storage;
};

struct BpLaunch {
    int kind;
    int storage;
    bool circular;
    bool setup_f64;
    uint32_t nx, ny, nz;
    uint32_t nu, nv;
    uint32_t first_view, num_views;
    uint32_t pitch, rows_pv;
    const void *filt;
    const BpRec *rec;
    const BpRec64 *rec64;
    float *vol;
    float inv_dgamma;
    float scale;                  // applied when finish
    bool accumulate, finish;
    uint32_t z_unroll;
    int variant;                  // GCT_BP_DIRECT or GCT_BP_LDS
    uint32_t view_block;          // LDS kernel shape (a row of kLdsShapes, resolved by the plan):
    uint32_t lds_rows, lds_pairs; //   views per staged block, staged rows, pairs per staged row
    const uint32_t *zwin;         // spiral: [first, last) view window per z slice (device); null otherwise
    bool rows_inside;             // every voxel's row lies inside the detector: LDS kernel without row clamp
    bool narrow_fan;              // arc kinds, FP32 set-up: every voxel in front of the source and the fan
                                  //   within +-0.6 rad (fan_atan_narrow instead of the octant-reduced form)
    bool box_inner;               // a tile's voxels stay within a row of its corners' rows (Box::inner applies)
    bool round_robin;             // LDS kernel: tiles dealt to the XCDs round robin, not in contiguous runs
    // The launch covers slices [z0, z0 + zcount) of the nz-slice volume and `vol`
    // holds just those (a z slab; z0 a multiple of 32, so tiles and their boxes
    // are the whole volume's).  Whole volume: z0 = 0, zcount = nz.
    uint32_t z0, zcount;
};

// Shapes of the LDS backprojection kernel: voxels per thread along z, views
// staged per block, staged rows per view and (value, difference) pairs per
// staged row.  The LDS a workgroup stages (8 B per pair) sets how many
// workgroups share a CU: 6 for the first shape, 7 for the second, fourth and
// fifth, 5 for the third (boxes up to 48 pairs wide), 4 for the z-unroll-32
// shapes (their registers allow no more).  Deeper view blocks (fewer barriers
// per view) fit only small boxes: up to 34 pairs by 28 rows for three views of
// 16 slices, 28 by 24 for four; 32 slices span up to 48 rows on a 2x
// magnified flat cone, 34 on a 1.9x C-arm (whose boxes the last shape fits).
// The first shape of each z unroll is its default; the autotuner times them
// all.
struct LdsShape {
    uint32_t zu, vb, rows, pairs;
};
constexpr LdsShape kLdsShapes[] = {{16, 2, 40, 40}, {16, 2, 32, 40}, {16, 2, 40, 48},
                                   {16, 3, 28, 34}, {16, 4, 24, 28}, {32, 3, 48, 34},
                                   {32, 4, 36, 28}};
constexpr int kNumLdsShapes = static_cast<int>(sizeof(kLdsShapes) / sizeof(kLdsShapes[0]));

// The first shape of z unroll `zu` (0: 16) that agrees with every nonzero
// field; -1 when none does.
inline int lds_shape(uint32_t zu, uint32_t vb, uint32_t rows, uint32_t pairs)
{
    if (zu == 0u) zu = 16u;
    for (int i = 0; i < kNumLdsShapes; ++i) {
        const LdsShape &sh = kLdsShapes[i];
        if (sh.zu == zu && (vb == 0u || vb == sh.vb) && (rows == 0u || rows == sh.rows) &&
            (pairs == 0u || pairs == sh.pairs))
            return i;
    }
    return -1;
}

struct FpLaunch {
    int kind;
    uint32_t nx, ny, nz;
    float x0, y0, z0, dx, dy, dz;
    uint32_t nu, nv, nviews;
    float du, dv;
    const gct_proj_geom *views;   // device copy
    const float *vol;
    float *pad;                   // scratch for the zero-padded copies (joseph_pad_elems); null: read `vol` only
    float *proj;
};

// The forward projector's scratch: two zero-padded copies of the volume, one
// per dominant ray axis, so that the plane march clamps its taps instead of
// bounds-checking them.  Rays marching along y read [z+3][y][x+3] (voxel
// (x, y, z) at ((z+1) ny + y)(nx+3) + x+1); rays along x read the transposed
// [z+3][x][y+3] copy that follows it.  Returns 0 when the copies do not
// apply (2-D images) or do not fit 32-bit byte offsets and 24-bit strides;
// the projector then reads the volume with bounds checks.
inline uint64_t joseph_pad_elems(int kind, uint32_t nx, uint32_t ny, uint32_t nz)
{
    if (kind <= GCT_FAN_LINE_1D) return 0;
    const uint64_t sy = uint64_t(ny) * (nx + 3u), sx = uint64_t(nx) * (ny + 3u);
    const uint64_t n = (uint64_t(nz) + 3u) * (sy + sx);
    if (n >= (uint64_t(1) << 30) || sy >= (uint64_t(1) << 24) || sx >= (uint64_t(1) << 24)) return 0;
    return n;
}

void launch_redundancy(int kind, const float *fan_ang, const float *beta, uint32_t nu, uint32_t nviews,
                       float range, int parallel_value_mode, float parallel_value, float *red, hipStream_t s);
void launch_filter(const FilterLaunch &a, hipStream_t s);
void launch_backproject(const BpLaunch &a, hipStream_t s);
void launch_forward_project(const FpLaunch &a, hipStream_t s);
void launch_pack_filtered(const float *src, void *filt, int storage, uint32_t nu, uint32_t nv,
                          uint32_t nviews, uint32_t pitch, uint32_t rows_pv, hipStream_t s);
void launch_unpack_filtered(const void *filt, float *dst, int storage, uint32_t nu, uint32_t nv,
                            uint32_t nviews, uint32_t pitch, uint32_t rows_pv, hipStream_t s);

// z-extent of one work item of the persistent batched backprojection: LDS
// tiles for circular scans (the parallel kinds always are), direct tiles for
// general trajectories.
constexpr uint32_t kBatchLdsZU = 16, kBatchDirectZU = 8;
__host__ __device__ inline uint32_t batch_tile_depth(int kind, int circular)
{
    return (kind == GCT_PARALLEL_2D || kind == GCT_CONE_PARALLEL || circular) ? kBatchLdsZU : kBatchDirectZU;
}

// One work item of the persistent batched backprojection: a z-block of one scan.
struct alignas(16) BatchScan {
    BpLaunch bp;                  // per-scan parameters (views: all)
    uint32_t tiles_x, tiles_y, tiles_z;
    uint32_t tile_begin;          // prefix sum over scans
};
// Scans of a batch that share one specialised persistent kernel: a contiguous
// range of the batch's scan table, tile_begin counted from the range start.
struct BatchGroup {
    int kind, storage;
    bool lds;
    bool f64;                     // FP64 per-column set-up (the plans' own choice, kept so results match)
    uint32_t first_scan, num_scans, total_tiles;
};
void launch_backproject_batch_group(const BatchScan *scans_dev, const BatchGroup &g, uint32_t *counter, hipStream_t s);

// Per-geometry-family instantiations (bp_1d.hip, bp_parallel.hip, bp_cyl.hip,
// bp_flat.hip); `scans` points at the group's first scan.
void launch_backproject_1d(const BpLaunch &a, hipStream_t s);
void launch_backproject_parallel(const BpLaunch &a, hipStream_t s);
void launch_backproject_cyl(const BpLaunch &a, hipStream_t s);
void launch_backproject_flat(const BpLaunch &a, hipStream_t s);
void launch_batch_group_1d(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s);
void launch_batch_group_parallel(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid,
                                 hipStream_t s);
void launch_batch_group_cyl(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s);
void launch_batch_group_flat(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid,
                             hipStream_t s);
void launch_zero_counters(uint32_t *counters, uint32_t n, hipStream_t s);

}  // namespace gct
