// Device templates of the backprojection kernels (voxel-driven, direct and
// LDS-staged; persistent batch kernels) and their launch helpers.  Included
// by the per-geometry translation units bp_1d.hip, bp_parallel.hip,
// bp_cyl.hip and bp_flat.hip, each of which instantiates only its own
// geometry kinds, and by kernels.hip for the storage types.
//
// Reference stages replaced: backProjKern / backProjVoxDriven
// (back_proj.cuh:16-104) with its per-view texture copies
// (copy_host_dev.cuh:99-147) and the final setVal / operator*= on the volume
// (nd_array_dev.cuh:58-99).
#pragma once

#include "common.hpp"

#include <algorithm>
#include <cstdint>
#include <type_traits>

namespace gct {
namespace {

// ---- storage of filtered projections -------------------------------------

template <int S> struct Stor;
template <> struct Stor<GCT_STORE_F32> {
    using T = float;
    template <class P> __device__ static float ld(P p, uint32_t i) { return p[i]; }
    __device__ static T cv(float x) { return x; }
};
template <> struct Stor<GCT_STORE_F16> {
    using T = _Float16;
    template <class P> __device__ static float ld(P p, uint32_t i) { return static_cast<float>(p[i]); }
    __device__ static T cv(float x) { return static_cast<_Float16>(x); }
};
template <> struct Stor<GCT_STORE_BF16> {
    using T = uint16_t;
    template <class P> __device__ static float ld(P p, uint32_t i) { return __uint_as_float(uint32_t(p[i]) << 16); }
    __device__ static T cv(float x)
    {
        uint32_t u = __float_as_uint(x);
        u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
        return static_cast<uint16_t>(u >> 16);
    }
};

// Global-memory forms of a pointer.  Pointers that a kernel reads from a work
// table in memory (the persistent batch) are generic to the compiler, which
// then takes flat loads for what they address, and a wave-uniform load that a
// store earlier in the kernel might have clobbered stays a vector load.  Loads
// through GlobalPtr are global loads; through ConstPtr (the constant address
// space: memory no thread writes during the kernel, such as the view
// records) wave-uniform loads are scalar loads.
// (The host pass of a .hip file only type-checks these device functions: there
// the types are plain pointers.)
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using GlobalPtr = __attribute__((address_space(1))) T *;
template <class T>
using ConstPtr = __attribute__((address_space(4))) const T *;
#else
template <class T>
using GlobalPtr = T *;
template <class T>
using ConstPtr = const T *;
#endif
template <class T>
__device__ inline GlobalPtr<T> to_global(T *p)
{
    return (GlobalPtr<T>)p;
}
template <class T>
__device__ inline ConstPtr<T> to_const(const T *p)
{
    return (ConstPtr<T>)p;
}

// The horizontal difference t1 - t0 of two neighbouring filtered values, as
// every backprojection path takes it: in FP32, and for FP16 storage rounded to
// FP16 once (what the LDS kernels stage, LdsPair below), so that all paths of
// a storage form give the same bits.
template <int S>
__device__ inline float tap_diff(float t0, float t1)
{
    if constexpr (S == GCT_STORE_F16)
        return static_cast<float>(static_cast<_Float16>(t1 - t0));
    else
        return t1 - t0;
}

// A staged detector pixel of the LDS kernels: its value and the difference to
// its right neighbour.  Two FP32 values (8 B); for FP16 storage both in FP16
// (4 B: the value as stored, the difference rounded once), so that a bilinear
// update reads 8 B of LDS instead of 16 -- the LDS bandwidth binds these
// kernels -- and a horizontal tap is one v_fma_mix_f32 (FP16 operands, FP32
// arithmetic: the same bits as the FP32 FMA of the widened values).
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
template <int S> struct LdsPair {
    using T = float2;
    static constexpr uint32_t kBytes = 8;
    __device__ static T make(float v, float d) { return make_float2(v, d); }
    __device__ static float lerp(float alpha, T p) { return fmaf(alpha, p.y, p.x); }
};
template <> struct LdsPair<GCT_STORE_F16> {
    using T = uint32_t;
    static constexpr uint32_t kBytes = 4;
    __device__ static T make(float v, float d)
    {
        half2_t h;
        h.x = static_cast<_Float16>(v);   // exact: v is an FP16 value widened
        h.y = static_cast<_Float16>(d);   // tap_diff's rounding
        return __builtin_bit_cast(uint32_t, h);
    }
    __device__ static float lerp(float alpha, T p)
    {
        const half2_t h = __builtin_bit_cast(half2_t, p);
        return fmaf(alpha, static_cast<float>(h.y), static_cast<float>(h.x));
    }
};

// ---- backprojection -------------------------------------------------------

// Fan angle atan2(y, x) for the FP32 arc-detector forms.  The octant is
// reduced to t = min(|x|, |y|) / max(|x|, |y|) in [0, 1], where atan(t) is
// t * P(t^2) with P of degree 8 fitted by least squares on Chebyshev nodes
// (atan_fit.py prints the coefficients; the fit is within 6.3e-9 of atan in
// FP64 and within 1.2e-7 evaluated here, about 1.5 ulp at pi/4), and the
// octant is then restored.  atan2(0, 0) is 0, as the library's.
constexpr float kAtanCoef[9] = {9.999998808e-01f, -3.333265781e-01f, 1.998653710e-01f,
                                -1.416433305e-01f, 1.050731987e-01f, -7.247950882e-02f,
                                3.989956155e-02f, -1.445869729e-02f, 2.468246734e-03f};

__device__ inline float fan_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mn = fminf(ax, ay), mx = fmaxf(ax, ay);
    const float t = mx > 0.0f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;   // v_rcp_f32 (1 ulp)
    const float u = t * t;
    float p = kAtanCoef[8];
#pragma unroll
    for (int j = 7; j >= 0; --j) p = fmaf(p, u, kAtanCoef[j]);
    float r = t * p;
    r = ay > ax ? 1.57079637f - r : r;
    r = x < 0.0f ? 3.14159274f - r : r;
    return copysignf(r, y);
}

// Fan angle atan(nu / nw) for scans whose plan showed (BpLaunch::narrow_fan)
// that every voxel lies in front of the source in every view (nw > 0) and
// that the detector's padded columns stay within +-0.6 rad less half a column:
// no octant reduction.  t = nu / nw is clamped to +-tan(0.6), where atan(t) is
// t * P(t^2) with P of degree 6 fitted on [0, tan 0.6] (atan_fit.py: within
// 3.3e-9 of atan in FP64, 7.9e-8 evaluated here); a voxel beyond 0.6 rad takes
// the clamped angle, whose column lies outside the detector, so it reads the
// zero padding as with the exact angle.
constexpr float kAtanNarrowT = 6.841368079e-01f;   // tan(0.6) in FP32
constexpr float kAtanNarrowCoef[7] = {9.999999404e-01f, -3.333281577e-01f, 1.998677999e-01f, -1.413784027e-01f,
                                      1.024508923e-01f, -6.219936907e-02f, 2.111230977e-02f};

__device__ inline float fan_atan_narrow(float nu, float nw)
{
    const float t = __builtin_amdgcn_fmed3f(nu * __builtin_amdgcn_rcpf(nw), -kAtanNarrowT, kAtanNarrowT);
    const float u = t * t;
    float p = kAtanNarrowCoef[6];
#pragma unroll
    for (int j = 5; j >= 0; --j) p = fmaf(p, u, kAtanNarrowCoef[j]);
    return t * p;
}

// The arc detectors' fan angle of (nu, nw) in the FP32 forms.
__device__ inline float fan_angle(bool narrow, float nu, float nw)
{
    return narrow ? fan_atan_narrow(nu, nw) : fan_atan2(nu, nw);
}

struct Sample {
    uint32_t colbase;   // padded column index of the left tap
    float alpha;        // horizontal fraction
};

// Taps are addressed by padded coordinates (detector coordinate + 1), which
// the per-view records give.  Clamping one to [0, n + 1] keeps every tap
// inside the padded buffer; any coordinate outside the detector then lands on
// zero padding, which reproduces the border-zero texture
// (nd_array_texture.cuh:43-50).  The clamped coordinate is non-negative, so
// truncation is its floor and v_fract_f32 gives its fraction exactly.
__device__ inline Sample col_sample(float colp, float nup)
{
    const float c = __builtin_amdgcn_fmed3f(colp, 0.0f, nup);
    return Sample{static_cast<uint32_t>(c), __builtin_amdgcn_fractf(c)};
}

// The largest float below x > 0.
__device__ inline float below(float x)
{
    return __int_as_float(__float_as_int(x) - 1);
}

// The taps are addressed by a 32-bit byte offset from a wave-uniform row
// pointer, and the right neighbour by pointer arithmetic on that address, so a
// tap pair is one VGPR offset and two loads with immediate offsets off an SGPR
// base (a view's padded buffer is far below 4 GiB).
template <int S>
__device__ inline const typename Stor<S>::T *tap(const typename Stor<S>::T *p, uint32_t elem)
{
    return reinterpret_cast<const typename Stor<S>::T *>(reinterpret_cast<const char *>(p) +
                                                          elem * static_cast<uint32_t>(sizeof(typename Stor<S>::T)));
}

// `rowp` is the padded row coordinate (row + 1), clamped like a column.
template <int S>
__device__ inline float bilinear(const typename Stor<S>::T *__restrict__ p, uint32_t pitch, Sample cs, float rowp,
                                 float nvp)
{
    const float r = __builtin_amdgcn_fmed3f(rowp, 0.0f, nvp);
    const float beta = __builtin_amdgcn_fractf(r);
    const auto *q = tap<S>(p, static_cast<uint32_t>(r) * pitch + cs.colbase);
    const float t0 = Stor<S>::ld(q, 0), t1 = Stor<S>::ld(q, 1);
    const float b0 = Stor<S>::ld(q + pitch, 0), b1 = Stor<S>::ld(q + pitch, 1);
    const float top = fmaf(cs.alpha, tap_diff<S>(t0, t1), t0);
    const float bot = fmaf(cs.alpha, tap_diff<S>(b0, b1), b0);
    return fmaf(beta, bot - top, top);
}

template <int S>
__device__ inline float linear(const typename Stor<S>::T *__restrict__ p, uint32_t pitch, Sample cs)
{
    const auto *q = tap<S>(p + pitch, cs.colbase);
    const float t0 = Stor<S>::ld(q, 0), t1 = Stor<S>::ld(q, 1);
    return fmaf(cs.alpha, t1 - t0, t0);
}

// Per (thread, view) setup of a voxel column, shared by all z voxels.
struct ColSetup {
    float col;     // padded column coordinate (unclamped)
    float w;       // weight (constant along z unless cylindrical)
    float row0;    // padded row coordinate of the first z voxel
    float drow;    // row step per z voxel
    float rho2;    // cylindrical: in-plane squared distance
    float nv0;     // cylindrical: v numerator of the first z voxel
    float dnv;     // cylindrical: v numerator step
    double row0d, drowd, nv0d;   // row0, drow, nv0 before rounding (FP64 set-up)
};

// Z = false: the column is set up at z index 0 (fz is ignored), which drops
// the z terms without changing any result.
template <int KIND, bool F64, bool Z = true>
__device__ inline ColSetup setup_column(const BpLaunch &a, uint32_t v, float fx, float fy, float fz, float nuf)
{
    ColSetup s{};
    if constexpr (KIND == GCT_PARALLEL_2D || KIND == GCT_CONE_PARALLEL) {
        // The column is an affine form with a large constant (the detector
        // centre); evaluating it in FP64 once per column keeps it as accurate
        // as the reference's FP32 chain t = x cos + y sin, col = t / du + cu.
        const BpRec64 r = to_const(a.rec64)[v];
        s.col = static_cast<float>(fma(r.u[2], static_cast<double>(fy), fma(r.u[1], static_cast<double>(fx), r.u[0])));
        s.w = 1.0f;
        s.row0d = Z ? fma(r.v[3], static_cast<double>(fz), r.v[0]) : r.v[0];
        s.drowd = r.v[3];
        s.row0 = static_cast<float>(s.row0d);
        s.drow = static_cast<float>(r.v[3]);
    } else if constexpr (KIND == GCT_FLAT) {
        if constexpr (F64) {
            const BpRec64 r = to_const(a.rec64)[v];
            const double X = fx, Y = fy, W = fz;
            double nw = fma(r.w[2], Y, fma(r.w[1], X, r.w[0]));
            double nu = fma(r.u[2], Y, fma(r.u[1], X, r.u[0]));
            double nv = fma(r.v[2], Y, fma(r.v[1], X, r.v[0]));
            if constexpr (Z) {
                nw = fma(r.w[3], W, nw);
                nu = fma(r.u[3], W, nu);
                nv = fma(r.v[3], W, nv);
            }
            // 1 / nw: v_rcp_f64 and two Newton steps (each squares the relative
            // error) instead of the IEEE division sequence; the results are
            // rounded to FP32 next
            double rd = __builtin_amdgcn_rcp(nw);
            rd = fma(fma(-nw, rd, 1.0), rd, rd);
            rd = fma(fma(-nw, rd, 1.0), rd, rd);
            s.col = static_cast<float>(fma(nu, rd, r.k[0]));
            s.w = static_cast<float>(rd * rd);
            s.row0d = fma(nv, rd, r.k[1]);
            s.drowd = r.v[3] * rd;
            s.row0 = static_cast<float>(s.row0d);
            s.drow = static_cast<float>(s.drowd);
        } else {
            const BpRec r = to_const(a.rec)[v];
            float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
            float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
            float nv = fmaf(r.v[2], fy, fmaf(r.v[1], fx, r.v[0]));
            if constexpr (Z) {
                nw = fmaf(r.w[3], fz, nw);
                nu = fmaf(r.u[3], fz, nu);
                nv = fmaf(r.v[3], fz, nv);
            }
            const float rd = __builtin_amdgcn_rcpf(nw);   // v_rcp_f32 (1 ulp); nw = 0 gives inf as 1 / nw does
            s.col = fmaf(nu, rd, r.k[0]);
            s.w = rd * rd;
            s.row0 = fmaf(nv, rd, r.k[1]);
            s.drow = r.v[3] * rd;
        }
    } else if constexpr (KIND == GCT_CYLINDRICAL) {
        if constexpr (F64) {
            const BpRec64 r = to_const(a.rec64)[v];
            const double X = fx, Y = fy, W = fz;
            double nw = fma(r.w[2], Y, fma(r.w[1], X, r.w[0]));
            double nu = fma(r.u[2], Y, fma(r.u[1], X, r.u[0]));
            double nv = fma(r.v[2], Y, fma(r.v[1], X, r.v[0]));
            if constexpr (Z) {
                nw = fma(r.w[3], W, nw);
                nu = fma(r.u[3], W, nu);
                nv = fma(r.v[3], W, nv);
            }
            const double rho2 = fma(nu, nu, nw * nw);
            const double rr = 1.0 / sqrt(rho2);
            s.col = static_cast<float>(fma(atan2(nu, nw), static_cast<double>(a.inv_dgamma), r.k[0]));
            s.row0d = fma(r.k[2] * nv, rr, r.k[1]);
            s.drowd = r.k[2] * r.v[3] * rr;
            s.row0 = static_cast<float>(s.row0d);
            s.drow = static_cast<float>(s.drowd);
            s.rho2 = static_cast<float>(rho2);
            s.nv0d = nv;
            s.nv0 = static_cast<float>(nv);
            s.dnv = static_cast<float>(r.v[3]);
            s.w = 1.0f;   // the forms are scaled so that the weight is 1 / |r|^2
        } else {
            const BpRec r = to_const(a.rec)[v];
            float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
            float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
            float nv = fmaf(r.v[2], fy, fmaf(r.v[1], fx, r.v[0]));
            if constexpr (Z) {
                nw = fmaf(r.w[3], fz, nw);
                nu = fmaf(r.u[3], fz, nu);
                nv = fmaf(r.v[3], fz, nv);
            }
            const float rho2 = fmaf(nu, nu, nw * nw);
            const float rr = __builtin_amdgcn_rsqf(rho2);   // v_rsq_f32 (1 ulp)
            s.col = fmaf(fan_angle(a.narrow_fan, nu, nw), a.inv_dgamma, r.k[0]);
            s.row0 = fmaf(r.k[2] * nv, rr, r.k[1]);
            s.drow = r.k[2] * r.v[3] * rr;
            s.rho2 = rho2;
            s.nv0 = nv;
            s.dnv = r.v[3];
            s.w = 1.0f;   // the forms are scaled so that the weight is 1 / |r|^2 per voxel
        }
    }
    return s;
}

// Row coordinate (and cylindrical v numerator) of a column at the base zb of
// a z block: in FP64 when the set-up is (one rounding), else in FP32.
template <int KIND, bool F64>
__device__ inline float row_base(const ColSetup &s, float zb)
{
    if constexpr (F64 || KIND == GCT_PARALLEL_2D || KIND == GCT_CONE_PARALLEL)
        return static_cast<float>(fma(static_cast<double>(zb), s.drowd, s.row0d));
    else
        return fmaf(zb, s.drow, s.row0);
}

template <int KIND, bool F64>
__device__ inline float nv_base(const ColSetup &s, float zb)
{
    if constexpr (KIND != GCT_CYLINDRICAL)
        return 0.0f;
    else if constexpr (F64)
        return static_cast<float>(fma(static_cast<double>(zb), static_cast<double>(s.dnv), s.nv0d));
    else
        return fmaf(zb, s.dnv, s.nv0);
}

// A voxel column's row coordinate (and cylindrical v numerator) at slice
// z = zb + kk, zb a multiple of kZBlock: base = fma(zb, step, value at z 0),
// then fma(kk, step, base).  Every kernel splits z this way, so a result
// depends on z alone and not on how z is cut into tiles, and a view costs one
// fma per kZBlock slices besides one per slice.
constexpr uint32_t kZBlock = 16;

// Slices whose LDS taps the LDS backprojection reads before using the first:
// at z unroll 32 (4 resident waves, up to 128 VGPRs) four, at 16 (7 waves,
// up to 72) two -- more spill.  Cylindrical kernels, whose per-update weight
// holds more registers, read pairs at every z unroll.
constexpr int kLdsGroup = 4;

// Whether the row axis of a view stays parallel to z: then u, w and the
// weight are constant along a voxel column and only the row coordinate moves.
template <int KIND, bool CIRC>
constexpr bool column_invariant()
{
    return CIRC || KIND == GCT_PARALLEL_2D || KIND == GCT_CONE_PARALLEL;
}

// Flat panels whose row axis is not along z (views from projection matrices:
// tilted or turned detectors): u, v and w all move along a voxel column.  Per
// view and 16-slice z block, the numerators at the block's first slice zb are
// formed once (in FP64 with the FP64 set-up, then rounded); slice zb + kk takes
// fma(kk, step, base) of each, its reciprocal, column, row and weight -- the
// same evaluation in the direct and the LDS kernels, so they agree bit for
// bit, and a result depends on z alone, not on how z is tiled.
struct NcBase {
    float nw, nu, nv;   // numerators at the block's first slice
    float dw, du, dv;   // their steps per slice
    float k0, k1;       // padded detector centre (column, row)
};

template <bool F64>
__device__ inline NcBase nc_base(const BpLaunch &a, uint32_t v, float fx, float fy, float zb)
{
    NcBase b;
    if constexpr (F64) {
        const BpRec64 r = to_const(a.rec64)[v];
        const double X = fx, Y = fy, Z = zb;
        b.nw = static_cast<float>(fma(r.w[3], Z, fma(r.w[2], Y, fma(r.w[1], X, r.w[0]))));
        b.nu = static_cast<float>(fma(r.u[3], Z, fma(r.u[2], Y, fma(r.u[1], X, r.u[0]))));
        b.nv = static_cast<float>(fma(r.v[3], Z, fma(r.v[2], Y, fma(r.v[1], X, r.v[0]))));
    } else {
        const BpRec r = to_const(a.rec)[v];
        b.nw = fmaf(r.w[3], zb, fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0])));
        b.nu = fmaf(r.u[3], zb, fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0])));
        b.nv = fmaf(r.v[3], zb, fmaf(r.v[2], fy, fmaf(r.v[1], fx, r.v[0])));
    }
    const BpRec r = to_const(a.rec)[v];
    b.dw = r.w[3];
    b.du = r.u[3];
    b.dv = r.v[3];
    b.k0 = r.k[0];
    b.k1 = r.k[1];
    return b;
}

// Column, row (padded, unclamped) and weight of slice zb + kk; kk == 0 takes
// the bases as they are (fma(0, step, base) is base).
__device__ inline void nc_map(const NcBase &b, float kk, bool first, float &col, float &row, float &w)
{
    const float nw = first ? b.nw : fmaf(kk, b.dw, b.nw);
    const float nu = first ? b.nu : fmaf(kk, b.du, b.nu);
    const float nv = first ? b.nv : fmaf(kk, b.dv, b.nv);
    const float rd = __builtin_amdgcn_rcpf(nw);   // v_rcp_f32 (1 ulp), as the circular set-up
    col = fmaf(nu, rd, b.k0);
    row = fmaf(nv, rd, b.k1);
    w = rd * rd;
}

// One 16 x 16 x ZU voxel tile: each thread owns one (x, y) column of ZU
// voxels and accumulates over the views in order, so every voxel sums its
// views in the reference's order (fdk_axial.cuh:189-193) starting from 0 (or
// from the partial sums already in the volume when accumulating).
template <int KIND, int S, bool CIRC, bool F64, int ZU>
__device__ inline void bp_tile3d(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t tid)
{
    using T = typename Stor<S>::T;
    const uint32_t ix = bx * 16u + (tid & 15u);
    const uint32_t iy = by * 16u + (tid >> 4);
    const uint32_t iz0 = a.z0 + bz * ZU;   // global slice index (the launch may cover a z slab)
    if (ix >= a.nx || iy >= a.ny) return;
    const uint32_t zn = min(static_cast<uint32_t>(ZU), a.z0 + a.zcount - iz0);
    const size_t slice = static_cast<size_t>(a.nx) * a.ny;
    float *vp = a.vol + static_cast<size_t>(iz0 - a.z0) * slice + static_cast<size_t>(iy) * a.nx + ix;

    float acc[ZU];
#pragma unroll
    for (int k = 0; k < ZU; ++k) acc[k] = (a.accumulate && static_cast<uint32_t>(k) < zn) ? vp[k * slice] : 0.0f;

    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy), fz = static_cast<float>(iz0);
    static_assert(kZBlock % ZU == 0, "a direct tile lies within one z block");
    const float fzb = static_cast<float>(iz0 & ~(kZBlock - 1u)), fko = static_cast<float>(iz0 & (kZBlock - 1u));
    const float nuf = static_cast<float>(a.nu), nup = static_cast<float>(a.nu + 1u);
    const float nvp = static_cast<float>(a.nv + 1u);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vend = a.first_view + a.num_views;
    for (uint32_t v = a.first_view; v < vend; ++v) {
        const T *p = base + static_cast<size_t>(v) * view_elems;
        if constexpr (column_invariant<KIND, CIRC>()) {
            // set up at z index 0 and step by the absolute index, so the
            // result does not depend on how z is split into tiles
            const ColSetup s = setup_column<KIND, F64, false>(a, v, fx, fy, 0.0f, nuf);
            const Sample cs = col_sample(s.col, nup);
            const float rowb = row_base<KIND, F64>(s, fzb);
            const float nvb = nv_base<KIND, F64>(s, fzb);
#pragma unroll
            for (int k = 0; k < ZU; ++k) {
                const float kk = fko + static_cast<float>(k);
                const float val = bilinear<S>(p, a.pitch, cs, fmaf(kk, s.drow, rowb), nvp);
                float w = s.w;
                if constexpr (KIND == GCT_CYLINDRICAL) {
                    const float nvk = fmaf(kk, s.dnv, nvb);
                    w = __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                }
                acc[k] = fmaf(val, w, acc[k]);
            }
        } else if constexpr (KIND == GCT_FLAT) {
            const NcBase nb = nc_base<F64>(a, v, fx, fy, fzb);
#pragma unroll
            for (int k = 0; k < ZU; ++k) {
                float col, row, w;
                nc_map(nb, fko + static_cast<float>(k), (iz0 & (kZBlock - 1u)) + static_cast<uint32_t>(k) == 0u, col,
                       row, w);
                const float val = bilinear<S>(p, a.pitch, col_sample(col, nup), row, nvp);
                acc[k] = fmaf(val, w, acc[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < ZU; ++k) {
                const ColSetup s = setup_column<KIND, F64>(a, v, fx, fy, fz + static_cast<float>(k), nuf);
                const float val = bilinear<S>(p, a.pitch, col_sample(s.col, nup), s.row0, nvp);
                float w = s.w;
                if constexpr (KIND == GCT_CYLINDRICAL) w = __builtin_amdgcn_rcpf(fmaf(s.nv0, s.nv0, s.rho2));
                acc[k] = fmaf(val, w, acc[k]);
            }
        }
    }
    const float sc = a.scale;
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        if (static_cast<uint32_t>(k) < zn) vp[k * slice] = a.finish ? acc[k] * sc : acc[k];
}

// Column (padded, unclamped) and weight of pixel (fx, fy) in the view of
// record r, 1-D detectors.  NARROW: the arc's fan angle without octant
// reduction (BpLaunch::narrow_fan).
template <int KIND, bool NARROW>
__device__ inline void col2d(const BpLaunch &a, const BpRec &r, float fx, float fy, float &col, float &w)
{
    if constexpr (KIND == GCT_PARALLEL_1D) {
        col = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
        w = 1.0f;
    } else if constexpr (KIND == GCT_FAN_LINE_1D) {
        const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
        const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
        const float rd = __builtin_amdgcn_rcpf(nw);   // v_rcp_f32 (1 ulp), as for flat 3-D
        col = fmaf(nu, rd, r.k[0]);
        w = rd * rd;
    } else {   // GCT_FAN_ARC_1D
        const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
        const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
        col = fmaf(NARROW ? fan_atan_narrow(nu, nw) : fan_atan2(nu, nw), a.inv_dgamma, r.k[0]);
        w = __builtin_amdgcn_rcpf(fmaf(nu, nu, nw * nw));   // forms scaled by 1/sqrt(SID)
    }
}

// Calls f(std::bool_constant<NARROW>) with the launch's fan-angle form, so
// that a view loop carries no per-view branch on it.
template <int KIND, class F>
__device__ inline void with_fan_form(const BpLaunch &a, F &&f)
{
    if constexpr (KIND == GCT_FAN_ARC_1D) {
        if (a.narrow_fan) f(std::true_type{});
        else f(std::false_type{});
    } else {
        f(std::false_type{});
    }
}

// Views per group in the 2-D view loop: a group's taps are all requested
// before the first is accumulated, so their latencies overlap; the sum still
// runs view by view, so the result does not depend on the grouping.  Images
// of up to kSmall2d pixels fill the device with at most 2 waves per SIMD and
// take groups of 16 (a 256^2 slice: 0.0545 against 0.0584 ms; on larger
// images, and in the persistent batch with its 54-register budget, 16 gained
// nothing or lost).
constexpr uint32_t kViews2d = 8;
constexpr uint32_t kSmall2d = 1u << 17;

// One tile of a 2-D image (1-D detector): pixel (ix, iy) per thread.
template <int KIND, int S, uint32_t NVG = kViews2d>
__device__ inline void bp_tile2d(const BpLaunch &a, uint32_t ix, uint32_t iy)
{
    using T = typename Stor<S>::T;
    if (ix >= a.nx || iy >= a.ny) return;
    GlobalPtr<float> vp = to_global(a.vol) + static_cast<size_t>(iy) * a.nx + ix;
    float acc = a.accumulate ? *vp : 0.0f;
    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy);
    const float nup = static_cast<float>(a.nu + 1u);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const GlobalPtr<const T> base = to_global(static_cast<const T *>(a.filt)) + a.pitch;   // row 0 of view 0
    const ConstPtr<BpRec> rec = to_const(a.rec);
    const uint32_t vend = a.first_view + a.num_views;
    with_fan_form<KIND>(a, [&](auto narrow) {
        constexpr bool kNarrow = decltype(narrow)::value;
        // linear<S> on the view's row: taps at a 32-bit byte offset from the row
        auto sample = [&](uint32_t v, float &w) {
            const BpRec r = rec[v];
            float col;
            col2d<KIND, kNarrow>(a, r, fx, fy, col, w);
            const Sample cs = col_sample(col, nup);
            const auto q = (GlobalPtr<const T>)((GlobalPtr<const char>)(base + static_cast<size_t>(v) * view_elems) +
                                                cs.colbase * static_cast<uint32_t>(sizeof(T)));
            const float t0 = Stor<S>::ld(q, 0), t1 = Stor<S>::ld(q, 1);
            return fmaf(cs.alpha, t1 - t0, t0);
        };
        uint32_t v = a.first_view;
        for (; v + NVG <= vend; v += NVG) {
            float val[NVG], w[NVG];
#pragma unroll
            for (uint32_t j = 0; j < NVG; ++j) val[j] = sample(v + j, w[j]);
#pragma unroll
            for (uint32_t j = 0; j < NVG; ++j) acc = fmaf(val[j], w[j], acc);
        }
        for (; v < vend; ++v) {
            float w;
            const float val = sample(v, w);
            acc = fmaf(val, w, acc);
        }
    });
    *vp = a.finish ? acc * a.scale : acc;
}

// Blocks are dealt round-robin to the 8 XCDs; renumbering them so that each
// XCD gets a contiguous run of tiles keeps an XCD's tiles (and the detector
// rows they read) together in that XCD's L2.
__device__ inline uint32_t xcd_tile(uint32_t bid, uint32_t nblk)
{
    const uint32_t xcd = bid & 7u, local = bid >> 3;
    const uint32_t q = nblk >> 3, rmd = nblk & 7u;
    return xcd * q + min(xcd, rmd) + local;
}

template <int KIND, int S, bool CIRC, bool F64, int ZU>
__global__ __launch_bounds__(256) void bp3d_kernel(BpLaunch a)
{
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, rest = t / nbx;
    const uint32_t by = rest % nby, bz = rest / nby;
    bp_tile3d<KIND, S, CIRC, F64, ZU>(a, bx, by, bz, threadIdx.x);
}

template <int KIND, int S, uint32_t NVG>
__global__ __launch_bounds__(64) void bp2d_kernel(BpLaunch a)
{
    const uint32_t nbx = (a.nx + 7u) / 8u;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, by = t / nbx;
    bp_tile2d<KIND, S, NVG>(a, bx * 8u + (threadIdx.x & 7u), by * 8u + (threadIdx.x >> 3));
}

// ---- LDS-staged backprojection ----------------------------------------------
//
// A workgroup owns a 16 x 16 x ZU voxel tile.  For each block of VB views it
// projects the tile's 8 corner voxels, takes the bounding box of the detector
// pixels they reach (plus one pixel of margin), and stages that box from HBM
// into LDS as pairs (p[r][c], p[r][c+1] - p[r][c]) -- value and horizontal
// difference, FP32 for FP32 and BF16 storage (BF16 widened), FP16 for FP16
// storage (LdsPair) -- so each thread's bilinear tap is two LDS reads of a
// pair (one ds_read2) and one FMA per row.  The differences are those the
// direct kernel forms per update (tap_diff), so both kernels give the same
// bits.  Coordinates are clamped to the staged box: the box never
// extends past the zero padding of the filtered-projection buffer, so the
// clamp also reproduces the border-zero texture, and it keeps every LDS read
// inside the box.  A view whose box does not fit is read from HBM directly.


// Lane n + 1's value (a wavefront shift by one lane, one DPP move; lane 63
// reads zero).
__device__ inline float next_lane(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xF, 0xF, true));
}

// Minimum and maximum over each aligned group of 8 lanes, by DPP moves (no LDS
// round trip): quad_perm [1,0,3,2] and [2,3,0,1] reduce each quad, and
// row_half_mirror (lane i <-> 7 - i) pairs the two quads of a group.
template <int CTRL>
__device__ inline float dpp_move(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, 0xF, 0xF, false));
}
__device__ inline float min8(float x)
{
    x = fminf(x, dpp_move<0xB1>(x));
    x = fminf(x, dpp_move<0x4E>(x));
    return fminf(x, dpp_move<0x141>(x));
}
__device__ inline float max8(float x)
{
    x = fmaxf(x, dpp_move<0xB1>(x));
    x = fmaxf(x, dpp_move<0x4E>(x));
    return fmaxf(x, dpp_move<0x141>(x));
}

// A view's staged region.  Box row r, column c is padded pixel
// (rlo + 1 + r, clo + 1 + c).  The backprojection clamps padded coordinates
// to [c0, c1] and [r0, r1] (just below the last staged column and row, so
// both taps of a sample are staged), and the byte offset of pair (r, c) in
// the tile is 8 (ebase + padded row TP + padded column), TP being the
// staging area's pairs per row.  inner = 1: the box's rows were not cut at
// the detector's edges, so they span the rows of the tile's corners with at
// least one row of margin on each side; every voxel of a whole z tile then
// projects inside them when the plan has bounded how far a tile's interior
// exceeds its corners' rows (BpLaunch::box_inner: not at all for flat and
// parallel rows, under half a row for cylindrical ones), and the circular LDS
// kernel leaves the row clamp out for that view.
struct alignas(16) Box {
    int clo, rlo, ncols, nrows;   // ncols == 0: the view is read from HBM
    float c0, c1, r0, r1;
    int ebase, inner, pad1, pad2;
};

// PADZ: the tile's last z corner is taken past the volume's last slice, so
// the box also holds the rows of the slices a partial z tile computes and
// discards (needed where the kernel does not clamp rows).
// NC: a flat panel whose u and w move along z (nc_base); each corner is
// projected at its own z.
template <int KIND, int ZU, bool PADZ = false, bool NC = false>
__device__ inline void tile_boxes(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t vb0,
                                  uint32_t nvb, int TR, int TP, Box *boxes, uint32_t tid)
{
    if (tid >= 8u * nvb) return;
    const uint32_t j = tid >> 3, corner = tid & 7u;
    const uint32_t cx = min(bx * 16u + ((corner & 1u) ? 15u : 0u), a.nx - 1u);
    const uint32_t cy = min(by * 16u + ((corner & 2u) ? 15u : 0u), a.ny - 1u);
    const uint32_t czt = a.z0 + bz * ZU + ((corner & 4u) ? static_cast<uint32_t>(ZU - 1) : 0u);
    const uint32_t cz = PADZ ? czt : min(czt, a.nz - 1u);
    float col, row;
    if constexpr (NC) {
        float w;
        nc_map(nc_base<false>(a, vb0 + j, static_cast<float>(cx), static_cast<float>(cy), static_cast<float>(cz)),
               0.0f, true, col, row, w);
    } else {
        const float nuf = static_cast<float>(a.nu);
        const ColSetup s =
            setup_column<KIND, false, false>(a, vb0 + j, static_cast<float>(cx), static_cast<float>(cy), 0.0f, nuf);
        col = s.col;
        row = fmaf(static_cast<float>(cz), s.drow, s.row0);
    }
    // the 8 corners of a view are lanes 8j .. 8j + 7, all active
    const float cmin = min8(col), cmax = max8(col), rmin = min8(row), rmax = max8(row);
    if (corner != 0u) return;
    Box b{};
    const bool finite = cmin > -1.0e6f && cmax < 1.0e6f && rmin > -1.0e6f && rmax < 1.0e6f;
    if (finite) {
        // corner coordinates are padded (+1); clo .. chi, rlo .. rhi are detector
        // columns and rows, one beyond the corners' taps on each side
        const int clo = max(static_cast<int>(floorf(cmin)) - 2, -1);
        const int chi = min(static_cast<int>(floorf(cmax)) + 1, static_cast<int>(a.nu) + 1);
        const int rlo = max(static_cast<int>(floorf(rmin)) - 2, -1);
        const int rhi = min(static_cast<int>(floorf(rmax)) + 1, static_cast<int>(a.nv) + 1);
        const int nc = chi - clo + 1, nr = rhi - rlo + 1;
        if (nc >= 2 && nr >= 2 && nc - 1 <= TP && nr <= TR) {
            b.clo = clo;
            b.rlo = rlo;
            b.ncols = nc;
            b.nrows = nr;
            b.c0 = static_cast<float>(clo + 1);
            b.c1 = below(static_cast<float>(clo + nc));
            b.r0 = static_cast<float>(rlo + 1);
            b.r1 = below(static_cast<float>(rlo + nr));
            b.ebase = static_cast<int>(j) * (TR * TP) - (clo + 1) - (rlo + 1) * TP;
            b.inner = static_cast<int>(floorf(rmin)) - 2 >= -1 &&
                      static_cast<int>(floorf(rmax)) + 1 <= static_cast<int>(a.nv) + 1;
        }
    }
    boxes[j] = b;
}

// Staging one view's box into LDS as (value, right neighbour - value) pairs.
// The wave's rows are rw, rw + RSTEP, ... (wave-uniform; with HALF, lanes
// 32-63 take the row below), so row guards are scalar branches.  The loads
// are raw buffer loads from the view's first byte: the row's byte offset is a
// scalar operand and the lane's column offset a vector operand fixed for the
// view, so a row costs no address arithmetic.  The buffer's range check covers
// the lane's offset only, not the row's (a scalar operand), so every load must
// fall inside the filtered buffer by itself: lanes past the box's last column
// and the second half-row of a box's last row pair read at most pitch + 30
// elements past the view, which the plan's spare elements after the last view
// cover (filt_spare_elems).  Each lane loads one value and takes its right
// neighbour from the next lane (a row of m pairs needs m + 1 values).  Lanes up
// to TP store, whatever the box's width m: pairs at columns m .. TP - 1 of a
// row are never read (the clamped column keeps both taps below column m), and
// a store mask fixed for the kernel costs no compare per row.
template <int S>
__device__ inline float stage_ld(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff)
{
    if constexpr (S == GCT_STORE_F32) {
        return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(voff), static_cast<int>(soff), 0));
    } else {
        const uint32_t h = __builtin_amdgcn_raw_buffer_load_b16(rs, static_cast<int>(voff), static_cast<int>(soff), 0);
        if constexpr (S == GCT_STORE_F16)
            return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
        else
            return __uint_as_float(h << 16);
    }
}

// Rows past the box re-read the box's first row (a valid address; the value
// is not stored): no branch, and no value left undefined per row.
template <int S, int RSTEP, int NIT, bool HALF>
__device__ inline void stage_load(__amdgpu_buffer_rsrc_t rs, uint32_t row0, uint32_t row_bytes, uint32_t voff,
                                  int rw, int nrows, float *v)
{
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int r = rw + it * RSTEP;
        const uint32_t soff = r < nrows ? row0 + static_cast<uint32_t>(r) * row_bytes : row0;
        v[it] = stage_ld<S>(rs, voff, soff);
    }
}

template <int S, int TP, int RSTEP, int NIT, bool HALF>
__device__ inline void stage_store(typename LdsPair<S>::T *tl, int rw, int nrows, uint32_t lane, const float *v)
{
    const uint32_t c = HALF ? (lane & 31u) : lane;
    const uint32_t h = HALF ? (lane >> 5) : 0u;
    const bool ok_all = c < static_cast<uint32_t>(TP);   // both rows of a HALF step are in the box
    const bool ok_top = ok_all && h == 0u;              // only the upper one is
    typename LdsPair<S>::T *tw = tl + (static_cast<uint32_t>(rw) + h) * TP + c;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int r = rw + it * RSTEP;
        if (r < nrows) {
            const float d = next_lane(v[it]) - v[it];
            const bool both = !HALF || r + 1 < nrows;
            if (both ? ok_all : ok_top) tw[it * RSTEP * TP] = LdsPair<S>::make(v[it], d);
        }
    }
}

// Stages view block [vb0, vb0 + nvb) with boxes `bx`, one view at a time
// (all loads of a view are issued before its first store).  Box row r,
// column c is padded pixel (rlo + 1 + r, clo + 1 + c); a box of fewer than
// 32 pairs per row is staged two rows per wave.
template <int S, int TR, int VB, int TP>
__device__ inline void stage_block(const typename Stor<S>::T *base, size_t view_elems, uint32_t pitch, uint32_t vb0,
                                   uint32_t nvb, const Box *bx, typename LdsPair<S>::T *tile, int wave, uint32_t lane)
{
    using T = typename Stor<S>::T;
    // a view and the padded row after it (the next view's first, or the buffer's spare one)
    const uint32_t view_bytes = static_cast<uint32_t>((view_elems + pitch) * sizeof(T));   // < 2^31 (plan check)
    const uint32_t row_bytes = pitch * static_cast<uint32_t>(sizeof(T));
    for (uint32_t j = 0; j < nvb; ++j) {
        const int ncols = __builtin_amdgcn_readfirstlane(bx[j].ncols);
        if (ncols == 0) continue;
        const int clo = __builtin_amdgcn_readfirstlane(bx[j].clo);
        const int rlo = __builtin_amdgcn_readfirstlane(bx[j].rlo);
        const int nrows = __builtin_amdgcn_readfirstlane(bx[j].nrows);
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<T *>(base + static_cast<size_t>(vb0 + j) * view_elems), 0, static_cast<int>(view_bytes),
            0x00020000);
        const uint32_t row0 = (static_cast<uint32_t>(rlo + 1) * pitch + static_cast<uint32_t>(clo + 1)) *
                              static_cast<uint32_t>(sizeof(T));
        typename LdsPair<S>::T *tl = tile + j * (TR * TP);
        if (ncols - 1 < 32) {
            constexpr int kN = (TR + 7) / 8;
            float v[kN];
            const uint32_t voff = ((lane >> 5) * pitch + (lane & 31u)) * static_cast<uint32_t>(sizeof(T));
            stage_load<S, 8, kN, true>(rs, row0, row_bytes, voff, 2 * wave, nrows, v);
            stage_store<S, TP, 8, kN, true>(tl, 2 * wave, nrows, lane, v);
        } else {
            constexpr int kW = (TR + 3) / 4;
            float v[kW];
            const uint32_t voff = lane * static_cast<uint32_t>(sizeof(T));
            stage_load<S, 4, kW, false>(rs, row0, row_bytes, voff, wave, nrows, v);
            stage_store<S, TP, 4, kW, false>(tl, wave, nrows, lane, v);
        }
    }
}

// One 16 x 16 x ZU tile (bx, by, bz) of a circular 3-D scan; every thread of
// the 256-thread workgroup calls it (it holds workgroup barriers).  `tile`
// and `boxes` are the workgroup's LDS staging area of VB * TR * TP pairs
// (TR rows of TP pairs per view) and two slots of VB boxes.
// WIN: spiral scans -- slice z only takes the views [zwin[2z], zwin[2z+1]);
// the tile's view loop covers the union of its slices' windows, and the
// per-slice test is on wave-uniform values (scalar branches).
// CLAMP = false: the plan showed that every voxel's row lies inside the
// detector (BpLaunch::rows_inside), so each row lies inside its staged box
// and the per-update row clamp is left out.
template <int KIND, int S, bool F64, int ZU, int TR, int VB, int TP, bool WIN = false, bool CLAMP = true>
__device__ void bp_tile3d_lds(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t tid, float2 *tile,
                              Box *boxes)
{
    using T = typename Stor<S>::T;
    using Pair = typename LdsPair<S>::T;   // the staging area holds pairs of this type (kPB bytes)
    constexpr uint32_t kPB = LdsPair<S>::kBytes;
    Pair *ptile = reinterpret_cast<Pair *>(tile);
    // Each wave takes an 8 x 8 quadrant of the tile, and each 16 lanes of it a
    // 4 x 4 block: the lanes' detector columns then span about a quarter as
    // many LDS banks per view as a 16-voxel strip's do.
    const uint32_t ix = bx * 16u + ((tid >> 6) & 1u) * 8u + ((tid >> 4) & 1u) * 4u + (tid & 3u);
    const uint32_t iy = by * 16u + (tid >> 7) * 8u + ((tid >> 5) & 1u) * 4u + ((tid >> 2) & 3u);
    const uint32_t iz0 = a.z0 + bz * ZU;   // global slice index (the launch may cover a z slab)
    const bool active = ix < a.nx && iy < a.ny;
    const uint32_t zn = min(static_cast<uint32_t>(ZU), a.z0 + a.zcount - iz0);
    const size_t slice = static_cast<size_t>(a.nx) * a.ny;
    float *vp = a.vol + static_cast<size_t>(iz0 - a.z0) * slice + static_cast<size_t>(active ? iy : 0u) * a.nx +
                (active ? ix : 0u);

    float acc[ZU];
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        acc[k] = (active && a.accumulate && static_cast<uint32_t>(k) < zn) ? vp[k * slice] : 0.0f;

    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy), fz = static_cast<float>(iz0);
    static_assert(ZU % kZBlock == 0, "an LDS tile is whole z blocks");
    constexpr int kZB = ZU / static_cast<int>(kZBlock);
    const float nuf = static_cast<float>(a.nu), nup = static_cast<float>(a.nu + 1u);
    const float nvp = static_cast<float>(a.nv + 1u);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    uint32_t vfirst = a.first_view, vend = a.first_view + a.num_views;
    uint32_t wl[WIN ? ZU : 1], wh[WIN ? ZU : 1];
    if constexpr (WIN) {
        uint32_t lo = 0xFFFFFFFFu, hi = 0u;
#pragma unroll
        for (int k = 0; k < ZU; ++k) {
            const uint32_t z = min(iz0 + static_cast<uint32_t>(k), a.nz - 1u);
            wl[k] = a.zwin[2u * z];
            wh[k] = a.zwin[2u * z + 1u];
            if (static_cast<uint32_t>(k) < zn && wl[k] < wh[k]) {
                lo = min(lo, wl[k]);
                hi = max(hi, wh[k]);
            }
        }
        vfirst = max(vfirst, lo);
        vend = min(vend, hi);
    }

    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const uint32_t lane = tid & 63u;
    // Two barriers per view block: the boxes of block b + 1 are found after
    // block b is backprojected, into the other of the two box slots, so they
    // are ready at the barrier that also frees the staging area.
    if (vfirst < vend)
        tile_boxes<KIND, ZU, !CLAMP>(a, bx, by, bz, vfirst, min(static_cast<uint32_t>(VB), vend - vfirst), TR, TP,
                                     boxes, tid);
    __syncthreads();
    // One view block: stage, backproject, find the next block's boxes.  The
    // fast form serves blocks whose boxes all fit the staging area and has no
    // per-view branch; the general form reads a view whose box does not fit
    // from HBM.  A tile runs the fast form until its first block with such a
    // view and the general form from there on, so the accumulators are not
    // merged from two branches per view (the compiler then copies all of them
    // every view).  Both forms take the views in the same order.
    uint32_t cur = 0, vb0 = vfirst;
    auto view_block = [&](auto general) __attribute__((always_inline)) {
        constexpr bool kGeneral = decltype(general)::value;
        const uint32_t nvb = min(static_cast<uint32_t>(VB), vend - vb0);
        const Box *bc = boxes + cur * VB;
        stage_block<S, TR, VB, TP>(base, view_elems, a.pitch, vb0, nvb, bc, ptile, wave, lane);
        __syncthreads();
        if (active) {
            for (uint32_t j = 0; j < nvb; ++j) {
                const uint32_t v = vb0 + j;
                const ColSetup s = setup_column<KIND, F64, false>(a, v, fx, fy, 0.0f, nuf);
                const Box b = bc[j];
                if (!kGeneral || b.ncols != 0) {
                    // Padded coordinates, as in col_sample / bilinear.  The box clamps never
                    // bind for a coordinate of this tile, which the box covers with a margin.
                    const float clp = __builtin_amdgcn_fmed3f(s.col, b.c0, b.c1);
                    const float alpha = __builtin_amdgcn_fractf(clp);
                    // The per-view terms of the byte offset fold into one base (modulo 2^32:
                    // it is negative when the box starts below row 0 of view j's slot); the
                    // padded row index is added with one 24-bit multiply-add.
                    const uint32_t obase = kPB * static_cast<uint32_t>(b.ebase + static_cast<int>(clp));
                    const float rlo = b.r0, rhi = b.r1;
                    float rowb[kZB], nvb[kZB];
#pragma unroll
                    for (int g = 0; g < kZB; ++g) {
                        const float zb = fz + static_cast<float>(g * static_cast<int>(kZBlock));
                        rowb[g] = row_base<KIND, F64>(s, zb);
                        nvb[g] = nv_base<KIND, F64>(s, zb);
                    }
                    // The updates of the view, with or without the row clamp (kClamp).
                    auto updates = [&](auto clamp_t) __attribute__((always_inline)) {
                        constexpr bool kClamp = decltype(clamp_t)::value;
                        // the v step in a vector register, so that fma(k, dnv, base) takes k as a
                        // literal (with a scalar dnv each k != 0, 1, 2, 4 needs a move first)
                        float dnv = s.dnv;
                        if constexpr (KIND == GCT_CYLINDRICAL) asm("" : "=v"(dnv) : "0"(dnv));
                        // an opaque copy of the row step: the rows are formed here, one per
                        // update, not hoisted above the branch (16 more live registers)
                        float drow = s.drow;
                        asm("" : "+v"(drow));
                        // G slices at a time: their LDS reads are issued before the first is
                        // used, so a wave keeps several in flight (cylindrical tiles: pairs at
                        // every z-unroll; their per-update weight takes registers of its own)
                        constexpr int G = WIN ? 1 : (KIND == GCT_CYLINDRICAL || ZU < 32) ? 2 : kLdsGroup;
#pragma unroll
                        for (int k0 = 0; k0 < ZU; k0 += G) {
                            Pair tp[G], bt[G];
                            float beta[G];
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                const int k = k0 + i;
                                if constexpr (WIN) {
                                    if (v < wl[k] || v >= wh[k]) continue;
                                }
                                const float kk = static_cast<float>(k % static_cast<int>(kZBlock));
                                const int g = k / static_cast<int>(kZBlock);
                                // fma(0, drow, base) is base: slice 0 of a block takes the base as is
                                const float rk = k % static_cast<int>(kZBlock) == 0 ? rowb[g] : fmaf(kk, drow, rowb[g]);
                                const float row = kClamp ? __builtin_amdgcn_fmed3f(rk, rlo, rhi) : rk;
                                beta[i] = __builtin_amdgcn_fractf(row);
                                const uint32_t off = __umul24(static_cast<uint32_t>(row), kPB * TP) + obase;
                                const Pair *q = reinterpret_cast<const Pair *>(reinterpret_cast<const char *>(ptile) + off);
                                tp[i] = q[0];
                                bt[i] = q[TP];
                            }
                            // keep the group's reads ahead of its arithmetic (the scheduler
                            // otherwise moves each read next to its use)
                            if constexpr (G > 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                            for (int i = 0; i < G; ++i) {
                                const int k = k0 + i;
                                if constexpr (WIN) {
                                    if (v < wl[k] || v >= wh[k]) continue;
                                }
                                const float kk = static_cast<float>(k % static_cast<int>(kZBlock));
                                const int g = k / static_cast<int>(kZBlock);
                                const float top = LdsPair<S>::lerp(alpha, tp[i]);
                                const float bot = LdsPair<S>::lerp(alpha, bt[i]);
                                const float val = fmaf(beta[i], bot - top, top);
                                float w = s.w;
                                if constexpr (KIND == GCT_CYLINDRICAL) {
                                    const float nvk = k % static_cast<int>(kZBlock) == 0 ? nvb[g] : fmaf(kk, dnv, nvb[g]);
                                    w = __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                                }
                                acc[k] = fmaf(val, w, acc[k]);
                            }
                        }
                    };
                    // A whole z tile whose box the detector's edges did not cut needs no
                    // row clamp (Box::inner); the flag is the same for the whole workgroup.
                    if constexpr (CLAMP && !WIN) {
                        if (a.box_inner && __builtin_amdgcn_readfirstlane(b.inner) != 0 &&
                            iz0 + static_cast<uint32_t>(ZU) <= a.nz)
                            updates(std::false_type{});
                        else
                            updates(std::true_type{});
                    } else {
                        updates(std::bool_constant<CLAMP>{});
                    }
                } else {
                    const T *p = base + static_cast<size_t>(v) * view_elems;
                    const Sample cs = col_sample(s.col, nup);
                    float rowb[kZB], nvb[kZB];
#pragma unroll
                    for (int g = 0; g < kZB; ++g) {
                        const float zb = fz + static_cast<float>(g * static_cast<int>(kZBlock));
                        rowb[g] = row_base<KIND, F64>(s, zb);
                        nvb[g] = nv_base<KIND, F64>(s, zb);
                    }
#pragma unroll
                    for (int k = 0; k < ZU; ++k) {
                        if constexpr (WIN) {
                            if (v < wl[k] || v >= wh[k]) continue;
                        }
                        const float kk = static_cast<float>(k % static_cast<int>(kZBlock));
                        const int g = k / static_cast<int>(kZBlock);
                        const float val = bilinear<S>(p, a.pitch, cs, fmaf(kk, s.drow, rowb[g]), nvp);
                        float w = s.w;
                        if constexpr (KIND == GCT_CYLINDRICAL) {
                            const float nvk = fmaf(kk, s.dnv, nvb[g]);
                            w = __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                        }
                        acc[k] = fmaf(val, w, acc[k]);
                    }
                }
            }
        }
        const uint32_t vn = vb0 + VB;
        if (vn < vend)
            tile_boxes<KIND, ZU, !CLAMP>(a, bx, by, bz, vn, min(static_cast<uint32_t>(VB), vend - vn), TR, TP,
                                         boxes + (cur ^ 1u) * VB, tid);
        __syncthreads();
        vb0 = vn;
        cur ^= 1u;
    };
    for (; vb0 < vend;) {
        const uint32_t nvb = min(static_cast<uint32_t>(VB), vend - vb0);
        const Box *bc = boxes + cur * VB;
        bool fits = true;
#pragma unroll
        for (int j = 0; j < VB; ++j)
            fits = fits && (static_cast<uint32_t>(j) >= nvb || bc[j].ncols != 0);
        if (!__builtin_amdgcn_readfirstlane(fits)) break;
        view_block(std::false_type{});
    }
    for (; vb0 < vend;) view_block(std::true_type{});
    if (!active) return;
    const float sc = a.scale;
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        if (static_cast<uint32_t>(k) < zn) vp[k * slice] = a.finish ? acc[k] * sc : acc[k];
}

// Occupancy target of an LDS kernel variant: workgroups (one wave per SIMD
// each) that fit the 160 KiB of LDS per CU, capped by what its registers
// reach (z-unroll 16: 7, i.e. <= 72 VGPRs; 32: 4, <= 128).  More resident
// workgroups hide more of the staging latency: measured, the 6- and
// 7-workgroup shapes of z-unroll 16 beat the 5-workgroup one; z-unroll 32
// halves the per-view column set-up per update at 4.
template <int ZU, int TR, int VB, int TP>
constexpr int lds_waves_per_simd()
{
    constexpr int bytes = VB * TR * TP * 8 + 2 * VB * static_cast<int>(sizeof(Box));
    constexpr int by_lds = (160 * 1024) / bytes;
    constexpr int by_regs = ZU >= 32 ? 4 : 7;
    return by_lds < 1 ? 1 : (by_lds < by_regs ? by_lds : by_regs);
}

template <int KIND, int S, bool F64, int ZU, int TR, int VB, int TP, bool WIN = false, bool CLAMP = true>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(lds_waves_per_simd<ZU, TR, VB, TP>(), 8)))
void bp3d_lds_kernel(BpLaunch a)
{
    __shared__ float2 tile[VB * TR * TP];
    __shared__ Box boxes[2 * VB];
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    // Contiguous runs of tiles per XCD keep the detector region an XCD reads
    // small (a cylindrical slab: 12.0 against 13.8 ms); dealt round robin, all
    // XCDs read the same views' rows at once from the Infinity Cache they share
    // and finish together (the flat cone: 3.05 against 3.19 ms).  The
    // autotuner picks per scan shape.
    const uint32_t t = a.round_robin ? blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, rest = t / nbx;
    bp_tile3d_lds<KIND, S, F64, ZU, TR, VB, TP, WIN, CLAMP>(a, bx, rest % nby, rest / nby, threadIdx.x, tile, boxes);
}

// One 16 x 16 x 16 tile of a flat-panel scan whose row axis is not along z
// (views from projection matrices; NcBase): the staging of bp_tile3d_lds --
// boxes from the tile's corners, each projected at its own z, staged as
// (value, difference) pairs -- with column, row and weight evaluated per
// voxel (nc_map) and both coordinates clamped to the box.  The box covers
// the tile's taps with a margin, so the clamps bind only where the detector's
// edge cuts the box, where they reproduce its zero border; per update the
// taps and arithmetic are those of the direct kernel's non-circular path, so
// the two agree bit for bit.
template <int S, bool F64, int TR, int VB, int TP>
__device__ void bp_tile3d_lds_nc(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t tid,
                                 float2 *tile, Box *boxes)
{
    constexpr int ZU = static_cast<int>(kZBlock);
    using T = typename Stor<S>::T;
    using Pair = typename LdsPair<S>::T;   // the staging area holds pairs of this type (kPB bytes)
    constexpr uint32_t kPB = LdsPair<S>::kBytes;
    Pair *ptile = reinterpret_cast<Pair *>(tile);
    const uint32_t ix = bx * 16u + ((tid >> 6) & 1u) * 8u + ((tid >> 4) & 1u) * 4u + (tid & 3u);
    const uint32_t iy = by * 16u + (tid >> 7) * 8u + ((tid >> 5) & 1u) * 4u + ((tid >> 2) & 3u);
    const uint32_t iz0 = a.z0 + bz * ZU;   // global slice index, the first of a z block
    const bool active = ix < a.nx && iy < a.ny;
    const uint32_t zn = min(static_cast<uint32_t>(ZU), a.z0 + a.zcount - iz0);
    const size_t slice = static_cast<size_t>(a.nx) * a.ny;
    float *vp = a.vol + static_cast<size_t>(iz0 - a.z0) * slice + static_cast<size_t>(active ? iy : 0u) * a.nx +
                (active ? ix : 0u);

    float acc[ZU];
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        acc[k] = (active && a.accumulate && static_cast<uint32_t>(k) < zn) ? vp[k * slice] : 0.0f;

    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy), fz = static_cast<float>(iz0);
    const float nup = static_cast<float>(a.nu + 1u), nvp = static_cast<float>(a.nv + 1u);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vfirst = a.first_view, vend = a.first_view + a.num_views;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(tid >> 6));
    const uint32_t lane = tid & 63u;
    if (vfirst < vend)
        tile_boxes<GCT_FLAT, ZU, false, true>(a, bx, by, bz, vfirst, min(static_cast<uint32_t>(VB), vend - vfirst), TR,
                                              TP, boxes, tid);
    __syncthreads();
    // fast and general view-block forms as in bp_tile3d_lds
    uint32_t cur = 0, vb0 = vfirst;
    auto view_block = [&](auto general) __attribute__((always_inline)) {
        constexpr bool kGeneral = decltype(general)::value;
        const uint32_t nvb = min(static_cast<uint32_t>(VB), vend - vb0);
        const Box *bc = boxes + cur * VB;
        stage_block<S, TR, VB, TP>(base, view_elems, a.pitch, vb0, nvb, bc, ptile, wave, lane);
        __syncthreads();
        if (active) {
            for (uint32_t j = 0; j < nvb; ++j) {
                const uint32_t v = vb0 + j;
                const NcBase nb = nc_base<F64>(a, v, fx, fy, fz);
                const Box b = bc[j];
                if (!kGeneral || b.ncols != 0) {
                    // byte offset of pair (r, c): kPB (ebase + c) + kPB TP r, modulo 2^32
                    const uint32_t eb = kPB * static_cast<uint32_t>(b.ebase);
                    constexpr int G = 2;
#pragma unroll
                    for (int k0 = 0; k0 < ZU; k0 += G) {
                        Pair tp[G], bt[G];
                        float alpha[G], beta[G], w[G];
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const int k = k0 + i;
                            float col, row;
                            nc_map(nb, static_cast<float>(k), k == 0, col, row, w[i]);
                            const float cc = __builtin_amdgcn_fmed3f(col, b.c0, b.c1);
                            const float rc = __builtin_amdgcn_fmed3f(row, b.r0, b.r1);
                            alpha[i] = __builtin_amdgcn_fractf(cc);
                            beta[i] = __builtin_amdgcn_fractf(rc);
                            const uint32_t off = __umul24(static_cast<uint32_t>(rc), kPB * TP) + eb +
                                                 kPB * static_cast<uint32_t>(cc);
                            const Pair *q = reinterpret_cast<const Pair *>(reinterpret_cast<const char *>(ptile) + off);
                            tp[i] = q[0];
                            bt[i] = q[TP];
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int i = 0; i < G; ++i) {
                            const float top = LdsPair<S>::lerp(alpha[i], tp[i]);
                            const float bot = LdsPair<S>::lerp(alpha[i], bt[i]);
                            const float val = fmaf(beta[i], bot - top, top);
                            acc[k0 + i] = fmaf(val, w[i], acc[k0 + i]);
                        }
                    }
                } else {
                    const T *p = base + static_cast<size_t>(v) * view_elems;
#pragma unroll
                    for (int k = 0; k < ZU; ++k) {
                        float col, row, w;
                        nc_map(nb, static_cast<float>(k), k == 0, col, row, w);
                        const float val = bilinear<S>(p, a.pitch, col_sample(col, nup), row, nvp);
                        acc[k] = fmaf(val, w, acc[k]);
                    }
                }
            }
        }
        const uint32_t vn = vb0 + VB;
        if (vn < vend)
            tile_boxes<GCT_FLAT, ZU, false, true>(a, bx, by, bz, vn, min(static_cast<uint32_t>(VB), vend - vn), TR, TP,
                                                  boxes + (cur ^ 1u) * VB, tid);
        __syncthreads();
        vb0 = vn;
        cur ^= 1u;
    };
    for (; vb0 < vend;) {
        const uint32_t nvb = min(static_cast<uint32_t>(VB), vend - vb0);
        const Box *bc = boxes + cur * VB;
        bool fits = true;
#pragma unroll
        for (int j = 0; j < VB; ++j)
            fits = fits && (static_cast<uint32_t>(j) >= nvb || bc[j].ncols != 0);
        if (!__builtin_amdgcn_readfirstlane(fits)) break;
        view_block(std::false_type{});
    }
    for (; vb0 < vend;) view_block(std::true_type{});
    if (!active) return;
    const float sc = a.scale;
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        if (static_cast<uint32_t>(k) < zn) vp[k * slice] = a.finish ? acc[k] * sc : acc[k];
}

// Non-circular flat scans, LDS-staged: the shape (16, 2, 40, 40), at 5
// workgroups per CU (up to 96 VGPRs: the per-voxel column, row and weight
// spill at the 6 its LDS allows).
template <int S, bool F64>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
void bp3d_lds_nc_kernel(BpLaunch a)
{
    __shared__ float2 tile[2 * 40 * 40];
    __shared__ Box boxes[2 * 2];
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t t = a.round_robin ? blockIdx.x : xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, rest = t / nbx;
    bp_tile3d_lds_nc<S, F64, 40, 2, 40>(a, bx, rest % nby, rest / nby, threadIdx.x, tile, boxes);
}

// 2-D images: each block of views has its (padded) detector rows staged whole
// in LDS; a 16 x 16 pixel tile per workgroup.
template <int KIND, int S>
__global__ __launch_bounds__(256) void bp2d_lds_kernel(BpLaunch a, uint32_t vb)
{
    using T = typename Stor<S>::T;
    extern __shared__ float rows[];
    const uint32_t tid = threadIdx.x;
    const uint32_t nbx = (a.nx + 15u) / 16u;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t ix = (t % nbx) * 16u + (tid & 15u), iy = (t / nbx) * 16u + (tid >> 4);
    const bool active = ix < a.nx && iy < a.ny;
    float *vp = a.vol + static_cast<size_t>(active ? iy : 0u) * a.nx + (active ? ix : 0u);
    float acc = (active && a.accumulate) ? *vp : 0.0f;
    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy);
    const float nup = static_cast<float>(a.nu + 1u);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vend = a.first_view + a.num_views;
    for (uint32_t vb0 = a.first_view; vb0 < vend; vb0 += vb) {
        const uint32_t nvb = min(vb, vend - vb0);
        for (uint32_t j = 0; j < nvb; ++j) {
            const T *p = base + static_cast<size_t>(vb0 + j) * view_elems + a.pitch;
            for (uint32_t c = tid; c < a.pitch; c += 256u) rows[j * a.pitch + c] = Stor<S>::ld(p, c);
        }
        __syncthreads();
        if (active) {
            with_fan_form<KIND>(a, [&](auto narrow) {
                for (uint32_t j = 0; j < nvb; ++j) {
                    float col, w;
                    col2d<KIND, decltype(narrow)::value>(a, a.rec[vb0 + j], fx, fy, col, w);
                    const Sample cs = col_sample(col, nup);
                    const float *rr = rows + j * a.pitch + cs.colbase;
                    const float t0 = rr[0], t1 = rr[1];
                    acc = fmaf(fmaf(cs.alpha, t1 - t0, t0), w, acc);
                }
            });
        }
        __syncthreads();
    }
    if (active) *vp = a.finish ? acc * a.scale : acc;
}

// ---- persistent batched backprojection -------------------------------------

constexpr int kBatchZU = static_cast<int>(kBatchDirectZU);   // direct tiles (general trajectories)
// LDS tiles of batched circular scans: (16, 2, 32, 40), 7 workgroups per CU
// (the cylindrical slabs of the uneven batch: 11.95 against 12.5 ms per slab
// with (16, 2, 40, 40))
constexpr int kBatchVB = static_cast<int>(kLdsShapes[1].vb);
constexpr int kBatchTR = static_cast<int>(kLdsShapes[1].rows);
constexpr int kBatchTP = static_cast<int>(kLdsShapes[1].pairs);
static_assert(kLdsShapes[1].zu == kBatchLdsZU, "batched LDS tiles use a z-unroll-16 shape");

// Persistent backprojection over the scans of one group (same geometry kind,
// storage, column set-up precision and tile type; a batch launches one such
// kernel per group).  Work items are (scan, tile) pairs numbered scan by scan;
// each workgroup takes items from the group's device counter until none are
// left, so the grid drains on its own whatever the mix of scan sizes.  F64:
// the FP64 per-column set-up of the scans' own plans (cone scans of 512^3 and
// more), so that batched volumes equal the plans' own bit for bit.
template <int KIND, int S, bool LDS, bool F64 = false>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(LDS ? lds_waves_per_simd<static_cast<int>(kBatchLdsZU), kBatchTR, kBatchVB, kBatchTP>()
                                       : 1, 8)))
void bp_batch_kernel(const BatchScan *__restrict__ scans, uint32_t nscans,
                                                       uint32_t total, uint32_t *counter)
{
    constexpr bool k1d = KIND <= GCT_FAN_LINE_1D;
    __shared__ uint32_t item;
    __shared__ float2 tile[LDS ? kBatchVB * kBatchTR * kBatchTP : 1];
    __shared__ Box boxes[2 * kBatchVB];
    const uint32_t tid = threadIdx.x;
    for (;;) {
        if (tid == 0) item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t t = item;
        __syncthreads();
        if (t >= total) break;
        uint32_t lo = 0, hi = nscans;   // last scan with tile_begin <= t
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (scans[mid].tile_begin <= t) lo = mid; else hi = mid;
        }
        const BatchScan &sc = scans[lo];
        const BpLaunch &a = sc.bp;
        const uint32_t local = t - sc.tile_begin;
        if constexpr (k1d) {
            const uint32_t bx = local % sc.tiles_x, by = local / sc.tiles_x;
            bp_tile2d<KIND, S>(a, bx * 16u + (tid & 15u), by * 16u + (tid >> 4));
        } else {
            const uint32_t bx = local % sc.tiles_x, rest = local / sc.tiles_x;
            const uint32_t by = rest % sc.tiles_y, bz = rest / sc.tiles_y;
            if constexpr (LDS)
                bp_tile3d_lds<KIND, S, F64, static_cast<int>(kBatchLdsZU), kBatchTR, kBatchVB, kBatchTP>(
                    a, bx, by, bz, tid, tile, boxes);
            else
                bp_tile3d<KIND, S, false, F64, kBatchZU>(a, bx, by, bz, tid);
        }
    }
}

// ---- launch helpers ---------------------------------------------------------

template <int KIND, int S, bool CIRC, bool F64>
void launch_bp3d_zu(const BpLaunch &a, hipStream_t s)
{
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t zu = a.z_unroll;
    const uint32_t nbz = (a.zcount + zu - 1u) / zu;
    const uint32_t blocks = nbx * nby * nbz;
    switch (zu) {
    case 4: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 4>), dim3(blocks), dim3(256), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 16>), dim3(blocks), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 8>), dim3(blocks), dim3(256), 0, s, a); break;
    }
}

template <int KIND, int S, bool F64, bool CLAMP>
void launch_lds_shape(int shape, dim3 g, dim3 b, const BpLaunch &a, hipStream_t s)
{
    switch (shape) {
    case 1: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 32, 2, 40, false, CLAMP>), g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 40, 2, 48, false, CLAMP>), g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 28, 3, 34, false, CLAMP>), g, b, 0, s, a); break;
    case 4: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 24, 4, 28, false, CLAMP>), g, b, 0, s, a); break;
    case 5: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 32, 48, 3, 34, false, CLAMP>), g, b, 0, s, a); break;
    case 6: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 32, 36, 4, 28, false, CLAMP>), g, b, 0, s, a); break;
    default: hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 40, 2, 40, false, CLAMP>), g, b, 0, s, a); break;
    }
}

template <int KIND, int S, bool F64>
void launch_bp3d_lds(const BpLaunch &a, hipStream_t s)
{
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const dim3 b(256);
    if constexpr (KIND == GCT_FLAT || KIND == GCT_CYLINDRICAL) {
        if (a.zwin) {   // spiral: per-slice view windows, default shape
            hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 40, 2, 40, true>),
                               dim3(nbx * nby * ((a.zcount + 15u) / 16u)), b, 0, s, a);
            return;
        }
    }
    // the plan resolved (z_unroll, view_block, lds_rows, lds_pairs) to a row of kLdsShapes
    const int shape = lds_shape(a.z_unroll, a.view_block, a.lds_rows, a.lds_pairs);
    const uint32_t zu = shape >= 0 ? kLdsShapes[shape].zu : 16u;
    const dim3 g(nbx * nby * ((a.zcount + zu - 1u) / zu));
    static_assert(kLdsShapes[0].zu == 16 && kLdsShapes[0].vb == 2 && kLdsShapes[0].rows == 40 && kLdsShapes[0].pairs == 40);
    static_assert(kLdsShapes[1].zu == 16 && kLdsShapes[1].vb == 2 && kLdsShapes[1].rows == 32 && kLdsShapes[1].pairs == 40);
    static_assert(kLdsShapes[2].zu == 16 && kLdsShapes[2].vb == 2 && kLdsShapes[2].rows == 40 && kLdsShapes[2].pairs == 48);
    static_assert(kLdsShapes[3].zu == 16 && kLdsShapes[3].vb == 3 && kLdsShapes[3].rows == 28 && kLdsShapes[3].pairs == 34);
    static_assert(kLdsShapes[4].zu == 16 && kLdsShapes[4].vb == 4 && kLdsShapes[4].rows == 24 && kLdsShapes[4].pairs == 28);
    static_assert(kLdsShapes[5].zu == 32 && kLdsShapes[5].vb == 3 && kLdsShapes[5].rows == 48 && kLdsShapes[5].pairs == 34);
    static_assert(kLdsShapes[6].zu == 32 && kLdsShapes[6].vb == 4 && kLdsShapes[6].rows == 36 && kLdsShapes[6].pairs == 28);
    static_assert(kNumLdsShapes == 7);
    if (a.rows_inside) launch_lds_shape<KIND, S, F64, false>(shape, g, b, a, s);
    else launch_lds_shape<KIND, S, F64, true>(shape, g, b, a, s);
}

template <int KIND, int S>
void launch_bp3d_variant(const BpLaunch &a, hipStream_t s)
{
    const bool lds = a.variant == GCT_BP_LDS;
    if constexpr (KIND == GCT_FLAT || KIND == GCT_CYLINDRICAL) {
        if (lds && a.circular) {
            if (a.setup_f64) launch_bp3d_lds<KIND, S, true>(a, s);
            else launch_bp3d_lds<KIND, S, false>(a, s);
        } else if (KIND == GCT_FLAT && lds) {
            const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
            const dim3 g(nbx * nby * ((a.zcount + kZBlock - 1u) / kZBlock));
            if (a.setup_f64) hipLaunchKernelGGL((bp3d_lds_nc_kernel<S, true>), g, dim3(256), 0, s, a);
            else hipLaunchKernelGGL((bp3d_lds_nc_kernel<S, false>), g, dim3(256), 0, s, a);
        } else if (a.setup_f64) {
            if (a.circular) launch_bp3d_zu<KIND, S, true, true>(a, s);
            else launch_bp3d_zu<KIND, S, false, true>(a, s);
        } else {
            if (a.circular) launch_bp3d_zu<KIND, S, true, false>(a, s);
            else launch_bp3d_zu<KIND, S, false, false>(a, s);
        }
    } else {
        if (lds) launch_bp3d_lds<KIND, S, false>(a, s);
        else launch_bp3d_zu<KIND, S, true, false>(a, s);
    }
}

template <int KIND, int S>
void launch_bp2d(const BpLaunch &a, hipStream_t s)
{
    if (a.variant == GCT_BP_LDS) {
        const uint32_t vb = std::max(1u, std::min(16u, 32768u / (a.pitch * 4u)));
        const uint32_t blocks = ((a.nx + 15u) / 16u) * ((a.ny + 15u) / 16u);
        hipLaunchKernelGGL((bp2d_lds_kernel<KIND, S>), dim3(blocks), dim3(256), vb * a.pitch * sizeof(float), s, a, vb);
    } else {
        const uint32_t blocks = ((a.nx + 7u) / 8u) * ((a.ny + 7u) / 8u);
        if (static_cast<uint64_t>(a.nx) * a.ny <= kSmall2d)
            hipLaunchKernelGGL((bp2d_kernel<KIND, S, 16>), dim3(blocks), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL((bp2d_kernel<KIND, S, kViews2d>), dim3(blocks), dim3(64), 0, s, a);
    }
}
template <int KIND, int S>
void launch_batch_kind(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    if constexpr (KIND <= GCT_FAN_LINE_1D) {
        hipLaunchKernelGGL((bp_batch_kernel<KIND, S, false>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                           g.total_tiles, counter);
    } else if constexpr (KIND == GCT_FLAT || KIND == GCT_CYLINDRICAL) {
        if (g.lds && g.f64)
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, true, true>), dim3(grid), dim3(256), 0, s, scans,
                               g.num_scans, g.total_tiles, counter);
        else if (g.lds)
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, true>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                               g.total_tiles, counter);
        else if (g.f64)
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, false, true>), dim3(grid), dim3(256), 0, s, scans,
                               g.num_scans, g.total_tiles, counter);
        else
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, false>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                               g.total_tiles, counter);
    } else {
        hipLaunchKernelGGL((bp_batch_kernel<KIND, S, true>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                           g.total_tiles, counter);
    }
}

// Backprojection of one geometry kind (all storage types).
template <int KIND>
void backproject_kind(const BpLaunch &a, hipStream_t s)
{
    switch (a.storage) {
    case GCT_STORE_F16:
        if constexpr (KIND <= GCT_FAN_LINE_1D) launch_bp2d<KIND, GCT_STORE_F16>(a, s);
        else launch_bp3d_variant<KIND, GCT_STORE_F16>(a, s);
        break;
    case GCT_STORE_BF16:
        if constexpr (KIND <= GCT_FAN_LINE_1D) launch_bp2d<KIND, GCT_STORE_BF16>(a, s);
        else launch_bp3d_variant<KIND, GCT_STORE_BF16>(a, s);
        break;
    default:
        if constexpr (KIND <= GCT_FAN_LINE_1D) launch_bp2d<KIND, GCT_STORE_F32>(a, s);
        else launch_bp3d_variant<KIND, GCT_STORE_F32>(a, s);
        break;
    }
}

// Persistent batch kernel of one geometry kind (all storage types).
template <int KIND>
void batch_group_kind(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    switch (g.storage) {
    case GCT_STORE_F16: launch_batch_kind<KIND, GCT_STORE_F16>(scans, g, counter, grid, s); break;
    case GCT_STORE_BF16: launch_batch_kind<KIND, GCT_STORE_BF16>(scans, g, counter, grid, s); break;
    default: launch_batch_kind<KIND, GCT_STORE_F32>(scans, g, counter, grid, s); break;
    }
}

}  // namespace
}  // namespace gct
