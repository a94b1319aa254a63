// HIP kernels of the FDK path for gfx950 (CDNA4): fused pre-weighting and
// ramp filtering (in-LDS Stockham FFT), voxel-driven backprojection, the
// Joseph forward projector and the batched persistent backprojection.
//
// Reference stages replaced (satidev/gct):
//   cosWeightKernel* (cos_weighting.cuh:155-263), parkerWeightKern and
//   parkerWeighting*Kernel (parker_weighting.cu:5-52), the cuFFT ramp filter
//   (ramp_filtering_fft.cuh:157-219, ramp_filtering_fft.cu:14-96),
//   backProjKern (back_proj.cuh:16-104) with its texture copies
//   (copy_host_dev.cuh:99-147), setVal/operator*= on the volume
//   (nd_array_dev.cuh:58-99).
#include "common.hpp"

#include <algorithm>
#include <cstdint>

namespace gct {

namespace {

constexpr float kPiRef = 3.141592654f;   // the reference's pi(), constants.cuh:6-9

// ---- storage of filtered projections -------------------------------------

template <int S> struct Stor;
template <> struct Stor<GCT_STORE_F32> {
    using T = float;
    __device__ static float ld(const T *p, uint32_t i) { return p[i]; }
    __device__ static T cv(float x) { return x; }
};
template <> struct Stor<GCT_STORE_F16> {
    using T = _Float16;
    __device__ static float ld(const T *p, uint32_t i) { return static_cast<float>(p[i]); }
    __device__ static T cv(float x) { return static_cast<_Float16>(x); }
};
template <> struct Stor<GCT_STORE_BF16> {
    using T = uint16_t;
    __device__ static float ld(const T *p, uint32_t i) { return __uint_as_float(uint32_t(p[i]) << 16); }
    __device__ static T cv(float x)
    {
        uint32_t u = __float_as_uint(x);
        u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
        return static_cast<uint16_t>(u >> 16);
    }
};

// ---- redundancy (Parker / parallel) weights --------------------------------

// parkerWeightFan, parker_weighting.cuh:80-105
__device__ float parker_fan(float gam, float beta, float range)
{
    const float t2 = 2.0f * gam + range - kPiRef;
    const float t3 = 2.0f * gam + kPiRef;
    if (beta >= 0.0f && beta <= t2) {
        float s = sinf(kPiRef * beta / (2.0f * (range - kPiRef) + 4.0f * gam));
        return s * s;
    }
    if (beta >= t2 && beta <= t3) return 1.0f;
    if (beta >= t3 && beta <= range) {
        float s = sinf(kPiRef * (range - beta) / (2.0f * (range - kPiRef) - 4.0f * gam));
        return s * s;
    }
    return 0.0f;
}

__global__ void redundancy_kernel(const float *fan, const float *beta, uint32_t nu, uint32_t nviews, float range,
                                  int mode, float value, float *red)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nu * nviews) return;
    const uint32_t v = i / nu, c = i - v * nu;
    red[i] = (mode == 0) ? value : parker_fan(fan[c], beta[v], range);
}

// ---- in-LDS Stockham FFT -------------------------------------------------
//
// Radix-R Stockham autosort step (Bainville's formulation): with T = L/R
// butterflies and p the product of the radices already applied, butterfly i
// reads x[i + j T], multiplies by exp(-2 pi i j k / (R p)), k = i mod p,
// applies a length-R DFT and writes y[(i - k) R + k + j p].  Each thread keeps
// its butterflies in registers across one barrier, so a single LDS buffer of
// L complex values suffices.

__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <int R> __device__ inline void dft(float2 *u);

template <> __device__ inline void dft<2>(float2 *u)
{
    float2 a = u[0], b = u[1];
    u[0] = make_float2(a.x + b.x, a.y + b.y);
    u[1] = make_float2(a.x - b.x, a.y - b.y);
}

template <> __device__ inline void dft<4>(float2 *u)
{
    float2 a = make_float2(u[0].x + u[2].x, u[0].y + u[2].y);
    float2 b = make_float2(u[0].x - u[2].x, u[0].y - u[2].y);
    float2 c = make_float2(u[1].x + u[3].x, u[1].y + u[3].y);
    float2 d = make_float2(u[1].x - u[3].x, u[1].y - u[3].y);
    u[0] = make_float2(a.x + c.x, a.y + c.y);
    u[2] = make_float2(a.x - c.x, a.y - c.y);
    u[1] = make_float2(b.x + d.y, b.y - d.x);   // b - i d
    u[3] = make_float2(b.x - d.y, b.y + d.x);   // b + i d
}

template <> __device__ inline void dft<8>(float2 *u)
{
    float2 e[4] = {u[0], u[2], u[4], u[6]};
    float2 o[4] = {u[1], u[3], u[5], u[7]};
    dft<4>(e);
    dft<4>(o);
    const float r = 0.70710678118654752f;
    // o[m] *= W8^m, W8 = exp(-i pi / 4)
    float2 o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
    float2 o2 = make_float2(o[2].y, -o[2].x);
    float2 o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    u[0] = make_float2(e[0].x + o[0].x, e[0].y + o[0].y);
    u[4] = make_float2(e[0].x - o[0].x, e[0].y - o[0].y);
    u[1] = make_float2(e[1].x + o1.x, e[1].y + o1.y);
    u[5] = make_float2(e[1].x - o1.x, e[1].y - o1.y);
    u[2] = make_float2(e[2].x + o2.x, e[2].y + o2.y);
    u[6] = make_float2(e[2].x - o2.x, e[2].y - o2.y);
    u[3] = make_float2(e[3].x + o3.x, e[3].y + o3.y);
    u[7] = make_float2(e[3].x - o3.x, e[3].y - o3.y);
}

template <uint32_t L, uint32_t R, uint32_t P, uint32_t NT>
__device__ inline void stockham_pass(float2 *buf, const float2 *__restrict__ tw, uint32_t tid)
{
    constexpr uint32_t T = L / R;
    constexpr uint32_t B = (T + NT - 1) / NT;
    float2 u[B][R];
#pragma unroll
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t i = tid + b * NT;
        if (T % NT == 0 || i < T) {
#pragma unroll
            for (uint32_t j = 0; j < R; ++j) u[b][j] = buf[i + j * T];
        }
    }
    __syncthreads();
#pragma unroll
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t i = tid + b * NT;
        if (T % NT == 0 || i < T) {
            const uint32_t k = i & (P - 1u);
            if constexpr (P > 1) {
                constexpr uint32_t stride = L / (R * P);
#pragma unroll
                for (uint32_t j = 1; j < R; ++j) u[b][j] = cmul(u[b][j], tw[j * k * stride]);
            }
            dft<R>(u[b]);
            const uint32_t base = (i - k) * R + k;
#pragma unroll
            for (uint32_t j = 0; j < R; ++j) buf[base + j * P] = u[b][j];
        }
    }
    __syncthreads();
}

template <uint32_t L, uint32_t P, uint32_t NT>
__device__ inline void fft_from(float2 *buf, const float2 *__restrict__ tw, uint32_t tid)
{
    if constexpr (P < L) {
        constexpr uint32_t rem = L / P;
        constexpr uint32_t R = (rem == 16u) ? 4u : (rem >= 8u ? 8u : rem);
        stockham_pass<L, R, P, NT>(buf, tw, tid);
        fft_from<L, P * R, NT>(buf, tw, tid);
    }
}

// ---- fused pre-weighting + ramp filter ------------------------------------

template <int KIND>
__device__ inline float cos_weight(const FiltRec &f, float du, float dv, uint32_t c, uint32_t r)
{
    // cos_weighting.cuh:16-153; PCSToDCS = (pix - centre) * width
    const float u = (static_cast<float>(c) - f.cu) * du;
    const float v = (static_cast<float>(r) - f.cv) * dv;
    if constexpr (KIND == GCT_FAN_ARC_1D) return cosf(u);
    if constexpr (KIND == GCT_FAN_LINE_1D) return f.sdd / sqrtf(u * u + f.sdd * f.sdd);
    if constexpr (KIND == GCT_CONE_PARALLEL) return f.sdd / sqrtf(v * v + f.sdd * f.sdd);
    if constexpr (KIND == GCT_CYLINDRICAL) return cosf(u) * (f.sdd / sqrtf(v * v + f.sdd * f.sdd));
    if constexpr (KIND == GCT_FLAT) return f.sdd / sqrtf(u * u + v * v + f.sdd * f.sdd);
    return 1.0f;
}

constexpr uint32_t kFiltThreads = 64;

// One wavefront filters two detector rows as one complex signal z = a + i b:
// IFFT(FFT(z) H) = (a * h) + i (b * h) because the kernel h is real.  H is the
// FFT of the N-tap kernel circularly shifted by floor(N/2) and scaled by 1/L,
// so output sample c is the reference's y[c + floor(N/2)] / L
// (ramp_filtering_fft.cu:45-62).  The inverse FFT is conj(FFT(conj(.))).
template <int KIND, int S, uint32_t L>
__global__ __launch_bounds__(kFiltThreads) void filter_kernel(FilterLaunch a)
{
    __shared__ float2 buf[L];
    using St = Stor<S>;
    const uint32_t tid = threadIdx.x;
    const uint32_t nv = a.nv, nu = a.nu;
    const uint32_t nrows = a.num_views * nv;
    const uint32_t r0 = 2u * blockIdx.x, r1 = r0 + 1u;
    const bool has1 = r1 < nrows;
    const uint32_t v0 = a.first_view + r0 / nv, j0 = r0 % nv;
    const uint32_t v1 = a.first_view + (has1 ? r1 : r0) / nv, j1 = (has1 ? r1 : r0) % nv;
    const float *pa = a.proj + (static_cast<size_t>(v0 - a.proj_view_offset) * nv + j0) * nu;
    const float *pb = a.proj + (static_cast<size_t>(v1 - a.proj_view_offset) * nv + j1) * nu;
    const FiltRec f0 = a.frec[v0], f1 = a.frec[v1];
    const float *red0 = a.red + static_cast<size_t>(v0) * nu;
    const float *red1 = a.red + static_cast<size_t>(v1) * nu;

    for (uint32_t c = tid; c < L; c += kFiltThreads) {
        float x = 0.0f, y = 0.0f;
        if (c < nu) {
            // fdk_axial.cuh:171-173: p *= cosine weight, then p *= redundancy weight
            x = pa[c] * cos_weight<KIND>(f0, a.du, a.dv, c, j0);
            x = x * red0[c];
            if (has1) {
                y = pb[c] * cos_weight<KIND>(f1, a.du, a.dv, c, j1);
                y = y * red1[c];
            }
        }
        buf[c] = make_float2(x, y);
    }
    __syncthreads();
    fft_from<L, 1, kFiltThreads>(buf, a.tw, tid);
    for (uint32_t c = tid; c < L; c += kFiltThreads) {
        const float2 z = cmul(buf[c], a.H[c]);
        buf[c] = make_float2(z.x, -z.y);
    }
    __syncthreads();
    fft_from<L, 1, kFiltThreads>(buf, a.tw, tid);

    typename St::T *out = static_cast<typename St::T *>(a.filt);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    typename St::T *oa = out + static_cast<size_t>(v0) * view_elems + static_cast<size_t>(j0 + 1u) * a.pitch + 1u;
    typename St::T *ob = out + static_cast<size_t>(v1) * view_elems + static_cast<size_t>(j1 + 1u) * a.pitch + 1u;
    for (uint32_t c = tid; c < nu; c += kFiltThreads) {
        const float2 z = buf[c];
        oa[c] = St::cv(z.x);
        if (has1) ob[c] = St::cv(-z.y);
    }
}

// ---- backprojection -------------------------------------------------------

struct Sample {
    uint32_t colbase;   // padded column index of the left tap
    float alpha;        // horizontal fraction
};

// Clamping the coordinate to [-1, n] keeps every tap inside the padded
// buffer; any coordinate outside the detector then lands on zero padding,
// which reproduces the border-zero texture (nd_array_texture.cuh:43-50).
__device__ inline Sample col_sample(float col, float nu)
{
    const float c = __builtin_amdgcn_fmed3f(col, -1.0f, nu);
    const float f = floorf(c);
    return Sample{static_cast<uint32_t>(static_cast<int>(f) + 1), c - f};
}

template <int S>
__device__ inline float bilinear(const typename Stor<S>::T *__restrict__ p, uint32_t pitch, Sample cs, float row,
                                 float nvf)
{
    const float r = __builtin_amdgcn_fmed3f(row, -1.0f, nvf);
    const float fr = floorf(r);
    const float beta = r - fr;
    const uint32_t off = static_cast<uint32_t>(static_cast<int>(fr) + 1) * pitch + cs.colbase;
    const float t0 = Stor<S>::ld(p, off), t1 = Stor<S>::ld(p, off + 1u);
    const float b0 = Stor<S>::ld(p, off + pitch), b1 = Stor<S>::ld(p, off + pitch + 1u);
    const float top = fmaf(cs.alpha, t1 - t0, t0);
    const float bot = fmaf(cs.alpha, b1 - b0, b0);
    return fmaf(beta, bot - top, top);
}

template <int S>
__device__ inline float linear(const typename Stor<S>::T *__restrict__ p, uint32_t pitch, Sample cs)
{
    const uint32_t off = pitch + cs.colbase;
    const float t0 = Stor<S>::ld(p, off), t1 = Stor<S>::ld(p, off + 1u);
    return fmaf(cs.alpha, t1 - t0, t0);
}

// Per (thread, view) setup of a voxel column, shared by all z voxels.
struct ColSetup {
    float col;     // column coordinate (unclamped)
    float w;       // weight (constant along z unless cylindrical)
    float row0;    // row coordinate of the first z voxel
    float drow;    // row step per z voxel
    float rho2;    // cylindrical: in-plane squared distance
    float nv0;     // cylindrical: v numerator of the first z voxel
    float dnv;     // cylindrical: v numerator step
};

template <int KIND, bool F64>
__device__ inline ColSetup setup_column(const BpLaunch &a, uint32_t v, float fx, float fy, float fz, float nuf)
{
    ColSetup s{};
    if constexpr (KIND == GCT_PARALLEL_2D || KIND == GCT_CONE_PARALLEL) {
        // The column is an affine form with a large constant (the detector
        // centre); evaluating it in FP64 once per column keeps it as accurate
        // as the reference's FP32 chain t = x cos + y sin, col = t / du + cu.
        const BpRec64 &r = a.rec64[v];
        s.col = static_cast<float>(r.u[0] + r.u[1] * static_cast<double>(fx) + r.u[2] * static_cast<double>(fy));
        s.w = 1.0f;
        s.row0 = static_cast<float>(r.v[0] + r.v[3] * static_cast<double>(fz));
        s.drow = static_cast<float>(r.v[3]);
    } else if constexpr (KIND == GCT_FLAT) {
        if constexpr (F64) {
            const BpRec64 &r = a.rec64[v];
            const double X = fx, Y = fy, Z = fz;
            const double nw = r.w[0] + r.w[1] * X + r.w[2] * Y + r.w[3] * Z;
            const double nu = r.u[0] + r.u[1] * X + r.u[2] * Y + r.u[3] * Z;
            const double nv = r.v[0] + r.v[1] * X + r.v[2] * Y + r.v[3] * Z;
            const double rd = 1.0 / nw;
            s.col = static_cast<float>(r.k[0] + nu * rd);
            s.w = static_cast<float>(rd * rd);
            s.row0 = static_cast<float>(r.k[1] + nv * rd);
            s.drow = static_cast<float>(r.v[3] * rd);
        } else {
            const BpRec &r = a.rec[v];
            const float nw = fmaf(r.w[3], fz, fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0])));
            const float nu = fmaf(r.u[3], fz, fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0])));
            const float nv = fmaf(r.v[3], fz, fmaf(r.v[2], fy, fmaf(r.v[1], fx, r.v[0])));
            const float rd = 1.0f / nw;
            s.col = fmaf(nu, rd, r.k[0]);
            s.w = rd * rd;
            s.row0 = fmaf(nv, rd, r.k[1]);
            s.drow = r.v[3] * rd;
        }
    } else if constexpr (KIND == GCT_CYLINDRICAL) {
        if constexpr (F64) {
            const BpRec64 &r = a.rec64[v];
            const double X = fx, Y = fy, Z = fz;
            const double nw = r.w[0] + r.w[1] * X + r.w[2] * Y + r.w[3] * Z;
            const double nu = r.u[0] + r.u[1] * X + r.u[2] * Y + r.u[3] * Z;
            const double nv = r.v[0] + r.v[1] * X + r.v[2] * Y + r.v[3] * Z;
            const double rho2 = nu * nu + nw * nw;
            const double rr = 1.0 / sqrt(rho2);
            s.col = static_cast<float>(r.k[0] + atan2(nu, nw) * static_cast<double>(a.inv_dgamma));
            s.row0 = static_cast<float>(r.k[1] + r.k[2] * nv * rr);
            s.drow = static_cast<float>(r.k[2] * r.v[3] * rr);
            s.rho2 = static_cast<float>(rho2);
            s.nv0 = static_cast<float>(nv);
            s.dnv = static_cast<float>(r.v[3]);
            s.w = static_cast<float>(r.k[3]);
        } else {
            const BpRec &r = a.rec[v];
            const float nw = fmaf(r.w[3], fz, fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0])));
            const float nu = fmaf(r.u[3], fz, fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0])));
            const float nv = fmaf(r.v[3], fz, fmaf(r.v[2], fy, fmaf(r.v[1], fx, r.v[0])));
            const float rho2 = fmaf(nu, nu, nw * nw);
            const float rr = 1.0f / sqrtf(rho2);
            s.col = fmaf(atan2f(nu, nw), a.inv_dgamma, r.k[0]);
            s.row0 = fmaf(r.k[2] * nv, rr, r.k[1]);
            s.drow = r.k[2] * r.v[3] * rr;
            s.rho2 = rho2;
            s.nv0 = nv;
            s.dnv = r.v[3];
            s.w = r.k[3];   // SID; divided by |r|^2 per voxel
        }
    }
    return s;
}

// Whether the row axis of a view stays parallel to z: then u, w and the
// weight are constant along a voxel column and only the row coordinate moves.
template <int KIND, bool CIRC>
constexpr bool column_invariant()
{
    return CIRC || KIND == GCT_PARALLEL_2D || KIND == GCT_CONE_PARALLEL;
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
    const uint32_t iz0 = bz * ZU;
    if (ix >= a.nx || iy >= a.ny) return;
    const uint32_t zn = min(static_cast<uint32_t>(ZU), a.nz - iz0);
    const size_t slice = static_cast<size_t>(a.nx) * a.ny;
    float *vp = a.vol + static_cast<size_t>(iz0) * slice + static_cast<size_t>(iy) * a.nx + ix;

    float acc[ZU];
#pragma unroll
    for (int k = 0; k < ZU; ++k) acc[k] = (a.accumulate && static_cast<uint32_t>(k) < zn) ? vp[k * slice] : 0.0f;

    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy), fz = static_cast<float>(iz0);
    const float nuf = static_cast<float>(a.nu), nvf = static_cast<float>(a.nv);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vend = a.first_view + a.num_views;
    for (uint32_t v = a.first_view; v < vend; ++v) {
        const T *p = base + static_cast<size_t>(v) * view_elems;
        if constexpr (column_invariant<KIND, CIRC>()) {
            // set up at z index 0 and step by the absolute index, so the
            // result does not depend on how z is split into tiles
            const ColSetup s = setup_column<KIND, F64>(a, v, fx, fy, 0.0f, nuf);
            const Sample cs = col_sample(s.col, nuf);
#pragma unroll
            for (int k = 0; k < ZU; ++k) {
                const float zk = fz + static_cast<float>(k);
                const float row = fmaf(zk, s.drow, s.row0);
                const float val = bilinear<S>(p, a.pitch, cs, row, nvf);
                float w = s.w;
                if constexpr (KIND == GCT_CYLINDRICAL) {
                    const float nvk = fmaf(zk, s.dnv, s.nv0);
                    w = s.w * __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                }
                acc[k] = fmaf(val, w, acc[k]);
            }
        } else {
#pragma unroll
            for (int k = 0; k < ZU; ++k) {
                const ColSetup s = setup_column<KIND, F64>(a, v, fx, fy, fz + static_cast<float>(k), nuf);
                const float val = bilinear<S>(p, a.pitch, col_sample(s.col, nuf), s.row0, nvf);
                float w = s.w;
                if constexpr (KIND == GCT_CYLINDRICAL) w = s.w * __builtin_amdgcn_rcpf(fmaf(s.nv0, s.nv0, s.rho2));
                acc[k] = fmaf(val, w, acc[k]);
            }
        }
    }
    const float sc = a.scale;
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        if (static_cast<uint32_t>(k) < zn) vp[k * slice] = a.finish ? acc[k] * sc : acc[k];
}

// One tile of a 2-D image (1-D detector): pixel (ix, iy) per thread.
template <int KIND, int S>
__device__ inline void bp_tile2d(const BpLaunch &a, uint32_t ix, uint32_t iy)
{
    using T = typename Stor<S>::T;
    if (ix >= a.nx || iy >= a.ny) return;
    float *vp = a.vol + static_cast<size_t>(iy) * a.nx + ix;
    float acc = a.accumulate ? *vp : 0.0f;
    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy);
    const float nuf = static_cast<float>(a.nu);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vend = a.first_view + a.num_views;
    for (uint32_t v = a.first_view; v < vend; ++v) {
        const BpRec &r = a.rec[v];
        const T *p = base + static_cast<size_t>(v) * view_elems;
        float col, w;
        if constexpr (KIND == GCT_PARALLEL_1D) {
            col = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
            w = 1.0f;
        } else if constexpr (KIND == GCT_FAN_LINE_1D) {
            const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
            const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
            const float rd = 1.0f / nw;
            col = fmaf(nu, rd, r.k[0]);
            w = rd * rd;
        } else {   // GCT_FAN_ARC_1D
            const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
            const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
            col = fmaf(atan2f(nu, nw), a.inv_dgamma, r.k[0]);
            w = r.k[3] / fmaf(nu, nu, nw * nw);
        }
        const float val = linear<S>(p, a.pitch, col_sample(col, nuf));
        acc = fmaf(val, w, acc);
    }
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

template <int KIND, int S>
__global__ __launch_bounds__(64) void bp2d_kernel(BpLaunch a)
{
    const uint32_t nbx = (a.nx + 7u) / 8u;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, by = t / nbx;
    bp_tile2d<KIND, S>(a, bx * 8u + (threadIdx.x & 7u), by * 8u + (threadIdx.x >> 3));
}

// ---- LDS-staged backprojection ----------------------------------------------
//
// A workgroup owns a 16 x 16 x ZU voxel tile.  For each block of VB views it
// projects the tile's 8 corner voxels, takes the bounding box of the detector
// pixels they reach (plus one pixel of margin), and stages that box from HBM
// into LDS as pairs (p[r][c], p[r][c+1] - p[r][c]) -- value and horizontal
// difference, FP16/BF16 widened to FP32 -- so each thread's bilinear tap is
// two 64-bit LDS reads and one FMA per row.  The differences are the same
// FP32 subtractions the direct kernel performs per update, so both kernels
// give the same bits.  Coordinates are clamped to the staged box: the box never
// extends past the zero padding of the filtered-projection buffer, so the
// clamp also reproduces the border-zero texture, and it keeps every LDS read
// inside the box.  A view whose box does not fit is read from HBM directly.

constexpr int kTP = 48;   // staged (value, difference) pairs per row

struct Box {
    int clo, rlo, ncols, nrows;   // ncols == 0: the view is read from HBM
};

template <int KIND, int ZU>
__device__ inline void tile_boxes(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t vb0,
                                  uint32_t nvb, int TR, Box *boxes, uint32_t tid)
{
    if (tid >= 8u * nvb) return;
    const uint32_t j = tid >> 3, corner = tid & 7u;
    const uint32_t cx = min(bx * 16u + ((corner & 1u) ? 15u : 0u), a.nx - 1u);
    const uint32_t cy = min(by * 16u + ((corner & 2u) ? 15u : 0u), a.ny - 1u);
    const uint32_t cz = min(bz * ZU + ((corner & 4u) ? static_cast<uint32_t>(ZU - 1) : 0u), a.nz - 1u);
    const float nuf = static_cast<float>(a.nu);
    const ColSetup s = setup_column<KIND, false>(a, vb0 + j, static_cast<float>(cx), static_cast<float>(cy), 0.0f, nuf);
    const float row = fmaf(static_cast<float>(cz), s.drow, s.row0);
    float cmin = s.col, cmax = s.col, rmin = row, rmax = row;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
        cmin = fminf(cmin, __shfl_xor(cmin, off));
        cmax = fmaxf(cmax, __shfl_xor(cmax, off));
        rmin = fminf(rmin, __shfl_xor(rmin, off));
        rmax = fmaxf(rmax, __shfl_xor(rmax, off));
    }
    if (corner != 0u) return;
    Box b{0, 0, 0, 0};
    const bool finite = cmin > -1.0e6f && cmax < 1.0e6f && rmin > -1.0e6f && rmax < 1.0e6f;
    if (finite) {
        const int clo = max(static_cast<int>(floorf(cmin)) - 1, -1);
        const int chi = min(static_cast<int>(floorf(cmax)) + 2, static_cast<int>(a.nu) + 1);
        const int rlo = max(static_cast<int>(floorf(rmin)) - 1, -1);
        const int rhi = min(static_cast<int>(floorf(rmax)) + 2, static_cast<int>(a.nv) + 1);
        const int nc = chi - clo + 1, nr = rhi - rlo + 1;
        if (nc >= 2 && nr >= 2 && nc - 1 <= kTP && nr <= TR) b = Box{clo, rlo, nc, nr};
    }
    boxes[j] = b;
}

// One 16 x 16 x ZU tile (bx, by, bz) of a circular 3-D scan; every thread of
// the 256-thread workgroup calls it (it holds workgroup barriers).  `tile`
// and `boxes` are the workgroup's LDS staging area of VB * TR * kTP pairs and
// VB boxes.
template <int KIND, int S, bool F64, int ZU, int TR, int VB>
__device__ void bp_tile3d_lds(const BpLaunch &a, uint32_t bx, uint32_t by, uint32_t bz, uint32_t tid, float2 *tile,
                              Box *boxes)
{
    using T = typename Stor<S>::T;
    const uint32_t ix = bx * 16u + (tid & 15u), iy = by * 16u + (tid >> 4), iz0 = bz * ZU;
    const bool active = ix < a.nx && iy < a.ny;
    const uint32_t zn = min(static_cast<uint32_t>(ZU), a.nz - iz0);
    const size_t slice = static_cast<size_t>(a.nx) * a.ny;
    float *vp = a.vol + static_cast<size_t>(iz0) * slice + static_cast<size_t>(active ? iy : 0u) * a.nx + (active ? ix : 0u);

    float acc[ZU];
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        acc[k] = (active && a.accumulate && static_cast<uint32_t>(k) < zn) ? vp[k * slice] : 0.0f;

    const float fx = static_cast<float>(ix), fy = static_cast<float>(iy), fz = static_cast<float>(iz0);
    const float nuf = static_cast<float>(a.nu), nvf = static_cast<float>(a.nv);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    const T *base = static_cast<const T *>(a.filt);
    const uint32_t vend = a.first_view + a.num_views;

    for (uint32_t vb0 = a.first_view; vb0 < vend; vb0 += VB) {
        const uint32_t nvb = min(static_cast<uint32_t>(VB), vend - vb0);
        tile_boxes<KIND, ZU>(a, bx, by, bz, vb0, nvb, TR, boxes, tid);
        __syncthreads();
        for (uint32_t j = 0; j < nvb; ++j) {
            const Box b = boxes[j];
            if (b.ncols == 0) continue;
            const T *p = base + static_cast<size_t>(vb0 + j) * view_elems;
            float2 *tl = tile + j * (TR * kTP);
            const int m = b.ncols - 1;            // staged pairs per row
            const bool narrow = m <= 32;          // two rows per wave
            const int c = narrow ? static_cast<int>(tid & 31u) : static_cast<int>(tid & 63u);
            const int rstep = narrow ? 8 : 4;
            if (c < m) {
                for (int r = narrow ? static_cast<int>(tid >> 5) : static_cast<int>(tid >> 6); r < b.nrows; r += rstep) {
                    const uint32_t g = static_cast<uint32_t>(b.rlo + r + 1) * a.pitch + static_cast<uint32_t>(b.clo + c + 1);
                    const float v0 = Stor<S>::ld(p, g), v1 = Stor<S>::ld(p, g + 1u);
                    tl[r * kTP + c] = make_float2(v0, v1 - v0);
                }
            }
        }
        __syncthreads();
        if (active) {
            for (uint32_t j = 0; j < nvb; ++j) {
                const uint32_t v = vb0 + j;
                const ColSetup s = setup_column<KIND, F64>(a, v, fx, fy, 0.0f, nuf);
                const Box b = boxes[j];
                if (b.ncols != 0) {
                    const float cl = __builtin_amdgcn_fmed3f(s.col, static_cast<float>(b.clo),
                                                             static_cast<float>(b.clo + b.ncols - 2));
                    const float cf = floorf(cl);
                    const float alpha = cl - cf;
                    // Pair (r, c) of the box sits at byte 8 (j TR kTP + (r - rlo) kTP + (c - clo)).
                    // The per-view terms fold into one base; the row term is added in float,
                    // which is exact because every offset is an integer below 2^24.
                    // (signed: the base is negative when the box starts below row 0 of view j's slot)
                    const int ebase = static_cast<int>(j) * (TR * kTP) + (static_cast<int>(cf) - b.clo) - b.rlo * kTP;
                    const float obase = static_cast<float>(8 * ebase);
                    const float rlo = static_cast<float>(b.rlo), rhi = static_cast<float>(b.rlo + b.nrows - 2);
#pragma unroll
                    for (int k = 0; k < ZU; ++k) {
                        const float zk = fz + static_cast<float>(k);
                        const float row = __builtin_amdgcn_fmed3f(fmaf(zk, s.drow, s.row0), rlo, rhi);
                        const float rf = floorf(row);
                        const float beta = row - rf;
                        const uint32_t off = static_cast<uint32_t>(fmaf(rf, static_cast<float>(8 * kTP), obase));
                        const float2 *q = reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(tile) + off);
                        const float2 tp = q[0], bt = q[kTP];
                        const float top = fmaf(alpha, tp.y, tp.x);
                        const float bot = fmaf(alpha, bt.y, bt.x);
                        const float val = fmaf(beta, bot - top, top);
                        float w = s.w;
                        if constexpr (KIND == GCT_CYLINDRICAL) {
                            const float nvk = fmaf(zk, s.dnv, s.nv0);
                            w = s.w * __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                        }
                        acc[k] = fmaf(val, w, acc[k]);
                    }
                } else {
                    const T *p = base + static_cast<size_t>(v) * view_elems;
                    const Sample cs = col_sample(s.col, nuf);
#pragma unroll
                    for (int k = 0; k < ZU; ++k) {
                        const float zk = fz + static_cast<float>(k);
                        const float row = fmaf(zk, s.drow, s.row0);
                        const float val = bilinear<S>(p, a.pitch, cs, row, nvf);
                        float w = s.w;
                        if constexpr (KIND == GCT_CYLINDRICAL) {
                            const float nvk = fmaf(zk, s.dnv, s.nv0);
                            w = s.w * __builtin_amdgcn_rcpf(fmaf(nvk, nvk, s.rho2));
                        }
                        acc[k] = fmaf(val, w, acc[k]);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (!active) return;
    const float sc = a.scale;
#pragma unroll
    for (int k = 0; k < ZU; ++k)
        if (static_cast<uint32_t>(k) < zn) vp[k * slice] = a.finish ? acc[k] * sc : acc[k];
}

template <int KIND, int S, bool F64, int ZU, int TR, int VB>
__global__ __launch_bounds__(256) void bp3d_lds_kernel(BpLaunch a)
{
    __shared__ float2 tile[VB * TR * kTP];
    __shared__ Box boxes[VB];
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t bx = t % nbx, rest = t / nbx;
    bp_tile3d_lds<KIND, S, F64, ZU, TR, VB>(a, bx, rest % nby, rest / nby, threadIdx.x, tile, boxes);
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
    const float nuf = static_cast<float>(a.nu);
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
            for (uint32_t j = 0; j < nvb; ++j) {
                const BpRec &r = a.rec[vb0 + j];
                float col, w;
                if constexpr (KIND == GCT_PARALLEL_1D) {
                    col = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
                    w = 1.0f;
                } else if constexpr (KIND == GCT_FAN_LINE_1D) {
                    const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
                    const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
                    const float rd = 1.0f / nw;
                    col = fmaf(nu, rd, r.k[0]);
                    w = rd * rd;
                } else {
                    const float nw = fmaf(r.w[2], fy, fmaf(r.w[1], fx, r.w[0]));
                    const float nu = fmaf(r.u[2], fy, fmaf(r.u[1], fx, r.u[0]));
                    col = fmaf(atan2f(nu, nw), a.inv_dgamma, r.k[0]);
                    w = r.k[3] / fmaf(nu, nu, nw * nw);
                }
                const Sample cs = col_sample(col, nuf);
                const float *rr = rows + j * a.pitch + cs.colbase;
                const float t0 = rr[0], t1 = rr[1];
                acc = fmaf(fmaf(cs.alpha, t1 - t0, t0), w, acc);
            }
        }
        __syncthreads();
    }
    if (active) *vp = a.finish ? acc * a.scale : acc;
}

// ---- persistent batched backprojection -------------------------------------

constexpr int kBatchZU = static_cast<int>(kBatchDirectZU);   // direct tiles (general trajectories)
constexpr int kBatchTR = 40, kBatchVB = 3;                    // LDS tiles (circular), depth kBatchLdsZU

// Persistent backprojection over the scans of one group (same geometry kind,
// storage and tile type; a batch launches one such kernel per group).  Work
// items are (scan, tile) pairs numbered scan by scan; each workgroup takes
// items from the group's device counter until none are left, so the grid
// drains on its own whatever the mix of scan sizes.
template <int KIND, int S, bool LDS>
__global__ __launch_bounds__(256) void bp_batch_kernel(const BatchScan *__restrict__ scans, uint32_t nscans,
                                                       uint32_t total, uint32_t *counter)
{
    constexpr bool k1d = KIND <= GCT_FAN_LINE_1D;
    __shared__ uint32_t item;
    __shared__ float2 tile[LDS ? kBatchVB * kBatchTR * kTP : 1];
    __shared__ Box boxes[kBatchVB];
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
                bp_tile3d_lds<KIND, S, false, static_cast<int>(kBatchLdsZU), kBatchTR, kBatchVB>(a, bx, by, bz, tid,
                                                                                                 tile, boxes);
            else
                bp_tile3d<KIND, S, false, false, kBatchZU>(a, bx, by, bz, tid);
        }
    }
}

__global__ void zero_counters_kernel(uint32_t *c, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) c[i] = 0u;
}

// ---- packing between [V][Nv][Nu] and the padded layout ---------------------

template <int S>
__global__ void pack_kernel(const float *src, void *dst, uint32_t nu, uint32_t nv, uint32_t nviews, uint32_t pitch,
                            uint32_t rows_pv)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t n = static_cast<size_t>(nu) * nv * nviews;
    if (i >= n) return;
    const uint32_t c = static_cast<uint32_t>(i % nu);
    const size_t rr = i / nu;
    const uint32_t r = static_cast<uint32_t>(rr % nv);
    const size_t v = rr / nv;
    typename Stor<S>::T *o = static_cast<typename Stor<S>::T *>(dst);
    o[(v * rows_pv + r + 1u) * pitch + c + 1u] = Stor<S>::cv(src[i]);
}

template <int S>
__global__ void unpack_kernel(const void *src, float *dst, uint32_t nu, uint32_t nv, uint32_t nviews, uint32_t pitch,
                              uint32_t rows_pv)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t n = static_cast<size_t>(nu) * nv * nviews;
    if (i >= n) return;
    const uint32_t c = static_cast<uint32_t>(i % nu);
    const size_t rr = i / nu;
    const uint32_t r = static_cast<uint32_t>(rr % nv);
    const size_t v = rr / nv;
    const typename Stor<S>::T *s = static_cast<const typename Stor<S>::T *>(src);
    dst[i] = Stor<S>::ld(s, static_cast<uint32_t>((v * rows_pv + r + 1u) * pitch + c + 1u));
}

// ---- Joseph forward projection ----------------------------------------------

// Line integral of the image along {o + t d} by Joseph's method: step through
// the planes of voxel centres normal to the axis the ray is most aligned with
// (in units of voxels), interpolate bilinearly in each plane (zero outside),
// and scale by the path length per plane.  Rays from a source count t > 0.
// Sum over the planes of voxel centres normal to the ray's dominant axis of
// the bilinear sample in each plane, times the step length.  Plane i is
// crossed at t(i) = t0 + i dt; the two in-plane coordinates (in voxel units)
// are linear in i, so after an FP64 set-up per ray each plane costs two FMAs,
// two floors and four predicated taps (zero outside the volume).
__device__ float joseph_ray(const FpLaunch &a, const float o[3], const float d[3], bool from_source)
{
    const bool is3 = !is_1d(a.kind);
    const float first[3] = {a.x0, a.y0, a.z0};
    const float sp[3] = {a.dx, a.dy, a.dz};
    const int n[3] = {static_cast<int>(a.nx), static_cast<int>(a.ny), static_cast<int>(is3 ? a.nz : 1u)};
    const int dims = is3 ? 3 : 2;
    int ax = 0;
    float best = -1.0f;
    for (int i = 0; i < dims; ++i) {
        const float q = fabsf(d[i]) / sp[i];
        if (q > best) { best = q; ax = i; }
    }
    if (!(fabsf(d[ax]) > 0.0f)) return 0.0f;
    // clip the plane range to where the ray is inside the (one-voxel grown) box
    float tlo = -3.0e38f, thi = 3.0e38f;
    for (int i = 0; i < dims; ++i) {
        const float lo = first[i] - sp[i], hi = first[i] + static_cast<float>(n[i]) * sp[i];
        if (fabsf(d[i]) > 0.0f) {
            float t0 = (lo - o[i]) / d[i], t1 = (hi - o[i]) / d[i];
            if (t0 > t1) { const float tt = t0; t0 = t1; t1 = tt; }
            tlo = fmaxf(tlo, t0);
            thi = fminf(thi, t1);
        } else if (o[i] < lo || o[i] > hi) {
            return 0.0f;
        }
    }
    if (from_source) tlo = fmaxf(tlo, 0.0f);
    if (!(thi > tlo)) return 0.0f;
    const float pa0 = o[ax] + tlo * d[ax], pa1 = o[ax] + thi * d[ax];
    const float ia = (fminf(pa0, pa1) - first[ax]) / sp[ax], ib = (fmaxf(pa0, pa1) - first[ax]) / sp[ax];
    int i0 = max(0, static_cast<int>(floorf(ia))), i1 = min(n[ax] - 1, static_cast<int>(ceilf(ib)));
    const int oa = (ax == 0) ? 1 : 0, ob = (ax == 2) ? 1 : 2;

    const double dax = d[ax];
    const double t0 = (static_cast<double>(first[ax]) - o[ax]) / dax, dt = static_cast<double>(sp[ax]) / dax;
    if (from_source) {   // only planes in front of the source: t(i) > 0
        const double ic = -t0 / dt;
        if (dt > 0.0) i0 = max(i0, static_cast<int>(floor(ic)) + 1);
        else i1 = min(i1, static_cast<int>(ceil(ic)) - 1);
    }
    if (i1 < i0) return 0.0f;
    const float fa0 = static_cast<float>((o[oa] + t0 * d[oa] - first[oa]) / sp[oa]);
    const float dfa = static_cast<float>(dt * d[oa] / sp[oa]);
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + (is3 ? d[2] * d[2] : 0.0f));
    const float step = sp[ax] * len / fabsf(d[ax]);
    // Rays marching along x read the x/y-transposed copy, so that the lanes of
    // a wave (neighbouring detector columns) always spread along the
    // contiguous axis of the copy they read.
    const bool tr = ax == 0 && a.vol_t != nullptr;
    const uint32_t stride[3] = {tr ? a.ny : 1u, tr ? 1u : a.nx, a.nx * a.ny};
    const uint32_t na = static_cast<uint32_t>(n[oa]), sa = stride[oa];
    const float *plane = (tr ? a.vol_t : a.vol) + static_cast<size_t>(i0) * stride[ax];
    const size_t pstep = stride[ax];
    float acc = 0.0f;
    if (!is3) {
        for (int i = i0; i <= i1; ++i, plane += pstep) {
            const float fa = fmaf(static_cast<float>(i), dfa, fa0);
            const float ffa = floorf(fa);
            const int a0 = static_cast<int>(ffa);
            const float wa = fa - ffa;
            const float v0 = static_cast<uint32_t>(a0) < na ? plane[static_cast<uint32_t>(a0) * sa] : 0.0f;
            const float v1 = static_cast<uint32_t>(a0 + 1) < na ? plane[static_cast<uint32_t>(a0 + 1) * sa] : 0.0f;
            acc += fmaf(wa, v1 - v0, v0);
        }
        return acc * step;
    }
    const float fb0 = static_cast<float>((o[ob] + t0 * d[ob] - first[ob]) / sp[ob]);
    const float dfb = static_cast<float>(dt * d[ob] / sp[ob]);
    const uint32_t nb = static_cast<uint32_t>(n[ob]), sb = stride[ob];
#pragma unroll 2
    for (int i = i0; i <= i1; ++i, plane += pstep) {
        const float fi = static_cast<float>(i);
        const float fa = fmaf(fi, dfa, fa0), fb = fmaf(fi, dfb, fb0);
        const float ffa = floorf(fa), ffb = floorf(fb);
        const int a0 = static_cast<int>(ffa), b0 = static_cast<int>(ffb);
        const float wa = fa - ffa, wb = fb - ffb;
        const bool ia0 = static_cast<uint32_t>(a0) < na, ia1 = static_cast<uint32_t>(a0 + 1) < na;
        const bool ib0 = static_cast<uint32_t>(b0) < nb, ib1 = static_cast<uint32_t>(b0 + 1) < nb;
        const uint32_t o00 = static_cast<uint32_t>(a0) * sa + static_cast<uint32_t>(b0) * sb;
        const float v00 = (ia0 && ib0) ? plane[o00] : 0.0f;
        const float v10 = (ia1 && ib0) ? plane[o00 + sa] : 0.0f;
        const float v01 = (ia0 && ib1) ? plane[o00 + sb] : 0.0f;
        const float v11 = (ia1 && ib1) ? plane[o00 + sa + sb] : 0.0f;
        const float lo = fmaf(wa, v10 - v00, v00), hi = fmaf(wa, v11 - v01, v01);
        acc += fmaf(wb, hi - lo, lo);
    }
    return acc * step;
}

// [z][y][x] -> [z][x][y], 32 x 32 tiles through LDS (both sides coalesced).
__global__ __launch_bounds__(256) void transpose_xy_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                           uint32_t nx, uint32_t ny, uint32_t tiles_x, uint32_t tiles_y)
{
    __shared__ float t[32][33];
    const uint32_t b = blockIdx.x;
    const uint32_t tx = b % tiles_x, rest = b / tiles_x;
    const uint32_t ty = rest % tiles_y, z = rest / tiles_y;
    const size_t slice = static_cast<size_t>(nx) * ny;
    const float *sp = src + z * slice;
    float *dp = dst + z * slice;
    const uint32_t lx = threadIdx.x & 31u, ly = threadIdx.x >> 5;
    for (uint32_t k = 0; k < 32u; k += 8u) {
        const uint32_t x = tx * 32u + lx, y = ty * 32u + ly + k;
        if (x < nx && y < ny) t[ly + k][lx] = sp[static_cast<size_t>(y) * nx + x];
    }
    __syncthreads();
    for (uint32_t k = 0; k < 32u; k += 8u) {
        const uint32_t y = ty * 32u + lx, x = tx * 32u + ly + k;
        if (x < nx && y < ny) dp[static_cast<size_t>(x) * ny + y] = t[lx][ly + k];
    }
}

__global__ __launch_bounds__(256) void joseph_kernel(FpLaunch a)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t nv = is_1d(a.kind) ? 1u : a.nv;
    const size_t total = static_cast<size_t>(a.nviews) * nv * a.nu;
    if (idx >= total) return;
    const uint32_t c = static_cast<uint32_t>(idx % a.nu);
    const size_t rr = idx / a.nu;
    const uint32_t r = static_cast<uint32_t>(rr % nv);
    const uint32_t v = static_cast<uint32_t>(rr / nv);
    const gct_proj_geom g = a.views[v];
    const float uc = (static_cast<float>(c) - g.det_row_center) * a.du;
    const float vr = (static_cast<float>(r) - g.det_col_center) * a.dv;
    const float D = g.src_det_dist;
    float o[3] = {0.0f, 0.0f, 0.0f}, d[3] = {0.0f, 0.0f, 0.0f};
    bool from_src = true;
    switch (a.kind) {
    case GCT_PARALLEL_1D:
    case GCT_PARALLEL_2D:
    case GCT_CONE_PARALLEL: {
        // ray {t n + s d}, n = (-sin, cos), d = (cos, sin): the adjoint of
        // t = x cos(theta + pi/2) + y sin(theta + pi/2) (coord_trans_proj.cuh:50-59)
        float sn, cs;
        sincosf(g.rot_ang, &sn, &cs);
        o[0] = -uc * sn; o[1] = uc * cs;
        d[0] = cs; d[1] = sn;
        if (a.kind == GCT_PARALLEL_2D) o[2] = -vr;
        if (a.kind == GCT_CONE_PARALLEL) o[2] = -vr * g.src_iso_dist / D;
        from_src = false;
        break;
    }
    case GCT_FAN_LINE_1D: {
        const float w0 = -g.det_row_dir[1], w1 = g.det_row_dir[0];   // rot90(eu)
        o[0] = g.src_pos[0]; o[1] = g.src_pos[1];
        d[0] = D * w0 + uc * g.det_row_dir[0];
        d[1] = D * w1 + uc * g.det_row_dir[1];
        break;
    }
    case GCT_FAN_ARC_1D: {
        float sn, cs;
        sincosf(uc, &sn, &cs);
        o[0] = g.src_pos[0]; o[1] = g.src_pos[1];
        d[0] = cs * g.iso_ray_dir[0] + sn * g.det_row_dir[0];
        d[1] = cs * g.iso_ray_dir[1] + sn * g.det_row_dir[1];
        break;
    }
    case GCT_CYLINDRICAL: {
        float sn, cs;
        sincosf(uc, &sn, &cs);
        for (int i = 0; i < 3; ++i) {
            o[i] = g.src_pos[i];
            d[i] = D * (cs * g.iso_ray_dir[i] + sn * g.det_row_dir[i]) + vr * g.det_col_dir[i];
        }
        break;
    }
    default: {   // flat
        const float *eu = g.det_row_dir, *ev = g.det_col_dir;
        const float ew[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0]};
        for (int i = 0; i < 3; ++i) {
            o[i] = g.src_pos[i];
            d[i] = D * ew[i] + uc * eu[i] + vr * ev[i];
        }
        break;
    }
    }
    a.proj[idx] = joseph_ray(a, o, d, from_src);
}

// ---- residual of a reprojection -------------------------------------------------

// partials[2 b] = sum over block b of (a - b)^2, partials[2 b + 1] = sum of b^2.
__global__ __launch_bounds__(256) void residual_kernel(const float *__restrict__ a, const float *__restrict__ b,
                                                       size_t n, double *partials)
{
    __shared__ double s1[256], s2[256];
    double x = 0.0, y = 0.0;
    for (size_t i = static_cast<size_t>(blockIdx.x) * 256u + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * 256u) {
        const double d = static_cast<double>(a[i]) - static_cast<double>(b[i]);
        x += d * d;
        y += static_cast<double>(b[i]) * static_cast<double>(b[i]);
    }
    s1[threadIdx.x] = x;
    s2[threadIdx.x] = y;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2u * blockIdx.x] = s1[0];
        partials[2u * blockIdx.x + 1u] = s2[0];
    }
}

// ---- dispatch helpers -------------------------------------------------------

template <int KIND, int S>
void launch_filter_L(const FilterLaunch &a, hipStream_t s)
{
    const uint32_t nrows = a.num_views * (is_1d(KIND) ? 1u : a.nv);
    const uint32_t blocks = (nrows + 1u) / 2u;
    if (blocks == 0) return;
    switch (a.L) {
    case 8: hipLaunchKernelGGL((filter_kernel<KIND, S, 8>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 16: hipLaunchKernelGGL((filter_kernel<KIND, S, 16>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 32: hipLaunchKernelGGL((filter_kernel<KIND, S, 32>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 64: hipLaunchKernelGGL((filter_kernel<KIND, S, 64>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 128: hipLaunchKernelGGL((filter_kernel<KIND, S, 128>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 256: hipLaunchKernelGGL((filter_kernel<KIND, S, 256>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 512: hipLaunchKernelGGL((filter_kernel<KIND, S, 512>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((filter_kernel<KIND, S, 1024>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 2048: hipLaunchKernelGGL((filter_kernel<KIND, S, 2048>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    case 4096: hipLaunchKernelGGL((filter_kernel<KIND, S, 4096>), dim3(blocks), dim3(kFiltThreads), 0, s, a); break;
    default: fail(GCT_ERR_UNSUPPORTED, "FFT length outside 8..4096");
    }
}

template <int KIND>
void launch_filter_S(const FilterLaunch &a, hipStream_t s)
{
    switch (a.storage) {
    case GCT_STORE_F16: launch_filter_L<KIND, GCT_STORE_F16>(a, s); break;
    case GCT_STORE_BF16: launch_filter_L<KIND, GCT_STORE_BF16>(a, s); break;
    default: launch_filter_L<KIND, GCT_STORE_F32>(a, s); break;
    }
}

template <int KIND, int S, bool CIRC, bool F64>
void launch_bp3d_zu(const BpLaunch &a, hipStream_t s)
{
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t zu = a.z_unroll;
    const uint32_t nbz = (a.nz + zu - 1u) / zu;
    const uint32_t blocks = nbx * nby * nbz;
    switch (zu) {
    case 4: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 4>), dim3(blocks), dim3(256), 0, s, a); break;
    case 16: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 16>), dim3(blocks), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((bp3d_kernel<KIND, S, CIRC, F64, 8>), dim3(blocks), dim3(256), 0, s, a); break;
    }
}

template <int KIND, int S, bool F64>
void launch_bp3d_lds(const BpLaunch &a, hipStream_t s)
{
    const uint32_t nbx = (a.nx + 15u) / 16u, nby = (a.ny + 15u) / 16u;
    const uint32_t zu = (a.z_unroll == 32u) ? 32u : 16u;
    const uint32_t blocks = nbx * nby * ((a.nz + zu - 1u) / zu);
    if (zu == 32u) hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 32, 64, 2>), dim3(blocks), dim3(256), 0, s, a);
    else hipLaunchKernelGGL((bp3d_lds_kernel<KIND, S, F64, 16, 40, 3>), dim3(blocks), dim3(256), 0, s, a);
}

template <int KIND, int S>
void launch_bp3d_variant(const BpLaunch &a, hipStream_t s)
{
    const bool lds = a.variant == GCT_BP_LDS;
    if constexpr (KIND == GCT_FLAT || KIND == GCT_CYLINDRICAL) {
        if (lds && a.circular) {
            if (a.setup_f64) launch_bp3d_lds<KIND, S, true>(a, s);
            else launch_bp3d_lds<KIND, S, false>(a, s);
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
        hipLaunchKernelGGL((bp2d_kernel<KIND, S>), dim3(blocks), dim3(64), 0, s, a);
    }
}

template <int S>
void launch_bp_S(const BpLaunch &a, hipStream_t s)
{
    if (is_1d(a.kind)) {
        switch (a.kind) {
        case GCT_PARALLEL_1D: launch_bp2d<GCT_PARALLEL_1D, S>(a, s); break;
        case GCT_FAN_ARC_1D: launch_bp2d<GCT_FAN_ARC_1D, S>(a, s); break;
        default: launch_bp2d<GCT_FAN_LINE_1D, S>(a, s); break;
        }
        return;
    }
    switch (a.kind) {
    case GCT_PARALLEL_2D: launch_bp3d_variant<GCT_PARALLEL_2D, S>(a, s); break;
    case GCT_CONE_PARALLEL: launch_bp3d_variant<GCT_CONE_PARALLEL, S>(a, s); break;
    case GCT_CYLINDRICAL: launch_bp3d_variant<GCT_CYLINDRICAL, S>(a, s); break;
    default: launch_bp3d_variant<GCT_FLAT, S>(a, s); break;
    }
}

}  // namespace

void launch_redundancy(int kind, const float *fan, const float *beta, uint32_t nu, uint32_t nviews, float range,
                       int mode, float value, float *red, hipStream_t s)
{
    (void)kind;
    const uint32_t n = nu * nviews;
    hipLaunchKernelGGL(redundancy_kernel, dim3((n + 255u) / 256u), dim3(256), 0, s, fan, beta, nu, nviews, range, mode,
                       value, red);
    hip_check(hipGetLastError(), "redundancy weights");
}

void launch_filter(const FilterLaunch &a, hipStream_t s)
{
    switch (a.kind) {
    case GCT_PARALLEL_1D: launch_filter_S<GCT_PARALLEL_1D>(a, s); break;
    case GCT_FAN_ARC_1D: launch_filter_S<GCT_FAN_ARC_1D>(a, s); break;
    case GCT_FAN_LINE_1D: launch_filter_S<GCT_FAN_LINE_1D>(a, s); break;
    case GCT_PARALLEL_2D: launch_filter_S<GCT_PARALLEL_2D>(a, s); break;
    case GCT_CONE_PARALLEL: launch_filter_S<GCT_CONE_PARALLEL>(a, s); break;
    case GCT_CYLINDRICAL: launch_filter_S<GCT_CYLINDRICAL>(a, s); break;
    default: launch_filter_S<GCT_FLAT>(a, s); break;
    }
    hip_check(hipGetLastError(), "ramp filter");
}

void launch_backproject(const BpLaunch &a, hipStream_t s)
{
    if (a.num_views == 0 && !a.finish && a.accumulate) return;
    switch (a.storage) {
    case GCT_STORE_F16: launch_bp_S<GCT_STORE_F16>(a, s); break;
    case GCT_STORE_BF16: launch_bp_S<GCT_STORE_BF16>(a, s); break;
    default: launch_bp_S<GCT_STORE_F32>(a, s); break;
    }
    hip_check(hipGetLastError(), "backprojection");
}

void launch_forward_project(const FpLaunch &a, hipStream_t s)
{
    const size_t total = static_cast<size_t>(a.nviews) * (is_1d(a.kind) ? 1u : a.nv) * a.nu;
    if (total == 0) return;
    if (a.vol_t) {
        const uint32_t tx = (a.nx + 31u) / 32u, ty = (a.ny + 31u) / 32u;
        hipLaunchKernelGGL(transpose_xy_kernel, dim3(tx * ty * a.nz), dim3(256), 0, s, a.vol, a.vol_t, a.nx, a.ny, tx,
                           ty);
    }
    hipLaunchKernelGGL(joseph_kernel, dim3(static_cast<uint32_t>((total + 255) / 256)), dim3(256), 0, s, a);
    hip_check(hipGetLastError(), "forward projection");
}

void launch_pack_filtered(const float *src, void *filt, int storage, uint32_t nu, uint32_t nv, uint32_t nviews,
                          uint32_t pitch, uint32_t rows_pv, hipStream_t s)
{
    const size_t n = static_cast<size_t>(nu) * nv * nviews;
    const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
    if (!blocks) return;
    switch (storage) {
    case GCT_STORE_F16: hipLaunchKernelGGL(pack_kernel<GCT_STORE_F16>, dim3(blocks), dim3(256), 0, s, src, filt, nu, nv, nviews, pitch, rows_pv); break;
    case GCT_STORE_BF16: hipLaunchKernelGGL(pack_kernel<GCT_STORE_BF16>, dim3(blocks), dim3(256), 0, s, src, filt, nu, nv, nviews, pitch, rows_pv); break;
    default: hipLaunchKernelGGL(pack_kernel<GCT_STORE_F32>, dim3(blocks), dim3(256), 0, s, src, filt, nu, nv, nviews, pitch, rows_pv); break;
    }
    hip_check(hipGetLastError(), "pack filtered projections");
}

void launch_unpack_filtered(const void *filt, float *dst, int storage, uint32_t nu, uint32_t nv, uint32_t nviews,
                            uint32_t pitch, uint32_t rows_pv, hipStream_t s)
{
    const size_t n = static_cast<size_t>(nu) * nv * nviews;
    const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
    if (!blocks) return;
    switch (storage) {
    case GCT_STORE_F16: hipLaunchKernelGGL(unpack_kernel<GCT_STORE_F16>, dim3(blocks), dim3(256), 0, s, filt, dst, nu, nv, nviews, pitch, rows_pv); break;
    case GCT_STORE_BF16: hipLaunchKernelGGL(unpack_kernel<GCT_STORE_BF16>, dim3(blocks), dim3(256), 0, s, filt, dst, nu, nv, nviews, pitch, rows_pv); break;
    default: hipLaunchKernelGGL(unpack_kernel<GCT_STORE_F32>, dim3(blocks), dim3(256), 0, s, filt, dst, nu, nv, nviews, pitch, rows_pv); break;
    }
    hip_check(hipGetLastError(), "unpack filtered projections");
}

void launch_residual(const float *a, const float *b, size_t n, double *partials, uint32_t blocks, hipStream_t s)
{
    hipLaunchKernelGGL(residual_kernel, dim3(blocks), dim3(256), 0, s, a, b, n, partials);
    hip_check(hipGetLastError(), "residual");
}

template <int KIND, int S>
void launch_batch_kind(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    if constexpr (KIND <= GCT_FAN_LINE_1D) {
        hipLaunchKernelGGL((bp_batch_kernel<KIND, S, false>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                           g.total_tiles, counter);
    } else if constexpr (KIND == GCT_FLAT || KIND == GCT_CYLINDRICAL) {
        if (g.lds)
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, true>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                               g.total_tiles, counter);
        else
            hipLaunchKernelGGL((bp_batch_kernel<KIND, S, false>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                               g.total_tiles, counter);
    } else {
        hipLaunchKernelGGL((bp_batch_kernel<KIND, S, true>), dim3(grid), dim3(256), 0, s, scans, g.num_scans,
                           g.total_tiles, counter);
    }
}

template <int S>
void launch_batch_S(const BatchScan *scans, const BatchGroup &g, uint32_t *counter, uint32_t grid, hipStream_t s)
{
    switch (g.kind) {
    case GCT_PARALLEL_1D: launch_batch_kind<GCT_PARALLEL_1D, S>(scans, g, counter, grid, s); break;
    case GCT_FAN_ARC_1D: launch_batch_kind<GCT_FAN_ARC_1D, S>(scans, g, counter, grid, s); break;
    case GCT_FAN_LINE_1D: launch_batch_kind<GCT_FAN_LINE_1D, S>(scans, g, counter, grid, s); break;
    case GCT_PARALLEL_2D: launch_batch_kind<GCT_PARALLEL_2D, S>(scans, g, counter, grid, s); break;
    case GCT_CONE_PARALLEL: launch_batch_kind<GCT_CONE_PARALLEL, S>(scans, g, counter, grid, s); break;
    case GCT_CYLINDRICAL: launch_batch_kind<GCT_CYLINDRICAL, S>(scans, g, counter, grid, s); break;
    default: launch_batch_kind<GCT_FLAT, S>(scans, g, counter, grid, s); break;
    }
}

void launch_zero_counters(uint32_t *counters, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, counters, n);
    hip_check(hipGetLastError(), "zero counters");
}

void launch_backproject_batch_group(const BatchScan *scans_dev, const BatchGroup &g, uint32_t *counter, hipStream_t s)
{
    int dev = 0, cus = 256;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    const uint32_t grid = std::max(1u, std::min(static_cast<uint32_t>(cus) * 8u, g.total_tiles));
    const BatchScan *scans = scans_dev + g.first_scan;
    switch (g.storage) {
    case GCT_STORE_F16: launch_batch_S<GCT_STORE_F16>(scans, g, counter, grid, s); break;
    case GCT_STORE_BF16: launch_batch_S<GCT_STORE_BF16>(scans, g, counter, grid, s); break;
    default: launch_batch_S<GCT_STORE_F32>(scans, g, counter, grid, s); break;
    }
    hip_check(hipGetLastError(), "batched backprojection");
}

}  // namespace gct
