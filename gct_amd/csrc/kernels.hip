// HIP kernels of the FDK path for gfx950 (CDNA4): fused pre-weighting and
// ramp filtering (in-LDS Stockham FFT), the Joseph forward projector,
// packing of filtered projections, and the dispatch of the backprojection
// kernels (whose templates live in bp_device.hpp and are instantiated per
// geometry family in bp_1d.hip, bp_parallel.hip, bp_cyl.hip, bp_flat.hip).
//
// Reference stages replaced (satidev/gct):
//   cosWeightKernel* (cos_weighting.cuh:155-263), parkerWeightKern and
//   parkerWeighting*Kernel (parker_weighting.cu:5-52), the cuFFT ramp filter
//   (ramp_filtering_fft.cuh:157-219, ramp_filtering_fft.cu:14-96).
#include "bp_device.hpp"
#include "common.hpp"

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>

namespace gct {

namespace {

constexpr float kPiRef = 3.141592654f;   // the reference's pi(), constants.cuh:6-9

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
// L complex values suffices.  Twiddles are read from an LDS copy of the table
// exp(-2 pi i m / L), m < L.

__device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }   // a * (-i)

template <int R> __device__ inline void dft(float2 *u);

template <> __device__ inline void dft<2>(float2 *u)
{
    const float2 a = u[0], b = u[1];
    u[0] = cadd(a, b);
    u[1] = csub(a, b);
}

// X1 = a + W b + W^2 d, X2 = a + W^2 b + W d with W = exp(-2 pi i / 3) = -1/2 - i sqrt(3)/2:
// both are m -/+ i s (b - d) with m = a - (b + d) / 2, s = sqrt(3)/2.
template <> __device__ inline void dft<3>(float2 *u)
{
    const float s = 0.86602540378443865f;
    const float2 a = u[0], t = cadd(u[1], u[2]), w = csub(u[1], u[2]);
    const float2 m = make_float2(fmaf(-0.5f, t.x, a.x), fmaf(-0.5f, t.y, a.y));
    u[0] = cadd(a, t);
    u[1] = make_float2(fmaf(s, w.y, m.x), fmaf(-s, w.x, m.y));
    u[2] = make_float2(fmaf(-s, w.y, m.x), fmaf(s, w.x, m.y));
}

template <> __device__ inline void dft<4>(float2 *u)
{
    const float2 a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
    const float2 c = cadd(u[1], u[3]), d = csub(u[1], u[3]);
    u[0] = cadd(a, c);
    u[2] = csub(a, c);
    u[1] = cadd(b, mul_mi(d));   // b - i d
    u[3] = csub(b, mul_mi(d));   // b + i d
}

template <> __device__ inline void dft<8>(float2 *u)
{
    float2 e[4] = {u[0], u[2], u[4], u[6]};
    float2 o[4] = {u[1], u[3], u[5], u[7]};
    dft<4>(e);
    dft<4>(o);
    const float r = 0.70710678118654752f;
    // o[m] *= W8^m, W8 = exp(-i pi / 4)
    const float2 o1 = make_float2(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
    const float2 o2 = mul_mi(o[2]);
    const float2 o3 = make_float2(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
    u[0] = cadd(e[0], o[0]);
    u[4] = csub(e[0], o[0]);
    u[1] = cadd(e[1], o1);
    u[5] = csub(e[1], o1);
    u[2] = cadd(e[2], o2);
    u[6] = csub(e[2], o2);
    u[3] = cadd(e[3], o3);
    u[7] = csub(e[3], o3);
}

// 16-point DFT as 4 x 4: length-4 DFTs over the stride-4 subsequences, the
// twiddles W16^(n1 k2), then length-4 DFTs across (output index k2 + 4 k1).
template <> __device__ inline void dft<16>(float2 *u)
{
    float2 t[4][4];
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) {
        float2 v[4] = {u[n1], u[n1 + 4], u[n1 + 8], u[n1 + 12]};
        dft<4>(v);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) t[n1][k2] = v[k2];
    }
    const float c1 = 0.92387953251128674f, s1 = 0.38268343236508977f, r = 0.70710678118654752f;
    // W16^m = exp(-2 pi i m / 16) for the products n1 k2 in {1, 2, 3, 4, 6, 9}
    t[1][1] = cmul(t[1][1], make_float2(c1, -s1));
    t[1][2] = make_float2(r * (t[1][2].x + t[1][2].y), r * (t[1][2].y - t[1][2].x));
    t[1][3] = cmul(t[1][3], make_float2(s1, -c1));
    t[2][1] = make_float2(r * (t[2][1].x + t[2][1].y), r * (t[2][1].y - t[2][1].x));
    t[2][2] = mul_mi(t[2][2]);
    t[2][3] = make_float2(r * (t[2][3].y - t[2][3].x), -r * (t[2][3].x + t[2][3].y));
    t[3][1] = cmul(t[3][1], make_float2(s1, -c1));
    t[3][2] = make_float2(r * (t[3][2].y - t[3][2].x), -r * (t[3][2].x + t[3][2].y));
    t[3][3] = cmul(t[3][3], make_float2(-c1, s1));
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) {
        float2 v[4] = {t[0][k2], t[1][k2], t[2][k2], t[3][k2]};
        dft<4>(v);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) u[k2 + 4 * k1] = v[k1];
    }
}

// One radix-R Stockham pass over L points, P the product of the radices
// already done.  Butterfly i (0 <= i < L/R) takes points i + j L/R and gives
// points (i - k) R + k + j P, k = i mod P.  `ld(i, j)` supplies point
// i + j L/R and `st(idx, j, v)` takes output point idx: the group's LDS
// buffer for inner passes; global memory, the weights and the ramp spectrum
// at the ends, which saves the LDS round trips and barriers of separate load,
// multiply and store stages.  SYNC_MID: the pass reads and writes the same
// buffer, so every read completes before the first write; SYNC_END: the next
// pass reads what this one wrote.
template <uint32_t L, uint32_t R, uint32_t P, uint32_t NT, bool SYNC_MID, bool SYNC_END, class Ld, class St>
__device__ inline void stockham_io(const float2 *tw, uint32_t tid, Ld ld, St st)
{
    constexpr uint32_t T = L / R;
    constexpr uint32_t B = (T + NT - 1) / NT;
    float2 u[B][R];
#pragma unroll
    for (uint32_t b = 0; b < B; ++b) {
        const uint32_t i = tid + b * NT;
        if (T % NT == 0 || i < T) {
#pragma unroll
            for (uint32_t j = 0; j < R; ++j) u[b][j] = ld(i, j);
        }
    }
    if constexpr (SYNC_MID) __syncthreads();
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
            for (uint32_t j = 0; j < R; ++j) st(base + j * P, j, u[b][j]);
        }
    }
    if constexpr (SYNC_END) __syncthreads();
}

// Radix order: 16 while at least 16 remain (8 then 4 for 32), then the rest;
// for lengths 3 * 2^k the powers of two come first and the radix-3 pass last,
// so every pass before it has a power-of-two product P (k = i mod P by mask).
template <uint32_t rem>
constexpr uint32_t next_radix()
{
    if constexpr (rem % 3u == 0u) {
        constexpr uint32_t p2 = rem / 3u;
        return p2 >= 16u ? 16u : (p2 > 1u ? p2 : 3u);
    } else {
        return rem == 32u ? 8u : (rem >= 16u ? 16u : rem);
    }
}

// The radix of the last pass of an L-point FFT.
template <uint32_t L, uint32_t P = 1>
constexpr uint32_t last_radix()
{
    if constexpr (P * next_radix<L / P>() >= L)
        return next_radix<L / P>();
    else
        return last_radix<L, P * next_radix<L / P>()>();
}

// Whether the group buffers of an L-point filter are padded between the
// first two passes, both radix 16.  Lane i of the first pass writes points
// 16 i .. 16 i + 15, so the lanes of one store are 16 points (128 bytes) apart
// and meet in the same LDS banks (half of the 768-point filter's LDS cycles
// were bank conflicts).  Padded, point m of that buffer sits in slot
// m + m / 16: the lanes' stores are 17 slots apart.  The larger buffers can
// cost a workgroup per CU (1024 points: 3 instead of 4), and the filter is
// still faster (0.233 against 0.274 ms on a 248-view C-arm scan).
template <uint32_t L>
constexpr bool fft_pad()
{
    return L % 256u == 0u && next_radix<L>() == 16u && next_radix<L / 16u>() == 16u;
}

// The passes of an L-point FFT from product P on, in the group's LDS buffer
// `buf`.  The first pass reads through `ld` when LD_CUSTOM (else from
// `buf`); the last always writes through `st`, which writes `buf` when
// ST_LDS.  Passes that write `buf` end with a barrier for its next reader.
// PAD: the first pass writes `buf` padded (fft_pad) and the second reads it so.
template <uint32_t L, uint32_t P, uint32_t NT, bool LD_CUSTOM, bool ST_LDS, bool PAD, class Ld, class St>
__device__ inline void fft_io(float2 *buf, const float2 *tw, uint32_t tid, Ld ld, St st)
{
    if constexpr (P < L) {
        constexpr uint32_t R = next_radix<L / P>();
        constexpr bool first = P == 1u, last = P * R == L;
        // padded slots: written by the first pass (point 16 i + j of lane i in
        // slot 17 i + j), read by the second (point i + j L / 16 in slot
        // i + i / 16 + j (L / 16 + L / 256)), both as a per-lane base plus
        // constant offsets
        constexpr bool pad_out = PAD && first, pad_in = PAD && P == 16u;
        auto lds_ld = [buf](uint32_t i, uint32_t j) {
            if constexpr (pad_in) return buf[i + (i >> 4) + j * (L / R + L / R / 16u)];
            else return buf[i + j * (L / R)];
        };
        auto lds_st = [buf](uint32_t idx, uint32_t, float2 v) {
            if constexpr (pad_out) buf[idx + (idx >> 4)] = v;
            else buf[idx] = v;
        };
        constexpr bool in_lds = !(first && LD_CUSTOM), out_lds = !last || ST_LDS;
        constexpr bool sync_mid = in_lds && out_lds, sync_end = out_lds;
        if constexpr (first && LD_CUSTOM) {
            if constexpr (last)
                stockham_io<L, R, P, NT, sync_mid, sync_end>(tw, tid, ld, st);
            else
                stockham_io<L, R, P, NT, sync_mid, sync_end>(tw, tid, ld, lds_st);
        } else {
            if constexpr (last)
                stockham_io<L, R, P, NT, sync_mid, sync_end>(tw, tid, lds_ld, st);
            else
                stockham_io<L, R, P, NT, sync_mid, sync_end>(tw, tid, lds_ld, lds_st);
        }
        fft_io<L, P * R, NT, LD_CUSTOM, ST_LDS, PAD>(buf, tw, tid, ld, st);
    }
}

// ---- fused pre-weighting + ramp filter ------------------------------------

// cos_weighting.cuh:16-153 with PCSToDCS = (pix - centre) * width; D / sqrt(q)
// is evaluated as D * rsq(q) (1 ulp).
template <int KIND>
__device__ inline float cos_weight(const FiltRec &f, float du, float dv, uint32_t c, uint32_t r)
{
    const float u = (static_cast<float>(c) - f.cu) * du;
    const float v = (static_cast<float>(r) - f.cv) * dv;
    const float d2 = f.sdd * f.sdd;
    if constexpr (KIND == GCT_FAN_ARC_1D) return cosf(u);
    if constexpr (KIND == GCT_FAN_LINE_1D) return f.sdd * __builtin_amdgcn_rsqf(fmaf(u, u, d2));
    if constexpr (KIND == GCT_CONE_PARALLEL) return f.sdd * __builtin_amdgcn_rsqf(fmaf(v, v, d2));
    if constexpr (KIND == GCT_CYLINDRICAL) return cosf(u) * (f.sdd * __builtin_amdgcn_rsqf(fmaf(v, v, d2)));
    if constexpr (KIND == GCT_FLAT) return f.sdd * __builtin_amdgcn_rsqf(fmaf(u, u, fmaf(v, v, d2)));
    return 1.0f;
}

// Threads per row pair: from 256 to 2048 points, L / 16 -- one radix-16
// butterfly per thread and pass, so every thread is busy in those passes (64
// threads left three quarters of them idle at 256 points: 0.034 against
// 0.020 ms, a quarter at 768: 0.308 against 0.279 ms); beyond, 256 (192 at
// 3072 points, one workgroup of 3 waves, was slower: 0.233 against 0.212 ms);
// below, L under 64 and 64 from there.  A workgroup holds the 256 / NT row
// pairs that fit 256 threads (filt_block), which share the LDS twiddle table.
template <uint32_t L>
constexpr uint32_t filt_nt()
{
    if constexpr (L >= 256u)
        return L / 16u <= 128u ? L / 16u : 256u;
    else
        return L < 64u ? L : 64u;
}
constexpr uint32_t kFiltBlock = 256;
template <uint32_t L>
constexpr uint32_t filt_block()
{
    return kFiltBlock / filt_nt<L>() * filt_nt<L>();
}

// One group of NT threads filters two detector rows as one complex signal
// z = a + i b: IFFT(FFT(z) H) = (a * h) + i (b * h) because the kernel h is
// real.  H is the FFT of the N-tap kernel circularly shifted by floor(N/2)
// and scaled by 1/L, so output sample c is the reference's y[c + floor(N/2)] / L
// (ramp_filtering_fft.cu:45-62): the truncated linear convolution, which any
// L >= N + floor(N/2) gives without wrap-around (fft_len; the reference's L
// is nextPow2(2 N)).  The inverse FFT is conj(FFT(conj(.))).
template <int KIND, int S, uint32_t L>
__global__ __launch_bounds__(kFiltBlock) void filter_kernel(FilterLaunch a)
{
    constexpr uint32_t NT = filt_nt<L>(), G = kFiltBlock / NT;
    constexpr bool kPad = fft_pad<L>();
    __shared__ float2 tw[L];
    __shared__ float2 bufs[G][kPad ? L + L / 16u : L];
    using St = Stor<S>;
    for (uint32_t i = threadIdx.x; i < L; i += filt_block<L>()) tw[i] = a.tw[i];
    const uint32_t grp = threadIdx.x / NT, tid = threadIdx.x % NT;
    float2 *buf = bufs[grp];
    const uint32_t nv = a.nv, nu = a.nu;
    const uint32_t nrows = a.num_views * nv;
    const uint32_t r0 = 2u * (blockIdx.x * G + grp), r1 = r0 + 1u;
    const bool has0 = r0 < nrows, has1 = r1 < nrows;
    const uint32_t q0 = has0 ? r0 : 0u, q1 = has1 ? r1 : q0;
    const uint32_t v0 = a.first_view + q0 / nv, j0 = q0 % nv;
    const uint32_t v1 = a.first_view + q1 / nv, j1 = q1 % nv;
    const float *pa = a.proj + (static_cast<size_t>(v0 - a.proj_view_offset) * nv + j0) * nu;
    const float *pb = a.proj + (static_cast<size_t>(v1 - a.proj_view_offset) * nv + j1) * nu;
    const FiltRec f0 = a.frec[v0], f1 = a.frec[v1];
    const float *red0 = a.red + static_cast<size_t>(v0) * nu;
    const float *red1 = a.red + static_cast<size_t>(v1) * nu;

    // Forward FFT: the first pass loads the weighted rows straight from global
    // memory (points at or beyond kMaxN >= nu are zero: about a third of its
    // inputs are constant zeros); the last pass multiplies by the ramp spectrum
    // and conjugates before writing LDS.
    constexpr uint32_t kMaxN = (2u * L + 1u) / 3u;   // the widest row fft_len gives L for
    const float du = a.du, dv = a.dv;
    auto ld_rows = [&](uint32_t i, uint32_t j) -> float2 {
        constexpr uint32_t R0 = next_radix<L>();
        if (j * (L / R0) >= kMaxN) return make_float2(0.0f, 0.0f);   // j is a constant: no branch
        const uint32_t c = i + j * (L / R0);
        const bool in = c < nu && has0;
        const float xa = in ? pa[c] : 0.0f, wa = in ? red0[c] : 0.0f;
        const float xb = (in && has1) ? pb[c] : 0.0f, wb = (in && has1) ? red1[c] : 0.0f;
        // fdk_axial.cuh:171-173: p *= cosine weight, then p *= redundancy weight
        return make_float2((xa * cos_weight<KIND>(f0, du, dv, c, j0)) * wa,
                           (xb * cos_weight<KIND>(f1, du, dv, c, j1)) * wb);
    };
    const float2 *__restrict__ H = a.H;
    auto st_spectrum = [buf, H](uint32_t idx, uint32_t, float2 v) {
        const float2 z = cmul(v, H[idx]);
        buf[idx] = make_float2(z.x, -z.y);
    };
    // (the first pass uses no twiddles: the barrier that ends it also
    // publishes the table)
    fft_io<L, 1, NT, true, true, kPad>(buf, tw, tid, ld_rows, st_spectrum);

    // Inverse FFT (conj . FFT . conj): the last pass writes the filtered rows
    // straight into the padded layout; only points below nu are kept.
    typename St::T *out = static_cast<typename St::T *>(a.filt);
    const size_t view_elems = static_cast<size_t>(a.rows_pv) * a.pitch;
    typename St::T *oa = out + static_cast<size_t>(v0) * view_elems + static_cast<size_t>(j0 + 1u) * a.pitch + 1u;
    typename St::T *ob = out + static_cast<size_t>(v1) * view_elems + static_cast<size_t>(j1 + 1u) * a.pitch + 1u;
    auto st_rows = [&](uint32_t idx, uint32_t j, float2 z) {
        constexpr uint32_t RL = last_radix<L>();
        if (j * (L / RL) >= kMaxN) return;   // idx >= j L / RL >= kMaxN >= nu
        if (idx < nu && has0) {
            oa[idx] = St::cv(z.x);
            if (has1) ob[idx] = St::cv(-z.y);
        }
    };
    auto ld_none = [](uint32_t, uint32_t) { return make_float2(0.0f, 0.0f); };
    fft_io<L, 1, NT, false, false, kPad>(buf, tw, tid, ld_none, st_rows);
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
// The plane march of one ray: the axis its direction is most aligned with
// (in voxel units), the planes of voxel centres normal to it that the ray
// crosses inside the volume grown by one voxel (and, for rays from a source,
// in front of it), plane i crossed at t(i) = t0 + i dt, and the path length
// per plane.  ax = -1: the ray misses the volume.
struct March {
    int ax, oa, ob;
    int i0, i1;
    double t0, dt;
    float step;
};

__device__ March joseph_march(const FpLaunch &a, const float o[3], const float d[3], bool from_source)
{
    March m{};
    m.ax = -1;
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
    if (!(fabsf(d[ax]) > 0.0f)) return m;
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
            return m;
        }
    }
    if (from_source) tlo = fmaxf(tlo, 0.0f);
    if (!(thi > tlo)) return m;
    const float pa0 = o[ax] + tlo * d[ax], pa1 = o[ax] + thi * d[ax];
    const float ia = (fminf(pa0, pa1) - first[ax]) / sp[ax], ib = (fmaxf(pa0, pa1) - first[ax]) / sp[ax];
    int i0 = max(0, static_cast<int>(floorf(ia))), i1 = min(n[ax] - 1, static_cast<int>(ceilf(ib)));
    const double dax = d[ax];
    const double t0 = (static_cast<double>(first[ax]) - o[ax]) / dax, dt = static_cast<double>(sp[ax]) / dax;
    if (from_source) {   // only planes in front of the source: t(i) > 0
        const double ic = -t0 / dt;
        if (dt > 0.0) i0 = max(i0, static_cast<int>(floor(ic)) + 1);
        else i1 = min(i1, static_cast<int>(ceil(ic)) - 1);
    }
    if (i1 < i0) return m;
    const float len = sqrtf(d[0] * d[0] + d[1] * d[1] + (is3 ? d[2] * d[2] : 0.0f));
    m.ax = ax;
    m.oa = (ax == 0) ? 1 : 0;
    m.ob = (ax == 2) ? 1 : 2;
    m.i0 = i0;
    m.i1 = i1;
    m.t0 = t0;
    m.dt = dt;
    m.step = sp[ax] * len / fabsf(d[ax]);
    return m;
}

// In-plane padded coordinates of a ray marching along x or y through the
// zero-padded copy for its axis: a = the other of x, y, b = z, each linear
// in the plane index (coordinate + 1, so that clamping to [0, n + 1] lands
// every out-of-volume tap on zero padding).
struct PadLine {
    float fa0p, dfa, fb0p, dfb;
};

__device__ inline PadLine pad_line(const FpLaunch &a, const March &m, const float o[3], const float d[3])
{
    const float first[3] = {a.x0, a.y0, a.z0};
    const float sp[3] = {a.dx, a.dy, a.dz};
    const int oa = m.oa;
    PadLine l;
    l.fa0p = static_cast<float>((o[oa] + m.t0 * d[oa] - first[oa]) / sp[oa] + 1.0);
    l.dfa = static_cast<float>(m.dt * d[oa] / sp[oa]);
    l.fb0p = static_cast<float>((o[2] + m.t0 * d[2] - first[2]) / sp[2] + 1.0);
    l.dfb = static_cast<float>(m.dt * d[2] / sp[2]);
    return l;
}

__device__ float joseph_ray(const FpLaunch &a, const float o[3], const float d[3], bool from_source)
{
    const bool is3 = !is_1d(a.kind);
    const March m = joseph_march(a, o, d, from_source);
    if (m.ax < 0) return 0.0f;
    const float first[3] = {a.x0, a.y0, a.z0};
    const float sp[3] = {a.dx, a.dy, a.dz};
    const int n[3] = {static_cast<int>(a.nx), static_cast<int>(a.ny), static_cast<int>(is3 ? a.nz : 1u)};
    const int ax = m.ax, oa = m.oa, ob = m.ob, i0 = m.i0, i1 = m.i1;
    const double t0 = m.t0, dt = m.dt;
    const float step = m.step;
    if (is3 && ax != 2 && a.pad != nullptr) {
        // Rays marching along x or y read the zero-padded copy for their axis
        // (joseph_pad_elems): in-plane a = the other of x, y (contiguous in the
        // copy), b = z.  Padded coordinates (+1) clamped to [0, n + 1] keep both
        // taps inside the copy and give exactly the bounds-checked taps: a tap
        // outside the volume lands on zero padding.  Offsets are 32-bit from the
        // copies' base, the z stride fits 24 bits (checked on the host).
        const bool xm = ax == 0;
        const uint32_t na = xm ? a.ny : a.nx, w = na + 3u;
        const uint32_t sb = (xm ? a.nx : a.ny) * w;
        const uint32_t base = xm ? (a.nz + 3u) * (a.ny * (a.nx + 3u)) : 0u;
        const PadLine l = pad_line(a, m, o, d);
        const float fa0p = l.fa0p, dfa = l.dfa, fb0p = l.fb0p, dfb = l.dfb;
        const float nap = static_cast<float>(na + 1u), nbp = static_cast<float>(a.nz + 1u);
        // byte offsets (the copies hold < 2^30 floats): each tap is one load from
        // the uniform base plus a 32-bit offset, the +1 column an immediate
        const char *__restrict__ P = reinterpret_cast<const char *>(a.pad);
        const uint32_t sb4 = sb * 4u, w4 = w * 4u;
        uint32_t pb = (base + static_cast<uint32_t>(i0) * w) * 4u;
        float acc = 0.0f;
#pragma unroll 8   // independent planes: eight in flight per thread
        for (int i = i0; i <= i1; ++i, pb += w4) {
            const float fi = static_cast<float>(i);
            const float ca = __builtin_amdgcn_fmed3f(fmaf(fi, dfa, fa0p), 0.0f, nap);
            const float cb = __builtin_amdgcn_fmed3f(fmaf(fi, dfb, fb0p), 0.0f, nbp);
            const uint32_t b00 = (__umul24(static_cast<uint32_t>(cb), sb) + static_cast<uint32_t>(ca)) * 4u + pb;
            const float wa = __builtin_amdgcn_fractf(ca), wb = __builtin_amdgcn_fractf(cb);
            const uint32_t b01 = b00 + sb4;
            const float v00 = *reinterpret_cast<const float *>(P + b00);
            const float v10 = *reinterpret_cast<const float *>(P + b00 + 4);
            const float v01 = *reinterpret_cast<const float *>(P + b01);
            const float v11 = *reinterpret_cast<const float *>(P + b01 + 4);
            const float lo = fmaf(wa, v10 - v00, v00), hi = fmaf(wa, v11 - v01, v01);
            acc += fmaf(wb, hi - lo, lo);
        }
        return acc * step;
    }
    const float fa0 = static_cast<float>((o[oa] + t0 * d[oa] - first[oa]) / sp[oa]);
    const float dfa = static_cast<float>(dt * d[oa] / sp[oa]);
    // the volume itself, with bounds-checked taps
    const uint32_t stride[3] = {1u, a.nx, a.nx * a.ny};
    const uint32_t na = static_cast<uint32_t>(n[oa]), sa = stride[oa];
    const float *plane = a.vol + static_cast<size_t>(i0) * stride[ax];
    const size_t pstep = stride[ax];
    float acc = 0.0f;
    if (!is3) {
#pragma unroll 8
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
#pragma unroll 8   // independent planes: eight in flight per thread
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

// The forward projector's zero-padded copies (joseph_pad_elems).
// y-march copy [z+3][y][x+3]: one padded row per workgroup, x fastest on both sides.
__global__ __launch_bounds__(256) void pad_ymarch_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                         uint32_t nx, uint32_t ny, uint32_t nz)
{
    const uint32_t row = blockIdx.x, w = nx + 3u;
    const uint32_t zp = row / ny, y = row - zp * ny;
    const bool zin = zp >= 1u && zp <= nz;
    const float *sr = src + (static_cast<size_t>(zin ? zp - 1u : 0u) * ny + y) * nx;
    float *dr = dst + static_cast<size_t>(row) * w;
    for (uint32_t xp = threadIdx.x; xp < w; xp += 256u)
        dr[xp] = (zin && xp >= 1u && xp <= nx) ? sr[xp - 1u] : 0.0f;
}

// x-march copy [z+3][x][y+3], written after the y-march copy: 32 x 32 tiles
// of (x, padded y) through LDS, so both the reads (x fastest) and the writes
// (y fastest) are coalesced.
__global__ __launch_bounds__(256) void pad_xmarch_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                                         uint32_t nx, uint32_t ny, uint32_t nz, uint32_t tiles_x,
                                                         uint32_t tiles_y)
{
    __shared__ float t[32][33];
    const uint32_t b = blockIdx.x;
    const uint32_t tx = b % tiles_x, rest = b / tiles_x;
    const uint32_t ty = rest % tiles_y, zp = rest / tiles_y;
    const uint32_t w = ny + 3u;
    const bool zin = zp >= 1u && zp <= nz;
    const float *sp = src + static_cast<size_t>(zin ? zp - 1u : 0u) * nx * ny;
    float *dp = dst + static_cast<size_t>(zp) * nx * w;
    const uint32_t lx = threadIdx.x & 31u, ly = threadIdx.x >> 5;
    for (uint32_t k = 0; k < 32u; k += 8u) {
        const uint32_t x = tx * 32u + lx, yp = ty * 32u + ly + k;
        t[ly + k][lx] = (zin && x < nx && yp >= 1u && yp <= ny) ? sp[static_cast<size_t>(yp - 1u) * nx + x] : 0.0f;
    }
    __syncthreads();
    for (uint32_t k = 0; k < 32u; k += 8u) {
        const uint32_t yp = ty * 32u + lx, x = tx * 32u + ly + k;
        if (x < nx && yp < w) dp[static_cast<size_t>(x) * w + yp] = t[lx][ly + k];
    }
}

// The ray of detector pixel (r, c) in view v: origin o, direction d, and
// whether only t > 0 counts (rays from a source).
__device__ inline bool ray_of(const FpLaunch &a, uint32_t v, uint32_t r, uint32_t c, float o[3], float d[3])
{
    const gct_proj_geom g = a.views[v];
    const float uc = (static_cast<float>(c) - g.det_row_center) * a.du;
    const float vr = (static_cast<float>(r) - g.det_col_center) * a.dv;
    const float D = g.src_det_dist;
    for (int i = 0; i < 3; ++i) o[i] = d[i] = 0.0f;
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
    return from_src;
}

__global__ __launch_bounds__(256) void joseph_kernel(FpLaunch a)
{
    const size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t nv = is_1d(a.kind) ? 1u : a.nv;
    const size_t total = static_cast<size_t>(a.nviews) * nv * a.nu;
    if (idx >= total) return;
    // Rays are taken detector row by detector row, all views of a row in
    // turn: the rays of one row meet the volume in one z slab for every view,
    // so consecutive workgroups reuse that slab from L2 / MALL instead of
    // sweeping the whole volume once per view.
    const uint32_t c = static_cast<uint32_t>(idx % a.nu);
    const size_t rv = idx / a.nu;
    const uint32_t v = static_cast<uint32_t>(rv % a.nviews);
    const uint32_t r = static_cast<uint32_t>(rv / a.nviews);
    const size_t out = (static_cast<size_t>(v) * nv + r) * a.nu + c;
    float o[3], d[3];
    const bool from_src = ray_of(a, v, r, c, o, d);
    a.proj[out] = joseph_ray(a, o, d, from_src);
}

// ---- Joseph projection through LDS-staged bricks (3-D kinds) ------------------
//
// A workgroup traces the rays of kJVB views x kJTR detector rows x kJTC
// columns together (each thread: one detector pixel in kJVB consecutive
// views), plane segment by plane segment along the march axis the block
// shares.  For each segment of kJB planes it takes the bounding box of all
// its rays' taps in those planes (in the padded copy's in-plane
// coordinates), stages that brick of the zero-padded copy in LDS -- each
// voxel is read once per workgroup and segment, where joseph_kernel reads it
// once per tap, from L2 or beyond -- and takes the bilinear taps from LDS.
// The views of a block turn by a few tenths of a degree, so their rays
// through a tile share most of the brick.  Per plane the arithmetic and the
// order of the sums are joseph_ray's, so each ray's value is the same bit for
// bit.  Rays whose dominant axis is not the block's (z, or the other in-plane
// axis near 45 degrees) are traced by joseph_ray afterwards, and a segment
// whose brick would not fit takes its taps from the copy in global memory.
constexpr uint32_t kJTC = 32, kJTR = 8, kJVB = 4;   // rays of a workgroup: columns, rows, views
// LDS row pitch 96: a wave's lanes trace two detector rows, whose taps in a
// plane lie about one brick row apart; rows 96 floats apart put the two
// rows' taps 32 banks apart, where a pitch of 65 (one bank) made most of
// them meet in the same banks (18 % of LDS cycles bank conflicts; 89.0 against
// 84.1 ms on a 475^3 request).  12 rows keep the brick at 37 KB, 4 workgroups
// per CU.
constexpr uint32_t kJB = 8, kJHM = 12, kJWM = 64;   // brick: planes, at most 12 rows of 64 taps
static_assert(kJHM % 4u == 0u, "the 4 waves stage whole sets of 4 rows");
constexpr uint32_t kJWP = 96u;                      // LDS row pitch
using LdsF = __attribute__((address_space(3))) const float;   // an LDS float, addressed by its 32-bit offset

// 4 workgroups of 37 KB of LDS fit a CU: ask for the registers of 4 waves per SIMD
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
void joseph_lds_kernel(FpLaunch a, uint32_t tiles_c, uint32_t tiles_r)
{
    static_assert(kJTC * kJTR == 256, "one detector pixel per thread");
    __shared__ float brick[kJB * kJHM * kJWP];
    __shared__ float box[4][4];   // per wave: min a, max a, min b, max b
    __shared__ int span[2];       // lowest first plane, highest last plane of the block's staged rays
    __shared__ uint32_t votes[2]; // rays marching along x, along y
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    // view blocks fastest, dealt to the XCDs in contiguous runs: an XCD traces
    // neighbouring views of one tile in turn, which share most of their bricks
    const uint32_t nvb = (a.nviews + kJVB - 1u) / kJVB;
    const uint32_t t = xcd_tile(blockIdx.x, gridDim.x);
    const uint32_t vb = t % nvb, rest = t / nvb;
    const uint32_t c = (rest % tiles_c) * kJTC + (tid & (kJTC - 1u));
    const uint32_t r = (rest / tiles_c) * kJTR + tid / kJTC;
    const bool pix = c < a.nu && r < a.nv;
    if (tid == 0) {
        votes[0] = votes[1] = 0u;
        span[0] = 0x7FFFFFFF;
        span[1] = -1;
    }
    __syncthreads();
    // per ray: 0 misses the volume, 1 staged march, 2 joseph_ray afterwards
    uint32_t mode[kJVB];
    int axis_of[kJVB], i0[kJVB], i1[kJVB];
    float fa[kJVB], da[kJVB], fb[kJVB], db[kJVB], step[kJVB], acc[kJVB];
#pragma unroll
    for (uint32_t j = 0; j < kJVB; ++j) {
        mode[j] = 0u;
        acc[j] = 0.0f;
        const uint32_t v = vb * kJVB + j;
        if (!pix || v >= a.nviews) continue;
        float o[3], d[3];
        const bool from_src = ray_of(a, v, r, c, o, d);
        const March m = joseph_march(a, o, d, from_src);
        if (m.ax < 0) continue;
        axis_of[j] = m.ax;
        mode[j] = m.ax == 2 ? 2u : 1u;
        if (mode[j] == 1u) {
            atomicAdd(&votes[m.ax], 1u);
            const PadLine l = pad_line(a, m, o, d);
            fa[j] = l.fa0p;
            da[j] = l.dfa;
            fb[j] = l.fb0p;
            db[j] = l.dfb;
            i0[j] = m.i0;
            i1[j] = m.i1;
            step[j] = m.step;
        }
    }
    __syncthreads();
    const int axis = votes[0] >= votes[1] ? 0 : 1;
#pragma unroll
    for (uint32_t j = 0; j < kJVB; ++j) {
        if (mode[j] == 1u && axis_of[j] != axis) mode[j] = 2u;
        if (mode[j] == 1u) {
            atomicMin(&span[0], i0[j]);
            atomicMax(&span[1], i1[j]);
        }
    }
    __syncthreads();
    const int lo = span[0], hi = span[1];
    const bool xm = axis == 0;
    const uint32_t na = xm ? a.ny : a.nx, w = na + 3u;
    const uint32_t sb = (xm ? a.nx : a.ny) * w;
    const uint32_t base = xm ? (a.nz + 3u) * (a.ny * (a.nx + 3u)) : 0u;
    const float nap = static_cast<float>(na + 1u), nbp = static_cast<float>(a.nz + 1u);
    const float *__restrict__ P = a.pad;
    // both padded copies, for the staging loads' range check (< 2^30 floats: joseph_pad_elems)
    const uint32_t pad_bytes =
        4u * (a.nz + 3u) * (a.ny * (a.nx + 3u) + a.nx * (a.ny + 3u));
    const __amdgpu_buffer_rsrc_t prs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(P), 0, static_cast<int>(pad_bytes), 0x00020000);
    const uint32_t brick_lds = static_cast<uint32_t>(reinterpret_cast<uintptr_t>((LdsF *)brick));
    for (int s0 = lo; s0 <= hi; s0 += static_cast<int>(kJB)) {
        const int s1 = s0 + static_cast<int>(kJB) - 1;
        // bounding box of this thread's taps in the segment: the clamped padded
        // coordinates are monotone in the plane index, so their ends bound them
        float mina = 3.0e38f, maxa = -1.0f, minb = 3.0e38f, maxb = -1.0f;
#pragma unroll
        for (uint32_t j = 0; j < kJVB; ++j) {
            if (mode[j] != 1u) continue;
            const int pl = max(s0, i0[j]), ph = min(s1, i1[j]);
            if (pl > ph) continue;
            const float a0 = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(pl), da[j], fa[j]), 0.0f, nap);
            const float a1 = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(ph), da[j], fa[j]), 0.0f, nap);
            const float b0 = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(pl), db[j], fb[j]), 0.0f, nbp);
            const float b1 = __builtin_amdgcn_fmed3f(fmaf(static_cast<float>(ph), db[j], fb[j]), 0.0f, nbp);
            mina = fminf(mina, fminf(a0, a1));
            maxa = fmaxf(maxa, fmaxf(a0, a1));
            minb = fminf(minb, fminf(b0, b1));
            maxb = fmaxf(maxb, fmaxf(b0, b1));
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            mina = fminf(mina, __shfl_xor(mina, off));
            maxa = fmaxf(maxa, __shfl_xor(maxa, off));
            minb = fminf(minb, __shfl_xor(minb, off));
            maxb = fmaxf(maxb, __shfl_xor(maxb, off));
        }
        if (lane == 0) {
            box[wave][0] = mina;
            box[wave][1] = maxa;
            box[wave][2] = minb;
            box[wave][3] = maxb;
        }
        __syncthreads();   // also: every thread is done with the previous segment's brick
        mina = fminf(fminf(box[0][0], box[1][0]), fminf(box[2][0], box[3][0]));
        maxa = fmaxf(fmaxf(box[0][1], box[1][1]), fmaxf(box[2][1], box[3][1]));
        minb = fminf(fminf(box[0][2], box[1][2]), fminf(box[2][2], box[3][2]));
        maxb = fmaxf(fmaxf(box[0][3], box[1][3]), fmaxf(box[2][3], box[3][3]));
        const bool any = maxa >= 0.0f;   // some ray of the block has taps in this segment
        const uint32_t alo = any ? static_cast<uint32_t>(mina) : 0u, blo = any ? static_cast<uint32_t>(minb) : 0u;
        const uint32_t W = any ? static_cast<uint32_t>(maxa) - alo + 2u : 0u;
        const uint32_t H = any ? static_cast<uint32_t>(maxb) - blo + 2u : 0u;
        const bool fits = any && W <= kJWM && H <= kJHM;
        const uint32_t np = static_cast<uint32_t>(min(static_cast<int>(kJB), hi - s0 + 1));
        if (fits) {
            // Brick element (p, hb, la) = copy element base + (s0 + p) w + (blo + hb) sb + alo + la:
            // wave q stages (plane, row) pairs q, q + 4, ... -- pair 4 k + q is plane k / (kJHM / 4),
            // row 4 (k % (kJHM / 4)) + q -- one tap per lane, whole rows of 64 lanes and all kJB planes whatever
            // the box: elements past its columns, rows or planes land in the brick and are never
            // read (the traced taps stay within columns < W, rows < H and planes < np).  Skipping
            // the rows past the box (by scalar branches, or by an offset the range check drops)
            // fetched fewer bytes but took longer (measured with 16-row bricks): the per-row
            // selects cost more issue than the extra loads cost bandwidth.  The row's element index is wave-uniform
            // (scalar); the loads are raw buffer loads with the whole byte offset in the lane's
            // (range-checked) operand, so an element past the copies reads 0 instead of faulting.
            // In two halves: 12 loads in flight per lane, then their stores, each at an immediate
            // offset from the lane's column.
            const uint32_t wq = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(wave)));
            const uint32_t alo_s = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(alo)));
            const uint32_t blo_s = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(blo)));
            const uint32_t e0 = base + static_cast<uint32_t>(s0) * w + (blo_s + wq) * sb + alo_s;
            const uint32_t lane4 = lane * 4u;
            float *bw = brick + wq * kJWP + lane;
            constexpr uint32_t kRows = kJB * kJHM / 4u, kHalf = kRows / 2u;
#pragma unroll
            for (uint32_t h = 0; h < 2u; ++h) {
                float val[kHalf];
#pragma unroll
                for (uint32_t k = 0; k < kHalf; ++k) {
                    const uint32_t kk = h * kHalf + k, p = kk / (kJHM / 4u), hq = kk % (kJHM / 4u);
                    const uint32_t e = e0 + p * w + 4u * hq * sb;
                    val[k] = __uint_as_float(
                        __builtin_amdgcn_raw_buffer_load_b32(prs, static_cast<int>(e * 4u + lane4), 0, 0));
                }
#pragma unroll
                for (uint32_t k = 0; k < kHalf; ++k) {
                    const uint32_t kk = h * kHalf + k, p = kk / (kJHM / 4u), hq = kk % (kJHM / 4u);
                    bw[(p * kJHM + 4u * hq) * kJWP] = val[k];
                }
            }
        }
        __syncthreads();
        if (!any) continue;
        if (fits) {
            // LDS byte address of the brick element (0, 0, 0) less that of tap (alo, blo)
            const uint32_t org = brick_lds - 4u * (blo * kJWP + alo);
#pragma unroll
            for (uint32_t j = 0; j < kJVB; ++j) {
                if (mode[j] != 1u) continue;
                const int pl = max(s0, i0[j]), ph = min(s1, i1[j]);
                if (pl > ph) continue;
                float s = acc[j];
                const float fs0 = static_cast<float>(s0), fs1 = static_cast<float>(s1);
                // the unclamped coordinates at both ends: inside [0, n + 1], every plane's
                // are (they are monotone), and the clamps would return them unchanged
                const float ea0 = fmaf(fs0, da[j], fa[j]), ea1 = fmaf(fs1, da[j], fa[j]);
                const float eb0 = fmaf(fs0, db[j], fb[j]), eb1 = fmaf(fs1, db[j], fb[j]);
                const bool inside = fminf(ea0, ea1) >= 0.0f && fmaxf(ea0, ea1) <= nap && fminf(eb0, eb1) >= 0.0f &&
                                    fmaxf(eb0, eb1) <= nbp;
                if (pl == s0 && ph == s1 && inside) {
                    // the ray crosses the whole segment inside the volume (the common case):
                    // every tap lies in the box, no plane is masked, no clamp binds
#pragma unroll
                    for (uint32_t p = 0; p < kJB; ++p) {
                        const float fi = fs0 + static_cast<float>(p);   // exact: plane indices < 2^24
                        const float ca = fmaf(fi, da[j], fa[j]);
                        const float cb = fmaf(fi, db[j], fb[j]);
                        const float wa = __builtin_amdgcn_fractf(ca), wb = __builtin_amdgcn_fractf(cb);
                        // LDS byte address of tap (ia, ib) of plane p: one shift-add and one
                        // 24-bit multiply-add from the absolute tap indices
                        const uint32_t ba = __umul24(static_cast<uint32_t>(cb), 4u * kJWP) +
                                            ((static_cast<uint32_t>(ca) << 2) + org + p * (4u * kJHM * kJWP));
                        LdsF *q = reinterpret_cast<LdsF *>(static_cast<uintptr_t>(ba));
                        const float v00 = q[0], v10 = q[1], v01 = q[kJWP], v11 = q[kJWP + 1u];
                        const float lo2 = fmaf(wa, v10 - v00, v00), hi2 = fmaf(wa, v11 - v01, v01);
                        s += fmaf(wb, hi2 - lo2, lo2);
                    }
                } else {
#pragma unroll
                    for (uint32_t p = 0; p < kJB; ++p) {
                        const int i = s0 + static_cast<int>(p);
                        const bool in = i >= pl && i <= ph;
                        const float fi = fs0 + static_cast<float>(p);
                        const float ca = __builtin_amdgcn_fmed3f(fmaf(fi, da[j], fa[j]), 0.0f, nap);
                        const float cb = __builtin_amdgcn_fmed3f(fmaf(fi, db[j], fb[j]), 0.0f, nbp);
                        const float wa = __builtin_amdgcn_fractf(ca), wb = __builtin_amdgcn_fractf(cb);
                        // masked planes may lie outside the box: the unsigned clamps keep
                        // their reads inside the brick
                        const uint32_t la = min(static_cast<uint32_t>(ca) - alo, W - 2u);
                        const uint32_t lb = min(static_cast<uint32_t>(cb) - blo, H - 2u);
                        const float *q = brick + p * (kJHM * kJWP) + __umul24(lb, kJWP) + la;
                        const float v00 = q[0], v10 = q[1], v01 = q[kJWP], v11 = q[kJWP + 1u];
                        const float lo2 = fmaf(wa, v10 - v00, v00), hi2 = fmaf(wa, v11 - v01, v01);
                        const float smp = fmaf(wb, hi2 - lo2, lo2);
                        s = in ? s + smp : s;
                    }
                }
                acc[j] = s;
            }
        } else {
            // the brick would not fit: taps from the copy in global memory, as joseph_ray takes them
            for (uint32_t j = 0; j < kJVB; ++j) {
                if (mode[j] != 1u) continue;
                const int pl = max(s0, i0[j]), ph = min(s1, i1[j]);
                float s = acc[j];
                for (int i = pl; i <= ph; ++i) {
                    const float fi = static_cast<float>(i);
                    const float ca = __builtin_amdgcn_fmed3f(fmaf(fi, da[j], fa[j]), 0.0f, nap);
                    const float cb = __builtin_amdgcn_fmed3f(fmaf(fi, db[j], fb[j]), 0.0f, nbp);
                    const float wa = __builtin_amdgcn_fractf(ca), wb = __builtin_amdgcn_fractf(cb);
                    const uint32_t e = base + static_cast<uint32_t>(i) * w + __umul24(static_cast<uint32_t>(cb), sb) +
                                       static_cast<uint32_t>(ca);
                    const float v00 = P[e], v10 = P[e + 1u], v01 = P[e + sb], v11 = P[e + sb + 1u];
                    const float lo2 = fmaf(wa, v10 - v00, v00), hi2 = fmaf(wa, v11 - v01, v01);
                    s += fmaf(wb, hi2 - lo2, lo2);
                }
                acc[j] = s;
            }
        }
    }
    const uint32_t nvr = a.nv;
#pragma unroll
    for (uint32_t j = 0; j < kJVB; ++j) {
        const uint32_t v = vb * kJVB + j;
        if (!pix || v >= a.nviews) continue;
        float out = 0.0f;
        if (mode[j] == 1u) {
            out = acc[j] * step[j];
        } else if (mode[j] == 2u) {
            float o[3], d[3];
            const bool from_src = ray_of(a, v, r, c, o, d);
            out = joseph_ray(a, o, d, from_src);
        }
        a.proj[(static_cast<size_t>(v) * nvr + r) * a.nu + c] = out;
    }
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
    const uint32_t pairs = (nrows + 1u) / 2u;
    if (pairs == 0) return;
    // workgroups of 256 / filt_nt<L>() row pairs, filt_nt<L>() threads each
    auto go = [&](auto kern, uint32_t nt) {
        const uint32_t g = kFiltBlock / nt;
        hipLaunchKernelGGL(kern, dim3((pairs + g - 1u) / g), dim3(g * nt), 0, s, a);
    };
    switch (a.L) {
    case 48: go(filter_kernel<KIND, S, 48>, filt_nt<48>()); break;
    case 96: go(filter_kernel<KIND, S, 96>, filt_nt<96>()); break;
    case 192: go(filter_kernel<KIND, S, 192>, filt_nt<192>()); break;
    case 384: go(filter_kernel<KIND, S, 384>, filt_nt<384>()); break;
    case 768: go(filter_kernel<KIND, S, 768>, filt_nt<768>()); break;
    case 1536: go(filter_kernel<KIND, S, 1536>, filt_nt<1536>()); break;
    case 3072: go(filter_kernel<KIND, S, 3072>, filt_nt<3072>()); break;
    case 8: go(filter_kernel<KIND, S, 8>, filt_nt<8>()); break;
    case 16: go(filter_kernel<KIND, S, 16>, filt_nt<16>()); break;
    case 32: go(filter_kernel<KIND, S, 32>, filt_nt<32>()); break;
    case 64: go(filter_kernel<KIND, S, 64>, filt_nt<64>()); break;
    case 128: go(filter_kernel<KIND, S, 128>, filt_nt<128>()); break;
    case 256: go(filter_kernel<KIND, S, 256>, filt_nt<256>()); break;
    case 512: go(filter_kernel<KIND, S, 512>, filt_nt<512>()); break;
    case 1024: go(filter_kernel<KIND, S, 1024>, filt_nt<1024>()); break;
    case 2048: go(filter_kernel<KIND, S, 2048>, filt_nt<2048>()); break;
    case 4096: go(filter_kernel<KIND, S, 4096>, filt_nt<4096>()); break;
    default: fail(GCT_ERR_UNSUPPORTED, "FFT length not a power of two in 8..4096 or 3 x one in 48..3072");
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
    switch (a.kind) {
    case GCT_PARALLEL_1D:
    case GCT_FAN_ARC_1D:
    case GCT_FAN_LINE_1D: launch_backproject_1d(a, s); break;
    case GCT_PARALLEL_2D:
    case GCT_CONE_PARALLEL: launch_backproject_parallel(a, s); break;
    case GCT_CYLINDRICAL: launch_backproject_cyl(a, s); break;
    default: launch_backproject_flat(a, s); break;
    }
    hip_check(hipGetLastError(), "backprojection");
}

void launch_forward_project(const FpLaunch &a, hipStream_t s)
{
    const size_t total = static_cast<size_t>(a.nviews) * (is_1d(a.kind) ? 1u : a.nv) * a.nu;
    if (total == 0) return;
    if (a.pad) {
        hipLaunchKernelGGL(pad_ymarch_kernel, dim3((a.nz + 3u) * a.ny), dim3(256), 0, s, a.vol, a.pad, a.nx, a.ny, a.nz);
        const uint32_t tx = (a.nx + 31u) / 32u, ty = (a.ny + 3u + 31u) / 32u;
        hipLaunchKernelGGL(pad_xmarch_kernel, dim3(tx * ty * (a.nz + 3u)), dim3(256), 0, s, a.vol,
                           a.pad + static_cast<size_t>(a.nz + 3u) * a.ny * (a.nx + 3u), a.nx, a.ny, a.nz, tx, ty);
        // GCT_JOSEPH=per_ray selects the one-ray-per-thread kernel instead (the
        // two give the same bits; the switch is for comparing them)
        static const bool per_ray = [] {
            const char *e = std::getenv("GCT_JOSEPH");
            return e && std::string(e) == "per_ray";
        }();
        if (!per_ray) {
            const uint32_t tiles_c = (a.nu + kJTC - 1u) / kJTC, tiles_r = (a.nv + kJTR - 1u) / kJTR;
            const uint64_t blocks = uint64_t(tiles_c) * tiles_r * ((a.nviews + kJVB - 1u) / kJVB);
            if (blocks >= (uint64_t(1) << 31)) fail(GCT_ERR_UNSUPPORTED, "forward projection grid too large");
            hipLaunchKernelGGL(joseph_lds_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, s, a, tiles_c,
                               tiles_r);
            hip_check(hipGetLastError(), "forward projection");
            return;
        }
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
    switch (g.kind) {
    case GCT_PARALLEL_1D:
    case GCT_FAN_ARC_1D:
    case GCT_FAN_LINE_1D: launch_batch_group_1d(scans, g, counter, grid, s); break;
    case GCT_PARALLEL_2D:
    case GCT_CONE_PARALLEL: launch_batch_group_parallel(scans, g, counter, grid, s); break;
    case GCT_CYLINDRICAL: launch_batch_group_cyl(scans, g, counter, grid, s); break;
    default: launch_batch_group_flat(scans, g, counter, grid, s); break;
    }
    hip_check(hipGetLastError(), "batched backprojection");
}

}  // namespace gct
