// Host geometry of the FDK path: per-view factories, rotation analysis,
// validity checks, ramp-kernel taps and the flat-panel projection matrix.
// Everything here is FP32 where the reference computes in FP32, so results
// match the reference's host code bit for bit where its source fixes the
// operation order.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numbers>
#include <sstream>

namespace gct {

namespace {

constexpr float kPiF = std::numbers::pi_v<float>;

// Utils::wrapFrom0To2PI (src/cpp/geom/utils.h:338-348).  remainder() of two
// floats is exact, so evaluating it in double returns the same float.
float wrap_0_2pi(float a)
{
    constexpr float pi_2 = 2.0f * kPiF;
    while (a < 0.0f) a += pi_2;
    return (a >= pi_2) ? static_cast<float>(std::remainder(static_cast<double>(a), static_cast<double>(pi_2))) : a;
}

// Utils::quadrantIndex(Radians) (utils.h:370-389).
unsigned quadrant(float ang)
{
    float a = wrap_0_2pi(ang);
    if (a >= 0.0f && a < kPiF / 2.0f) return 1u;
    if (a >= kPiF / 2.0f && a < kPiF) return 2u;
    if (a >= kPiF && a < (3.0f / 2.0f) * kPiF) return 3u;
    return 4u;
}

}  // namespace

uint32_t next_pow2(uint32_t n)
{
    uint32_t v = 1u;
    while (v < n) v *= 2u;
    return v;
}

uint32_t fft_len(uint32_t n)
{
    const uint32_t need = std::max(8u, n + n / 2u);
    const uint32_t p2 = next_pow2(need);
    const uint32_t p3 = 3u * next_pow2((need + 2u) / 3u);   // 3 x 2^k >= need
    return (p3 >= 48u && p3 < p2) ? p3 : p2;
}

RotationInfo rotation_info(const gct_ct_geom &g)
{
    RotationInfo r;
    const uint32_t n = g.num_proj;
    r.rot_ang.resize(n);
    // Detail::rotAng (rotation_geom.h:32-36; rotation_geom.cpp:255-295)
    for (uint32_t v = 0; v < n; ++v) {
        const gct_proj_geom &p = g.proj_geom[v];
        r.rot_ang[v] = is_parallel_family(g.kind) ? wrap_0_2pi(p.rot_ang)
                                                  : wrap_0_2pi(std::atan2(p.src_pos[1], p.src_pos[0]));
    }
    // Detail::rotDir (rotation_geom.cpp:7-43)
    if (n <= 1u) fail(GCT_ERR_RUNTIME, "Invalid X-ray source trajectory: fewer than two views.");
    {
        float a = wrap_0_2pi(r.rot_ang[0]), b = wrap_0_2pi(r.rot_ang[1]);
        unsigned qa = quadrant(a), qb = quadrant(b);
        if (qa == 1u && qb == 4u) r.dir = -1;
        else if (qa == 4u && qb == 1u) r.dir = 1;
        else if (b - a > 0.0f) r.dir = 1;
        else if (b - a < 0.0f) r.dir = -1;
        else fail(GCT_ERR_RUNTIME, "Invalid X-ray source trajectory: stationary or ambiguous rotation.");
    }
    // Detail::relRotAngles (rotation_geom.cpp:45-74)
    r.rel_ang.resize(n);
    const float first = r.rot_ang[0];
    for (uint32_t v = 0; v < n; ++v)
        r.rel_ang[v] = wrap_0_2pi(r.dir < 0 ? first - r.rot_ang[v] : r.rot_ang[v] - first);
    // Detail::rotAngIncr: mean of wrapped increments, accumulated in order (:76-91)
    float sum = 0.0f;
    for (uint32_t v = 0; v + 1 < n; ++v) sum = sum + wrap_0_2pi(r.rel_ang[v + 1] - r.rel_ang[v]);
    r.incr = sum / static_cast<float>(n - 1);
    // Detail::rotAngRange (:93-104)
    r.range = r.incr + *std::max_element(r.rel_ang.begin(), r.rel_ang.end());
    // Detail::numProjPerRot (:106-134); rel[0] is always 0 here
    r.nppr = n;
    for (uint32_t v = 1; v < n; ++v) {
        float a = r.rel_ang[v];
        if (std::fabs(a) < 0.000001f) a = 0.0f;
        if (a == 0.0f) { r.nppr = v; break; }
    }
    return r;
}

void validate_geom(const gct_ct_geom &g)
{
    if (g.kind < GCT_PARALLEL_1D || g.kind > GCT_FLAT) fail(GCT_ERR_INVALID_ARGUMENT, "unknown geometry kind");
    if (g.det_row.num_samples == 0 || !(g.det_row.sampling_width > 0.0f))
        fail(GCT_ERR_INVALID_ARGUMENT, "detector row sampler must have samples and a positive width");
    if (!is_1d(g.kind) && (g.det_col.num_samples == 0 || !(g.det_col.sampling_width > 0.0f)))
        fail(GCT_ERR_INVALID_ARGUMENT, "detector column sampler must have samples and a positive width");
    if (g.num_proj == 0 || g.proj_geom == nullptr) fail(GCT_ERR_INVALID_ARGUMENT, "no projection geometry");
    if (g.det_row.num_samples > 4096u) fail(GCT_ERR_UNSUPPORTED, "more than 4096 detector columns");
    if (!is_1d(g.kind) && g.det_col.num_samples > 65535u) fail(GCT_ERR_UNSUPPORTED, "more than 65535 detector rows");
    for (uint32_t v = 0; v < g.num_proj; ++v) {
        const gct_proj_geom &p = g.proj_geom[v];
        bool needs_sdd = !(g.kind == GCT_PARALLEL_1D || g.kind == GCT_PARALLEL_2D);
        if (needs_sdd && !(p.src_det_dist > 0.0f)) fail(GCT_ERR_INVALID_ARGUMENT, "source-detector distance must be positive");
        if (g.kind == GCT_CONE_PARALLEL && !(p.src_iso_dist > 0.0f))
            fail(GCT_ERR_INVALID_ARGUMENT, "source-isocenter distance must be positive");
    }
}

// ReconImgGeom::isSamplingValid (recon_img_geom.h:57-70) and the image
// dimension each geometry reconstructs (projDim/reconImgDim, proj_meta_info.h).
void validate_image(const gct_ct_geom &g, const gct_image_geom &img)
{
    uint32_t want = is_1d(g.kind) ? 2u : 3u;
    if (img.dim != want) fail(GCT_ERR_INVALID_ARGUMENT, "image dimension does not match the geometry");
    if (img.x.num_samples == 0 || img.y.num_samples == 0 || !(img.x.sampling_width > 0.0f) ||
        !(img.y.sampling_width > 0.0f))
        fail(GCT_ERR_INVALID_ARGUMENT, "invalid image sampling");
    if (img.dim == 3 && (img.z.num_samples == 0 || !(img.z.sampling_width > 0.0f)))
        fail(GCT_ERR_INVALID_ARGUMENT, "invalid image sampling");
    uint64_t nvox = uint64_t(img.x.num_samples) * img.y.num_samples * (img.dim == 3 ? img.z.num_samples : 1u);
    if (nvox > (uint64_t(1) << 34)) fail(GCT_ERR_UNSUPPORTED, "image larger than 2^34 voxels");
}

// RampFilter::kernel (ramp_filter.h:29-56) and projFiltKernel
// (proj_filt_kernel.cpp:5-45).  Taps are computed in FP64 and stored FP32 as
// the reference does; the FP64 values are returned too (the arc weighting
// multiplies the FP32 tap in FP64, proj_filt_kernel.cpp:32-37).
std::vector<double> ramp_taps_f64(int kind, uint32_t n, float width, float cutoff, int window,
                                  std::vector<float> *taps32)
{
    if (window < GCT_WINDOW_NONE || window > GCT_WINDOW_SHEPP_LOGAN) fail(GCT_ERR_INVALID_ARGUMENT, "unknown window");
    if (!(cutoff > 0.0f)) fail(GCT_ERR_INVALID_ARGUMENT, "cut-off frequency must be positive");
    if (window == GCT_WINDOW_SHEPP_LOGAN && cutoff != 1.0f)
        fail(GCT_ERR_UNSUPPORTED, "the Shepp-Logan kernel is defined for cut-off 1 only");
    const double a = (window == GCT_WINDOW_HANN) ? 0.5 : (window == GCT_WINDOW_HAMMING) ? 0.54 : 1.0;
    const double c = cutoff, sw = width, pi = std::numbers::pi;
    auto M = [](double t) { return (t == 0.0) ? 0.5 : (std::sin(t) / t + (std::cos(t) - 1.0) / (t * t)); };
    const int half = static_cast<int>(n) / 2;
    std::vector<double> out(n);
    if (taps32) taps32->resize(n);
    for (uint32_t k = 0; k < n; ++k) {
        const int m = static_cast<int>(k) - half;
        const double nd = m;
        double h;
        if (window == GCT_WINDOW_SHEPP_LOGAN)
            h = -2.0 / (pi * pi * sw * (4.0 * nd * nd - 1.0));   // Shepp & Logan (1974)
        else
            h = ((c * c) / (2.0 * sw)) * (a * M(c * pi * nd) + ((1 - a) / 2.0) * M(pi * c * nd + pi) +
                                          ((1 - a) / 2.0) * M(pi * c * nd - pi));
        const float hf = static_cast<float>(h);
        if (is_arc_row(kind) && m != 0) {
            const double beta = nd * sw;
            const double wgt = std::pow(beta / std::sin(beta), 2);
            out[k] = wgt * h;
            if (taps32) (*taps32)[k] = static_cast<float>(wgt * static_cast<double>(hf));
        } else {
            out[k] = h;
            if (taps32) (*taps32)[k] = hf;
        }
    }
    return out;
}

}  // namespace gct

using namespace gct;

// ---- C ABI: geometry ------------------------------------------------------

extern "C" gct_status gct_geom_create_circular(const gct_circular_params *p, gct_proj_geom *out)
{
    return guarded([&] {
        if (!p || !out) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (p->kind < GCT_PARALLEL_1D || p->kind > GCT_FLAT) fail(GCT_ERR_INVALID_ARGUMENT, "unknown geometry kind");
        if (p->dir != 1 && p->dir != -1) fail(GCT_ERR_INVALID_ARGUMENT, "rotation direction must be +1 (CW) or -1 (CCW)");
        for (uint32_t i = 0; i < p->num_proj; ++i) {
            gct_proj_geom g;
            std::memset(&g, 0, sizeof(g));
            // rot_ang = init + incr * i in FP32 (proj_geom_factory_params.cpp:150-151, cb_factory.cpp:9-10)
            const float rot = p->init_ang + p->ang_incr * static_cast<float>(i);
            g.rot_ang = rot;
            g.det_row_center = p->det_row_center;
            g.det_col_center = is_1d(p->kind) ? p->det_row_center : p->det_col_center;
            g.src_det_dist = p->src_det_dist;
            g.src_iso_dist = p->src_iso_dist;
            if (!is_parallel_family(p->kind)) {
                // ProjGeomCBFactory::create (proj_geom_cb_factory.cpp:14-48): Rz(dir*theta)
                // as Eigen's AngleAxisf builds it, basis rows = implicit * Rz^T,
                // src = Rz * (SID, 0, -pitch*theta).
                const float phi = static_cast<float>(p->dir) * rot;
                const float c = std::cos(phi), s = std::sin(phi);
                const float rz22 = (1.0f - c) + c;
                const float sid = p->src_iso_dist;
                const float zs = -p->hel_pitch * rot;
                g.det_row_dir[0] = -s; g.det_row_dir[1] = c; g.det_row_dir[2] = 0.0f;
                g.det_col_dir[0] = -0.0f; g.det_col_dir[1] = -0.0f; g.det_col_dir[2] = -rz22;
                g.iso_ray_dir[0] = -c; g.iso_ray_dir[1] = -s; g.iso_ray_dir[2] = -0.0f;
                g.src_pos[0] = c * sid + (-s) * 0.0f + 0.0f * zs;
                g.src_pos[1] = s * sid + c * 0.0f + 0.0f * zs;
                g.src_pos[2] = 0.0f * sid + 0.0f * 0.0f + rz22 * zs;
                if (is_1d(p->kind)) g.src_pos[2] = 0.0f;
            }
            out[i] = g;
        }
    });
}

extern "C" gct_status gct_rotation_geom_compute(const gct_ct_geom *geom, gct_rotation_geom *out, float *rot_ang,
                                        float *rel_rot_ang)
{
    return guarded([&] {
        if (!geom || !geom->proj_geom || geom->num_proj == 0) fail(GCT_ERR_INVALID_ARGUMENT, "empty geometry");
        RotationInfo r = rotation_info(*geom);
        if (out) {
            out->dir = r.dir;
            out->rot_ang_incr = r.incr;
            out->rot_ang_range = r.range;
            out->num_proj_per_rot = r.nppr;
        }
        if (rot_ang) std::copy(r.rot_ang.begin(), r.rot_ang.end(), rot_ang);
        if (rel_rot_ang) std::copy(r.rel_ang.begin(), r.rel_ang.end(), rel_rot_ang);
    });
}

extern "C" gct_status gct_check_axial_recon(const gct_ct_geom *geom, uint32_t cols, uint32_t rows, uint32_t nproj)
{
    return guarded([&] {
        if (!geom) fail(GCT_ERR_INVALID_ARGUMENT, "null geometry");
        // checkProjDetConsistency (check_validity.h:25-64)
        if (cols != geom->det_row.num_samples) {
            std::ostringstream s;
            s << "Inconsistent projection and detector geometry data. Number of columns of projection image:: "
              << cols << " Number of detector columns:: " << geom->det_row.num_samples;
            fail(GCT_ERR_INVALID_ARGUMENT, s.str());
        }
        if (!is_1d(geom->kind) && rows != geom->det_col.num_samples) {
            std::ostringstream s;
            s << "Inconsistent projection and detector geometry data. Number of rows of projection image:: "
              << rows << " Number of detector rows:: " << geom->det_col.num_samples;
            fail(GCT_ERR_INVALID_ARGUMENT, s.str());
        }
        // checkProjViewConsistency (check_validity.h:77-107)
        if (nproj != geom->num_proj) {
            fail(GCT_ERR_INVALID_ARGUMENT, nproj > geom->num_proj
                                               ? "Projection geometry parameters are not available for all projection images."
                                               : "Projection images are not available for all projection geometries.");
        }
    });
}

extern "C" gct_status gct_ramp_kernel(const gct_ct_geom *geom, float cutoff, int32_t window, float *taps)
{
    return guarded([&] {
        if (!geom || !taps) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        std::vector<float> t32;
        ramp_taps_f64(geom->kind, geom->det_row.num_samples, geom->det_row.sampling_width, cutoff, window, &t32);
        std::copy(t32.begin(), t32.end(), taps);
    });
}

// ProjectionMatrix(pg, du, dv) (projection_matrix.cpp:66-96), evaluated in FP64.
extern "C" gct_status gct_projection_matrix(const gct_proj_geom *pg, float du, float dv, double *P)
{
    return guarded([&] {
        if (!pg || !P || !(du > 0.0f) || !(dv > 0.0f)) fail(GCT_ERR_INVALID_ARGUMENT, "invalid argument");
        const double D = pg->src_det_dist;
        const double K[3][3] = {{D / du, 0.0, pg->det_row_center}, {0.0, D / dv, pg->det_col_center}, {0, 0, 1}};
        const double eu[3] = {pg->det_row_dir[0], pg->det_row_dir[1], pg->det_row_dir[2]};
        const double ev[3] = {pg->det_col_dir[0], pg->det_col_dir[1], pg->det_col_dir[2]};
        const double ew[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0]};
        const double R[3][3] = {{eu[0], eu[1], eu[2]}, {ev[0], ev[1], ev[2]}, {ew[0], ew[1], ew[2]}};
        const double s[3] = {pg->src_pos[0], pg->src_pos[1], pg->src_pos[2]};
        double E[3][4];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) E[i][j] = R[i][j];
            E[i][3] = -(R[i][0] * s[0] + R[i][1] * s[1] + R[i][2] * s[2]);
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 4; ++j) P[i * 4 + j] = K[i][0] * E[0][j] + K[i][1] * E[1][j] + K[i][2] * E[2][j];
        const double n = P[11];
        if (n == 0.0) fail(GCT_ERR_INVALID_ARGUMENT, "source lies in the detector's principal plane");
        for (int i = 0; i < 12; ++i) P[i] /= n;
    });
}

// ProjectionMatrix(mat) decomposition (projection_matrix.cpp:4-65) as an RQ
// factorisation M = K R (K upper triangular with positive diagonal, R
// orthonormal) by Gram-Schmidt on M's rows from the bottom, and the flat view
// of ProjGeomFactory<DetFanLine,DetFanLine>::create(ProjectionMatrix)
// (proj_geom_factory_proj_mat.cpp:7-26) with the focal length scaled by the
// pixel width (the TODO at :13).
extern "C" gct_status gct_projection_matrix_decompose(const double *P, float pix_width, gct_proj_geom *out,
                                                      double *focal_uv)
{
    return guarded([&] {
        if (!P || !out) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        double m[3][3], p4[3];
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) m[i][j] = P[i * 4 + j];
            p4[i] = P[i * 4 + 3];
        }
        auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        auto nrm = [&](const double *a) { return std::sqrt(dot(a, a)); };
        double r[3][3], K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        K[2][2] = nrm(m[2]);
        if (!(K[2][2] > 0.0)) fail(GCT_ERR_INVALID_ARGUMENT, "degenerate projection matrix");
        for (int j = 0; j < 3; ++j) r[2][j] = m[2][j] / K[2][2];
        K[1][2] = dot(m[1], r[2]);
        double t[3];
        for (int j = 0; j < 3; ++j) t[j] = m[1][j] - K[1][2] * r[2][j];
        K[1][1] = nrm(t);
        if (!(K[1][1] > 0.0)) fail(GCT_ERR_INVALID_ARGUMENT, "degenerate projection matrix");
        for (int j = 0; j < 3; ++j) r[1][j] = t[j] / K[1][1];
        K[0][2] = dot(m[0], r[2]);
        K[0][1] = dot(m[0], r[1]);
        for (int j = 0; j < 3; ++j) t[j] = m[0][j] - K[0][1] * r[1][j] - K[0][2] * r[2][j];
        K[0][0] = nrm(t);
        if (!(K[0][0] > 0.0)) fail(GCT_ERR_INVALID_ARGUMENT, "degenerate projection matrix");
        for (int j = 0; j < 3; ++j) r[0][j] = t[j] / K[0][0];
        // source = -M^{-1} p4 (cofactor inverse)
        double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                     m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
        if (det == 0.0) fail(GCT_ERR_INVALID_ARGUMENT, "singular projection matrix");
        double inv[3][3];
        inv[0][0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) / det;
        inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) / det;
        inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) / det;
        inv[1][0] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) / det;
        inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) / det;
        inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) / det;
        inv[2][0] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) / det;
        inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) / det;
        inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) / det;
        gct_proj_geom g;
        std::memset(&g, 0, sizeof(g));
        for (int i = 0; i < 3; ++i) {
            g.src_pos[i] = static_cast<float>(-(inv[i][0] * p4[0] + inv[i][1] * p4[1] + inv[i][2] * p4[2]));
            g.det_row_dir[i] = static_cast<float>(r[0][i]);
            g.det_col_dir[i] = static_cast<float>(r[1][i]);
            g.iso_ray_dir[i] = static_cast<float>(r[2][i]);
        }
        const double fu = K[0][0] / K[2][2], fv = K[1][1] / K[2][2];
        g.det_row_center = static_cast<float>(K[0][2] / K[2][2]);
        g.det_col_center = static_cast<float>(K[1][2] / K[2][2]);
        g.src_det_dist = static_cast<float>(fu * static_cast<double>(pix_width));
        g.src_iso_dist = std::sqrt(g.src_pos[0] * g.src_pos[0] + g.src_pos[1] * g.src_pos[1] + g.src_pos[2] * g.src_pos[2]);
        *out = g;
        if (focal_uv) { focal_uv[0] = fu; focal_uv[1] = fv; }
    });
}

namespace gct {

// SpiralCTGeom (spiral_ct_geom.h:30-95) and the slice loop of the spiral FDK
// (fdk_spiral.cuh:53-70), in FP32 as the reference: slice z sits at
// leftEndPointPos(mz, nz, dz) + z dz; its centre view is the first one with
// the smallest |source z - slice z|; the window [centre - nppr/2, + nppr) in
// unsigned arithmetic is used only when first > 0 and last < V, otherwise the
// slice is left empty (zero).
std::vector<uint32_t> spiral_windows(const gct_ct_geom &g, const gct_image_geom &img)
{
    if (img.dim != 3) fail(GCT_ERR_INVALID_ARGUMENT, "spiral reconstruction needs a 3-D image");
    if (is_1d(g.kind)) fail(GCT_ERR_INVALID_ARGUMENT, "spiral reconstruction needs a 2-D detector");
    const RotationInfo r = rotation_info(g);
    const uint32_t V = g.num_proj;
    if (r.nppr == 0)
        fail(GCT_ERR_RUNTIME, "Invalid spiral CT geometry: number of projections per rotation should not be zero.");
    bool asc = true, desc = true;
    for (uint32_t v = 1; v < V; ++v) {
        if (g.proj_geom[v].src_pos[2] < g.proj_geom[v - 1].src_pos[2]) asc = false;
        if (g.proj_geom[v].src_pos[2] > g.proj_geom[v - 1].src_pos[2]) desc = false;
    }
    if (!asc && !desc)
        fail(GCT_ERR_RUNTIME, "Invalid spiral CT geometry: the z-coordinate of source position should be sorted "
                              "either in ascending  or descending order.");
    if (r.nppr >= V) fail(GCT_ERR_RUNTIME, "Invalid spiral CT geometry: fewer views than one rotation.");
    const float pitch = std::fabs(g.proj_geom[0].src_pos[2] - g.proj_geom[r.nppr].src_pos[2]) /
                        (2.0f * std::numbers::pi_v<float>);
    if (!(pitch > 0.0f)) fail(GCT_ERR_RUNTIME, "Invalid spiral CT geometry: spiral pitch cannot be zero or negative.");
    const uint32_t nz = img.z.num_samples;
    const float dz = img.z.sampling_width;
    const float first_pos = img.mid_point[2] - ((static_cast<float>(nz) - 1.0f) * dz / 2.0f);
    std::vector<uint32_t> win(2u * nz, 0u);
    for (uint32_t z = 0; z < nz; ++z) {
        const float sp = first_pos + static_cast<float>(z) * dz;
        uint32_t mid = 0;
        float best = std::fabs(g.proj_geom[0].src_pos[2] - sp);
        for (uint32_t v = 1; v < V; ++v) {
            const float d = std::fabs(g.proj_geom[v].src_pos[2] - sp);
            if (d < best) { best = d; mid = v; }
        }
        const uint32_t first = mid - r.nppr / 2u;
        const uint32_t last = first + r.nppr;
        if (first > 0u && last < V) { win[2u * z] = first; win[2u * z + 1u] = last; }
    }
    return win;
}

}  // namespace gct

extern "C" gct_status gct_spiral_windows(const gct_ct_geom *geom, const gct_image_geom *img, uint32_t *win)
{
    return guarded([&] {
        if (!geom || !img || !win) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        validate_geom(*geom);
        validate_image(*geom, *img);
        const std::vector<uint32_t> w = spiral_windows(*geom, *img);
        std::copy(w.begin(), w.end(), win);
    });
}
