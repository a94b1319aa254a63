// Host orchestration of the FDK path: plans (per-view tables, filter
// spectrum, device buffers), the fused filter and backprojection stages, the
// one-shot reconstruct with its high/low-memory paths, the autotuner and
// batches captured in HIP graphs.
//
// Reference: Recon::CUDA::reconFDK and Detail::reconFDKHighMem/LowMem
// (src/cuda/recon/fdk_axial.cuh:53-205), CTGeomDev (src/cuda/geom/ct_geom_dev.cu),
// RampFilteringFFT (src/cuda/recon/ramp_filtering_fft.cuh:79-111),
// getParkerWeights (src/cuda/recon/parker_weighting.cuh:132-200).
#include "plan_internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numbers>
#include <tuple>
#include <tuple>

namespace gct {

size_t storage_bytes(int storage) { return storage == GCT_STORE_F32 ? 4u : 2u; }

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

namespace {

// ---- device memory --------------------------------------------------------

// ---- host FFT (FP64) for the filter spectrum ------------------------------

void fft64(std::vector<std::complex<double>> &a);

// FFT of length 2^k or 3 x 2^k in FP64: three interleaved length-n/3
// transforms joined by one radix-3 stage, X_k = sum_r W_n^(r k) S_r[k mod n/3].
void fft64_any(std::vector<std::complex<double>> &a)
{
    const size_t n = a.size();
    if ((n & (n - 1)) == 0) {
        fft64(a);
        return;
    }
    const size_t m = n / 3;
    std::vector<std::complex<double>> s[3];
    for (size_t r = 0; r < 3; ++r) {
        s[r].resize(m);
        for (size_t i = 0; i < m; ++i) s[r][i] = a[3 * i + r];
        fft64(s[r]);
    }
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc = s[0][k % m];
        for (size_t r = 1; r < 3; ++r)
            acc += s[r][k % m] * std::polar(1.0, -2.0 * std::numbers::pi * static_cast<double>((r * k) % n) /
                                                     static_cast<double>(n));
        a[k] = acc;
    }
}

void fft64(std::vector<std::complex<double>> &a)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * std::numbers::pi / static_cast<double>(len);
        for (size_t i = 0; i < n; i += len)
            for (size_t j = 0; j < len / 2; ++j) {
                const std::complex<double> w(std::cos(ang * static_cast<double>(j)), std::sin(ang * static_cast<double>(j)));
                const std::complex<double> u = a[i + j], v = a[i + j + len / 2] * w;
                a[i + j] = u + v;
                a[i + j + len / 2] = u - v;
            }
    }
}

struct Tuning {
    uint32_t z_unroll;
    int variant;
    uint32_t view_block, lds_rows, lds_pairs;
    bool round_robin;
};
std::mutex g_tune_mu;
std::map<std::tuple<int, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, int, int, int>, Tuning> g_tune;

}  // namespace

}  // namespace gct

using namespace gct;

namespace gct {
namespace {

void build_records(gct_plan &p, std::vector<BpRec> &rec, std::vector<BpRec64> &rec64, std::vector<FiltRec> &frec)
{
    const gct_ct_geom &g = p.geom;
    const gct_image_geom &im = p.img;
    const double x0 = double(im.mid_point[0]) - (double(im.x.num_samples) - 1.0) * double(im.x.sampling_width) / 2.0;
    const double y0 = double(im.mid_point[1]) - (double(im.y.num_samples) - 1.0) * double(im.y.sampling_width) / 2.0;
    const double z0 = (im.dim == 3) ? double(im.mid_point[2]) - (double(im.z.num_samples) - 1.0) * double(im.z.sampling_width) / 2.0 : 0.0;
    const double dx = im.x.sampling_width, dy = im.y.sampling_width, dz = (im.dim == 3) ? im.z.sampling_width : 0.0;
    const double du = g.det_row.sampling_width, dv = is_1d(g.kind) ? 1.0 : g.det_col.sampling_width;
    rec.assign(p.V, BpRec{});
    rec64.assign(p.V, BpRec64{});
    frec.assign(p.V, FiltRec{});
    p.circular = true;
    // affine form (c0 + c1 ix + c2 iy + c3 iz) of (position - src) . e, scaled by f
    auto affine = [&](const double s[3], const double e[3], double f, double out[4]) {
        out[0] = ((x0 - s[0]) * e[0] + (y0 - s[1]) * e[1] + (z0 - s[2]) * e[2]) * f;
        out[1] = dx * e[0] * f;
        out[2] = dy * e[1] * f;
        out[3] = dz * e[2] * f;
    };
    for (uint32_t v = 0; v < p.V; ++v) {
        const gct_proj_geom &pg = p.views[v];
        BpRec64 r{};
        // The records give padded detector coordinates (column + 1, row + 1):
        // a filtered view carries one zero column and row before its first, and
        // the backprojection addresses its taps from there.  Adding the 1 to the
        // centre here, in FP64, keeps it from costing a rounding per column.
        const double cu = double(pg.det_row_center) + 1.0, cv = double(pg.det_col_center) + 1.0;
        const double D = pg.src_det_dist;
        double s[3] = {pg.src_pos[0], pg.src_pos[1], is_1d(g.kind) ? 0.0 : pg.src_pos[2]};
        double eu[3] = {pg.det_row_dir[0], pg.det_row_dir[1], is_1d(g.kind) ? 0.0 : pg.det_row_dir[2]};
        double ev[3] = {pg.det_col_dir[0], pg.det_col_dir[1], pg.det_col_dir[2]};
        double ew[3] = {pg.iso_ray_dir[0], pg.iso_ray_dir[1], is_1d(g.kind) ? 0.0 : pg.iso_ray_dir[2]};
        frec[v] = FiltRec{pg.det_row_center, pg.det_col_center, pg.src_det_dist, 0.0f};
        switch (g.kind) {
        case GCT_PARALLEL_1D:
        case GCT_PARALLEL_2D:
        case GCT_CONE_PARALLEL: {
            // t = x cos(theta + pi/2) + y sin(theta + pi/2) (coord_trans_proj.cuh:50-59);
            // col = t / du + cu; 2-D rows: -z (:149) or -z D / SID (:168), row = . / dv + cv
            const double th = double(pg.rot_ang) + std::numbers::pi / 2.0;
            const double c = std::cos(th), sn = std::sin(th);
            r.u[0] = cu + (x0 * c + y0 * sn) / du;
            r.u[1] = dx * c / du;
            r.u[2] = dy * sn / du;
            const double m = (g.kind == GCT_CONE_PARALLEL) ? D / double(pg.src_iso_dist) : 1.0;
            r.v[0] = cv - z0 * m / dv;
            r.v[3] = -dz * m / dv;
            break;
        }
        case GCT_FAN_LINE_1D:
        case GCT_FLAT: {
            // u = D (r.eu)/(r.ew), v = D (r.ev)/(r.ew), w = D SID/(r.ew)^2
            // (coord_trans_proj.cuh:108-134, 211-238): every form is divided by
            // sqrt(D SID) so that 1/(W.X)^2 is the weight directly.
            if (g.kind == GCT_FAN_LINE_1D) { ew[0] = -eu[1]; ew[1] = eu[0]; ew[2] = 0.0; }   // rot90 (ct_geom_dev.cu:154-156)
            else {
                ew[0] = eu[1] * ev[2] - eu[2] * ev[1];   // eu x ev (ct_geom_dev.cu:407-411)
                ew[1] = eu[2] * ev[0] - eu[0] * ev[2];
                ew[2] = eu[0] * ev[1] - eu[1] * ev[0];
            }
            const double sid = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
            const double k = 1.0 / std::sqrt(D * sid);
            affine(s, ew, k, r.w);
            affine(s, eu, D / du * k, r.u);
            if (g.kind == GCT_FLAT) affine(s, ev, D / dv * k, r.v);
            r.k[0] = cu;
            r.k[1] = cv;
            break;
        }
        case GCT_FAN_ARC_1D:
        case GCT_CYLINDRICAL: {
            // gamma = atan2(r.eu, r.ew); v = D (r^.ev)/sqrt(1 - (r^.ev)^2) = D (r.ev)/rho
            // for an orthonormal basis; w = SID / |r|^2 (coord_trans_proj.cuh:80-105, 175-208).
            // gamma and v are unchanged when the three forms are scaled together, so
            // they are scaled by 1/sqrt(SID): the weight is then 1 / |r|^2 directly.
            const double sid = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
            const double c = 1.0 / std::sqrt(sid);
            affine(s, ew, c, r.w);
            affine(s, eu, c, r.u);
            if (g.kind == GCT_CYLINDRICAL) affine(s, ev, c, r.v);
            r.k[0] = cu;
            r.k[1] = cv;
            r.k[2] = D / dv;
            r.k[3] = 1.0;
            break;
        }
        }
        rec64[v] = r;
        BpRec &f = rec[v];
        for (int i = 0; i < 4; ++i) {
            f.u[i] = static_cast<float>(r.u[i]);
            f.v[i] = static_cast<float>(r.v[i]);
            f.w[i] = static_cast<float>(r.w[i]);
            f.k[i] = static_cast<float>(r.k[i]);
        }
        if (g.kind == GCT_FLAT || g.kind == GCT_CYLINDRICAL)
            if (f.u[3] != 0.0f || f.w[3] != 0.0f) p.circular = false;
    }
}

// True when, for every view, every voxel of the volume -- z padded to whole
// 32-slice tiles, the slices an LDS tile computes past the last -- projects
// strictly between padded rows 1 and nv + 1, with a margin for the FP32 set-up.
// Then the LDS kernel's staged box of a tile is never cut by the detector's
// edge and contains the rows of all the tile's voxels (the box spans the rows
// of the tile's corners with a margin of one row): its row clamp can never
// bind, and the kernel is built without it.  In FP64 from the FP64 records:
//   flat: row = nv / nw + k1 is linear-fractional; when nw has one sign over
//     the volume, its extremes over any box are at corners;
//   parallel kinds: the row is affine in z;
//   cylindrical: row = k1 + k2 nv / rho with rho = |(nu, nw)|: bounded from
//     nv's corner range and rho's range over the volume's footprint; a tile's
//     corners can miss rho's minimum d over the tile (inside an edge) by at
//     most sqrt(d^2 + h^2) - d <= h^2 / (2 d) (h: half a tile edge), which moves a
//     row by at most |k2 nv| h^2 / (2 rho_min^3); twice that must stay below
//     half a row, and it widens the bounds.
bool rows_inside_detector(const gct_plan &p, const std::vector<BpRec64> &rec64)
{
    const gct_ct_geom &g = p.geom;
    if (!p.circular || p.img.dim != 3 || is_1d(g.kind)) return false;
    const double xm = double(p.img.x.num_samples) - 1.0, ym = double(p.img.y.num_samples) - 1.0;
    const double zm = double((p.img.z.num_samples + 31u) / 32u * 32u) - 1.0;
    const double lo = 1.05, hi = double(p.nv) + 0.95;
    const double X[2] = {0.0, xm}, Y[2] = {0.0, ym}, Z[2] = {0.0, zm};
    for (const BpRec64 &r : rec64) {
        auto form = [](const double f[4], double x, double y, double z) { return f[0] + f[1] * x + f[2] * y + f[3] * z; };
        double rmin = 1e300, rmax = -1e300;
        if (g.kind == GCT_PARALLEL_2D || g.kind == GCT_CONE_PARALLEL) {
            for (double z : Z) {
                const double row = r.v[0] + r.v[3] * z;
                rmin = std::min(rmin, row);
                rmax = std::max(rmax, row);
            }
        } else if (g.kind == GCT_FLAT) {
            double wmin = 1e300, wmax = -1e300;
            for (double x : X)
                for (double y : Y)
                    for (double z : Z) {
                        const double nw = form(r.w, x, y, z);
                        wmin = std::min(wmin, nw);
                        wmax = std::max(wmax, nw);
                        const double row = form(r.v, x, y, z) / nw + r.k[1];
                        rmin = std::min(rmin, row);
                        rmax = std::max(rmax, row);
                    }
            if (!(wmin > 0.0 || wmax < 0.0)) return false;
        } else if (g.kind == GCT_CYLINDRICAL) {
            // (nu, nw) over the footprint: a parallelogram; rho's range over it
            double pu[4], pw[4], nvmin = 1e300, nvmax = -1e300;
            for (int c = 0; c < 4; ++c) {
                const double x = X[c & 1], y = Y[c >> 1];
                pu[c] = form(r.u, x, y, 0.0);
                pw[c] = form(r.w, x, y, 0.0);
            }
            for (double x : X)
                for (double y : Y)
                    for (double z : Z) {
                        const double nv = form(r.v, x, y, z);
                        nvmin = std::min(nvmin, nv);
                        nvmax = std::max(nvmax, nv);
                    }
            double rho_max = 0.0, rho_min = 1e300;
            const int edge[4][2] = {{0, 1}, {1, 3}, {3, 2}, {2, 0}};
            double cross_sign = 0.0;
            bool inside = true;
            for (const auto &e : edge) {
                const double ax = pu[e[0]], ay = pw[e[0]], bx = pu[e[1]], by = pw[e[1]];
                rho_max = std::max(rho_max, std::hypot(ax, ay));
                const double ex = bx - ax, ey = by - ay, len2 = ex * ex + ey * ey;
                const double t = len2 > 0.0 ? std::clamp(-(ax * ex + ay * ey) / len2, 0.0, 1.0) : 0.0;
                rho_min = std::min(rho_min, std::hypot(ax + t * ex, ay + t * ey));
                const double cr = ax * ey - ay * ex;   // origin on the same side of every edge: inside
                if (cross_sign == 0.0) cross_sign = cr;
                else if (cr * cross_sign <= 0.0) inside = false;
            }
            if (inside || !(rho_min > 0.0)) return false;
            const double k2 = r.k[2];
            const double q[4] = {nvmin / rho_min, nvmin / rho_max, nvmax / rho_min, nvmax / rho_max};
            for (double v : q) {
                rmin = std::min(rmin, r.k[1] + k2 * v);
                rmax = std::max(rmax, r.k[1] + k2 * v);
            }
            // a tile edge in the forms' units: 15 voxels times the longer step
            const double step = std::max(std::hypot(r.u[1], r.w[1]), std::hypot(r.u[2], r.w[2]));
            const double h = 7.5 * step;
            const double dev = std::fabs(k2) * std::max(std::fabs(nvmin), std::fabs(nvmax)) * h * h /
                               (rho_min * rho_min * rho_min);
            if (!(dev < 0.5)) return false;
            rmin -= dev;
            rmax += dev;
        } else {
            return false;
        }
        if (!(rmin >= lo && rmax <= hi)) return false;
    }
    return true;
}

// Whether the LDS kernel may leave the row clamp out for a view whose staged
// box the detector's edges did not cut (Box::inner).  Such a box spans the rows
// of its tile's corners with at least one row of margin on each side; flat and
// parallel rows take their extremes over a tile at its corners (linear-
// fractional with a denominator of one sign over the volume, checked here, or
// affine), and cylindrical rows k1 + k2 nv / rho can exceed their corners'
// range inside a tile by at most |k2 nv| h^2 / (2 rho_min^3), h half a tile
// edge (rows_inside_detector).  That bound, over the whole volume and every
// view, must stay below half a row.
bool box_inner_ok(const gct_plan &p, const std::vector<BpRec64> &rec64)
{
    const gct_ct_geom &g = p.geom;
    if (g.kind != GCT_CYLINDRICAL && g.kind != GCT_FLAT) return true;
    if (!p.circular || p.img.dim != 3) return false;
    const double xm = double(p.img.x.num_samples) - 1.0, ym = double(p.img.y.num_samples) - 1.0;
    const double zm = double((p.img.z.num_samples + 31u) / 32u * 32u) - 1.0;
    const double X[2] = {0.0, xm}, Y[2] = {0.0, ym}, Z[2] = {0.0, zm};
    for (const BpRec64 &r : rec64) {
        auto form = [](const double f[4], double x, double y, double z) { return f[0] + f[1] * x + f[2] * y + f[3] * z; };
        if (g.kind == GCT_FLAT) {
            // the denominator of the flat rows keeps one sign over the volume
            double wmin = 1e300, wmax = -1e300;
            for (double x : X)
                for (double y : Y)
                    for (double z : Z) {
                        const double nw = form(r.w, x, y, z);
                        wmin = std::min(wmin, nw);
                        wmax = std::max(wmax, nw);
                    }
            if (!(wmin > 0.0 || wmax < 0.0)) return false;
            continue;
        }
        double pu[4], pw[4], nvmax = 0.0;
        for (int c = 0; c < 4; ++c) {
            pu[c] = form(r.u, X[c & 1], Y[c >> 1], 0.0);
            pw[c] = form(r.w, X[c & 1], Y[c >> 1], 0.0);
        }
        for (double x : X)
            for (double y : Y)
                for (double z : Z) nvmax = std::max(nvmax, std::fabs(form(r.v, x, y, z)));
        // rho's minimum over the footprint (a parallelogram in (nu, nw)): zero when it holds the origin
        const int edge[4][2] = {{0, 1}, {1, 3}, {3, 2}, {2, 0}};
        double rho_min = 1e300, cross_sign = 0.0;
        bool inside = true;
        for (const auto &e : edge) {
            const double ax = pu[e[0]], ay = pw[e[0]], bx = pu[e[1]], by = pw[e[1]];
            const double ex = bx - ax, ey = by - ay, len2 = ex * ex + ey * ey;
            const double t = len2 > 0.0 ? std::clamp(-(ax * ex + ay * ey) / len2, 0.0, 1.0) : 0.0;
            rho_min = std::min(rho_min, std::hypot(ax + t * ex, ay + t * ey));
            const double cr = ax * ey - ay * ex;
            if (cross_sign == 0.0) cross_sign = cr;
            else if (cr * cross_sign <= 0.0) inside = false;
        }
        if (inside || !(rho_min > 0.0)) return false;
        const double step = std::max(std::hypot(r.u[1], r.w[1]), std::hypot(r.u[2], r.w[2]));
        const double h = 7.5 * step;
        const double dev = std::fabs(r.k[2]) * nvmax * h * h / (rho_min * rho_min * rho_min);
        if (!(dev < 0.5)) return false;
    }
    return true;
}

// Whether the arc kinds' FP32 kernels may take the fan angle without octant
// reduction (fan_atan_narrow): in every view every voxel lies in front of the
// source (nw > 0: nw is affine over the image, so its minimum is at a corner of
// the image's footprint, the z term included for a cylindrical detector; with a
// margin of 1e-5 of the form's magnitude, so that the FP32 form the kernels
// evaluate is positive too), and
// both padded edge columns of the detector, 0 and nu + 1, lie within the
// clamped angle's reach with half a column to spare -- then a voxel whose angle
// is clamped projects outside the detector, as its exact angle does.  In FP64
// from the FP64 records.
bool narrow_fan_ok(const gct_plan &p, const std::vector<BpRec64> &rec64)
{
    const gct_ct_geom &g = p.geom;
    if ((g.kind != GCT_FAN_ARC_1D && g.kind != GCT_CYLINDRICAL) || p.setup_f64) return false;
    const double reach = 0.5999999 - 0.5 * double(g.det_row.sampling_width);   // the clamp's angle, less half a column
    if (!(g.det_row.sampling_width > 0.0f) || !(reach > 0.0)) return false;
    const double xm = double(p.img.x.num_samples) - 1.0, ym = double(p.img.y.num_samples) - 1.0;
    const double zm = (p.img.dim == 3) ? double(p.img.z.num_samples) - 1.0 : 0.0;
    for (const BpRec64 &r : rec64) {
        const double mag = std::fabs(r.w[0]) + std::fabs(r.w[1]) * xm + std::fabs(r.w[2]) * ym + std::fabs(r.w[3]) * zm;
        for (int c = 0; c < 8; ++c) {
            const double x = (c & 1) ? xm : 0.0, y = (c & 2) ? ym : 0.0, z = (c & 4) ? zm : 0.0;
            if (!(r.w[0] + r.w[1] * x + r.w[2] * y + r.w[3] * z > 1e-5 * mag)) return false;
        }
        // padded column = gamma / dgamma + k0 (k0 = centre + 1)
        const double dg = g.det_row.sampling_width;
        if (!(r.k[0] * dg < reach && (double(p.nu) + 1.0 - r.k[0]) * dg < reach)) return false;
    }
    return true;
}

// The LDS shape a plan without shape options starts from (the autotuner can
// still replace it): the first of the shapes, in the order they rank on the
// flat-cone and C-arm workloads, whose staging area holds the boxes of at
// least 99 % of the (tile, view) pairs -- a box that does not fit sends its
// view through HBM.  The boxes follow the kernel's rule (the tile's corners,
// a row and column of margin, cut at the padded detector), here in FP64 from
// the FP64 records, over a sample of tiles (every 8th along each axis, and the
// last) and of views (24, evenly spaced).
int default_lds_shape(const gct_plan &p, const std::vector<BpRec64> &rec64)
{
    const gct_ct_geom &g = p.geom;
    const uint32_t nx = p.img.x.num_samples, ny = p.img.y.num_samples, nz = p.img.z.num_samples;
    const uint32_t nbx = (nx + 15u) / 16u, nby = (ny + 15u) / 16u;
    auto samples = [](uint32_t n) {
        std::vector<uint32_t> s;
        const uint32_t step = std::max(1u, n / 8u);
        for (uint32_t i = 0; i < n; i += step) s.push_back(i);
        if (s.back() != n - 1u) s.push_back(n - 1u);
        return s;
    };
    std::vector<uint32_t> views;
    for (uint32_t i = 0; i < 24u && i < p.V; ++i) views.push_back(uint32_t(uint64_t(i) * p.V / std::min(24u, p.V)));
    // padded (column, row) of voxel (x, y, z) in view v, as setup_column computes them
    auto project = [&](const BpRec64 &r, double x, double y, double z, double &col, double &row) {
        auto f = [&](const double q[4]) { return q[0] + q[1] * x + q[2] * y + q[3] * z; };
        if (g.kind == GCT_FLAT) {
            const double nw = f(r.w);
            col = f(r.u) / nw + r.k[0];
            row = f(r.v) / nw + r.k[1];
        } else if (g.kind == GCT_CYLINDRICAL) {
            const double nu = f(r.u), nw = f(r.w);
            col = std::atan2(nu, nw) * double(p.inv_dgamma) + r.k[0];
            row = r.k[2] * f(r.v) / std::sqrt(nu * nu + nw * nw) + r.k[1];
        } else {
            col = r.u[0] + r.u[1] * x + r.u[2] * y;
            row = r.v[0] + r.v[3] * z;
        }
    };
    // fraction of the sampled tile boxes that shape si stages whole
    auto fit_fraction = [&](int si) {
        const LdsShape &sh = kLdsShapes[si];
        const uint32_t nbz = (nz + sh.zu - 1u) / sh.zu;
        uint64_t total = 0, fit = 0;
        for (uint32_t v : views)
            for (uint32_t bz : samples(nbz))
                for (uint32_t by : samples(nby))
                    for (uint32_t bx : samples(nbx)) {
                        double cmin = 1e300, cmax = -1e300, rmin = 1e300, rmax = -1e300;
                        for (int c = 0; c < 8; ++c) {
                            const double x = std::min(bx * 16u + ((c & 1) ? 15u : 0u), nx - 1u);
                            const double y = std::min(by * 16u + ((c & 2) ? 15u : 0u), ny - 1u);
                            const double z = std::min(bz * sh.zu + ((c & 4) ? sh.zu - 1u : 0u), nz - 1u);
                            double col, row;
                            project(rec64[v], x, y, z, col, row);
                            cmin = std::min(cmin, col);
                            cmax = std::max(cmax, col);
                            rmin = std::min(rmin, row);
                            rmax = std::max(rmax, row);
                        }
                        ++total;
                        if (!(cmin > -1e6 && cmax < 1e6 && rmin > -1e6 && rmax < 1e6)) continue;
                        const int64_t clo = std::max<int64_t>(int64_t(std::floor(cmin)) - 2, -1);
                        const int64_t chi = std::min<int64_t>(int64_t(std::floor(cmax)) + 1, int64_t(p.nu) + 1);
                        const int64_t rlo = std::max<int64_t>(int64_t(std::floor(rmin)) - 2, -1);
                        const int64_t rhi = std::min<int64_t>(int64_t(std::floor(rmax)) + 1, int64_t(p.nv) + 1);
                        const int64_t nc = chi - clo + 1, nr = rhi - rlo + 1;
                        if (nc >= 2 && nr >= 2 && nc - 1 <= int64_t(sh.pairs) && nr <= int64_t(sh.rows)) ++fit;
                    }
        return total > 0 ? double(fit) / double(total) : 0.0;
    };
    static constexpr int kOrder[] = {6, 5, 4, 3, 1};
    for (int si : kOrder)
        if (fit_fraction(si) >= 0.99) return si;
    // No shape stages 99 % of the boxes whole (wide boxes): shapes 1 and 0
    // stage the same 40 pairs per row, and shape 1 8 fewer rows (32 against
    // 40), so it is the faster of the two wherever rows are not what overflows.
    return fit_fraction(1) >= fit_fraction(0) - 0.005 ? 1 : 0;
}

// The arc/cylindrical mapping uses rho = |(r.eu, r.ew)|, which equals the
// reference's |r| sqrt(1 - (r^.ev)^2) only for an orthonormal detector basis.
void check_basis(const gct_plan &p)
{
    if (!is_arc_row(p.geom.kind)) return;
    for (uint32_t v = 0; v < p.V; ++v) {
        const gct_proj_geom &pg = p.views[v];
        const int n = (p.geom.kind == GCT_CYLINDRICAL) ? 3 : 2;
        auto dot = [&](const float *a, const float *b) {
            double s = 0.0;
            for (int i = 0; i < n; ++i) s += double(a[i]) * double(b[i]);
            return s;
        };
        bool ok = std::fabs(dot(pg.det_row_dir, pg.det_row_dir) - 1.0) < 1e-5 &&
                  std::fabs(dot(pg.iso_ray_dir, pg.iso_ray_dir) - 1.0) < 1e-5 &&
                  std::fabs(dot(pg.det_row_dir, pg.iso_ray_dir)) < 1e-5;
        if (n == 3)
            ok = ok && std::fabs(dot(pg.det_col_dir, pg.det_col_dir) - 1.0) < 1e-5 &&
                 std::fabs(dot(pg.det_col_dir, pg.det_row_dir)) < 1e-5 &&
                 std::fabs(dot(pg.det_col_dir, pg.iso_ray_dir)) < 1e-5;
        if (!ok) fail(GCT_ERR_UNSUPPORTED, "arc and cylindrical detectors need an orthonormal detector basis");
    }
}

void upload(gct_plan &p, DevBuf &b, const void *src, size_t bytes)
{
    b = DevBuf(bytes, p.stream, p.ordered);
    hip_check(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, p.stream), "hipMemcpyAsync H2D");
}

// Builds a plan.  All device work (uploads, the redundancy table, clearing the
// filtered buffer) runs on `stream` and is complete when this returns; with
// ordered = true the buffers come from the stream-ordered pool.
}  // namespace

std::unique_ptr<gct_plan> make_plan(const gct_ct_geom &geom, const gct_image_geom &img, const gct_recon_opts &o,
                                    uint32_t filt_views, hipStream_t stream, bool ordered,
                                    const std::vector<uint32_t> *spiral_win)
{
    validate_geom(geom);
    validate_image(geom, img);
    if (geom.det_row.num_samples > kMaxDetCols) fail(GCT_ERR_UNSUPPORTED, "more than 2730 detector columns");
    // one padded filtered view and the row after it (FP32: (Nv + 4) x pitch x 4 B) are
    // addressed by 31-bit buffer offsets
    if (!is_1d(geom.kind) &&
        (uint64_t(geom.det_col.num_samples) + 4u) * ((kMaxDetCols + 3u + 15u) & ~15u) * 4u >= (uint64_t(1) << 31))
        fail(GCT_ERR_UNSUPPORTED, "more than 196000 detector rows");
    if (o.storage < GCT_STORE_F32 || o.storage > GCT_STORE_BF16) fail(GCT_ERR_INVALID_ARGUMENT, "unknown storage");
    if (o.bp_kernel < GCT_BP_AUTO || o.bp_kernel > GCT_BP_LDS) fail(GCT_ERR_INVALID_ARGUMENT, "unknown bp_kernel");
    // 2-D images default to the direct kernel: the LDS one stages whole
    // detector rows per view, many times what a 16 x 16 pixel tile reads
    const int variant = (o.bp_kernel == GCT_BP_DIRECT || (o.bp_kernel == GCT_BP_AUTO && is_1d(geom.kind)))
                            ? GCT_BP_DIRECT
                            : GCT_BP_LDS;
    if (o.z_unroll != 0 && !(variant == GCT_BP_DIRECT ? (o.z_unroll == 4 || o.z_unroll == 8 || o.z_unroll == 16)
                                                      : (o.z_unroll == 16 || o.z_unroll == 32)))
        fail(GCT_ERR_INVALID_ARGUMENT, "z_unroll must be 4, 8 or 16 (direct kernel) or 16 or 32 (LDS kernel)");
    if ((o.view_block || o.lds_rows || o.lds_pairs) &&
        lds_shape(o.z_unroll, o.view_block, o.lds_rows, o.lds_pairs) < 0)
        fail(GCT_ERR_INVALID_ARGUMENT,
             "no LDS backprojection shape matches (z_unroll, view_block, lds_rows, lds_pairs); the shapes are "
             "(16, 2, 40, 40), (16, 2, 32, 40), (16, 2, 40, 48), (16, 3, 28, 34), (16, 4, 24, 28), (32, 3, 48, 34), (32, 4, 36, 28)");
    auto p = std::make_unique<gct_plan>();
    p->device = o.device;
    p->stream = stream;
    p->ordered = ordered;
    p->opts = o;
    p->views.assign(geom.proj_geom, geom.proj_geom + geom.num_proj);
    p->geom = geom;
    p->geom.proj_geom = p->views.data();
    p->img = img;
    p->V = geom.num_proj;
    p->nu = geom.det_row.num_samples;
    p->nv = is_1d(geom.kind) ? 1u : geom.det_col.num_samples;
    p->L = fft_len(p->nu);
    p->rot = rotation_info(p->geom);   // throws for an invalid trajectory (ct_geom.h:26-32)
    p->scale = (p->rot.incr > 0.0f) ? p->rot.incr : 1.0f;   // fdk_axial.cuh:196-199
    p->inv_dgamma = static_cast<float>(1.0 / double(geom.det_row.sampling_width));
    const uint64_t nvox = uint64_t(img.x.num_samples) * img.y.num_samples * (img.dim == 3 ? img.z.num_samples : 1u);
    const bool cone = geom.kind == GCT_FLAT || geom.kind == GCT_CYLINDRICAL;
    p->setup_f64 = cone && (o.setup_precision == GCT_SETUP_F64 ||
                            (o.setup_precision == GCT_SETUP_AUTO && nvox >= uint64_t(512) * 512 * 512));
    p->pitch = (p->nu + 3u + 15u) & ~15u;
    p->rows_pv = p->nv + 3u;
    p->filt_views = filt_views ? std::min(filt_views, p->V) : p->V;
    check_basis(*p);

    DeviceGuard dg(p->device);
    std::vector<BpRec> rec;
    std::vector<BpRec64> rec64;
    std::vector<FiltRec> frec;
    build_records(*p, rec, rec64, frec);
    p->rows_inside = rows_inside_detector(*p, rec64);
    p->narrow_fan = narrow_fan_ok(*p, rec64);
    p->box_inner = box_inner_ok(*p, rec64);
    const bool shape_free = !o.z_unroll && !o.view_block && !o.lds_rows && !o.lds_pairs;
    const int auto_shape = (shape_free && variant == GCT_BP_LDS && p->circular && !is_1d(geom.kind))
                               ? default_lds_shape(*p, rec64)
                               : -1;
    upload(*p, p->d_brec, rec.data(), rec.size() * sizeof(BpRec));
    upload(*p, p->d_brec64, rec64.data(), rec64.size() * sizeof(BpRec64));
    upload(*p, p->d_frec, frec.data(), frec.size() * sizeof(FiltRec));
    upload(*p, p->d_views, p->views.data(), p->views.size() * sizeof(gct_proj_geom));

    // Filter spectrum: FFT_L of the N-tap kernel placed at n mod L, scaled by
    // 1/L (the cuFFT C2R normalisation of ramp_filtering_fft.cu:59), from the
    // FP32 taps the reference uses, evaluated in FP64 and stored FP32.
    std::vector<float> taps32;
    ramp_taps_f64(geom.kind, p->nu, geom.det_row.sampling_width, o.cut_off_freq, o.window, &taps32);
    std::vector<std::complex<double>> h(p->L, 0.0);
    const int half = static_cast<int>(p->nu) / 2;
    for (uint32_t k = 0; k < p->nu; ++k) {
        const int m = static_cast<int>(k) - half;
        h[static_cast<uint32_t>((m + static_cast<int>(p->L)) % static_cast<int>(p->L))] = taps32[k];
    }
    fft64_any(h);
    std::vector<float2> H(p->L), tw(p->L);
    for (uint32_t k = 0; k < p->L; ++k) {
        H[k] = make_float2(static_cast<float>(h[k].real() / p->L), static_cast<float>(h[k].imag() / p->L));
        const double a = -2.0 * std::numbers::pi * static_cast<double>(k) / static_cast<double>(p->L);
        tw[k] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    upload(*p, p->d_H, H.data(), H.size() * sizeof(float2));
    upload(*p, p->d_tw, tw.data(), tw.size() * sizeof(float2));

    // Redundancy table [V][Nu] (getParkerWeights, parker_weighting.cuh:132-200):
    // fan angles from view 0 on the host in FP32 as the reference does
    // (fanAngles :61-77), the table on the device.
    p->d_red = DevBuf(size_t(p->V) * p->nu * sizeof(float), p->stream, p->ordered);
    {
        int mode = 0;
        float value = 1.0f;
        std::vector<float> fan(p->nu, 0.0f), beta(p->V, 0.0f);
        if (spiral_win) {
            // the spiral FDK applies no redundancy weighting (fdk_spiral.cuh:71-82)
        } else if (is_parallel_family(geom.kind)) {
            const float pi_f = std::numbers::pi_v<float>;
            value = (p->rot.incr > 0.0f && p->rot.range <= (p->rot.incr + pi_f)) ? 1.0f : 0.5f;
        } else if (p->rot.range > std::numbers::pi_v<float>) {
            mode = 1;
            const gct_proj_geom &v0 = p->views[0];
            for (uint32_t c = 0; c < p->nu; ++c) {
                const float d = (static_cast<float>(c) - v0.det_row_center) * geom.det_row.sampling_width;
                fan[c] = is_arc_row(geom.kind) ? d : std::atan2(d, v0.src_det_dist);
            }
            for (uint32_t v = 0; v < p->V; ++v) beta[v] = (p->rot.dir < 0) ? p->rot.rel_ang[p->V - 1 - v] : p->rot.rel_ang[v];
        }
        // range <= pi for a fan geometry: the reference leaves the table
        // uninitialised (parker_weighting.cuh:163-175); this library uses 1.
        DevBuf dfan, dbeta;
        upload(*p, dfan, fan.data(), fan.size() * sizeof(float));
        upload(*p, dbeta, beta.data(), beta.size() * sizeof(float));
        launch_redundancy(geom.kind, dfan.as<float>(), dbeta.as<float>(), p->nu, p->V, p->rot.range, mode, value,
                          p->d_red.as<float>(), p->stream);
        // spare zero elements after the last view for the LDS staging loads
        const size_t fbytes = (size_t(p->filt_views) * p->rows_pv * p->pitch + filt_spare_elems(p->pitch)) *
                              storage_bytes(o.storage);
        p->d_filt = DevBuf(fbytes, p->stream, p->ordered);
        hip_check(hipMemsetAsync(p->d_filt.p, 0, fbytes, p->stream), "hipMemsetAsync");   // zero padding stays zero
        // host vectors above must outlive their copies; dfan/dbeta free in order
        hip_check(hipStreamSynchronize(p->stream), "plan set-up");
    }

    if (spiral_win) {
        if (!p->circular) fail(GCT_ERR_UNSUPPORTED, "spiral reconstruction needs views whose u and w do not depend on z");
        p->spiral = true;
        upload(*p, p->d_zwin, spiral_win->data(), spiral_win->size() * sizeof(uint32_t));
        hip_check(hipStreamSynchronize(p->stream), "spiral windows");
    }
    p->variant = variant;
    p->z_unroll = o.z_unroll ? o.z_unroll : (variant == GCT_BP_LDS ? 16u : 8u);
    p->view_block = o.view_block;
    p->lds_rows = o.lds_rows;
    p->lds_pairs = o.lds_pairs;
    if (auto_shape >= 0) {
        const LdsShape &sh = kLdsShapes[auto_shape];
        p->z_unroll = sh.zu;
        p->view_block = sh.vb;
        p->lds_rows = sh.rows;
        p->lds_pairs = sh.pairs;
    }
    // Views whose u and w move along z: flat panels use the LDS kernel of that
    // case, which has the one shape (16, 2, 40, 40); other kinds the direct kernel.
    if (!p->circular && p->variant == GCT_BP_LDS && img.dim == 3) {
        if (geom.kind == GCT_FLAT) {
            const LdsShape &sh = kLdsShapes[0];
            p->z_unroll = sh.zu;
            p->view_block = sh.vb;
            p->lds_rows = sh.rows;
            p->lds_pairs = sh.pairs;
        } else {
            p->variant = GCT_BP_DIRECT;
            p->z_unroll = 8u;
        }
    }
    // Untuned LDS plans deal their tiles round robin unless the volume is at
    // most 4 tiles deep: measured, round robin is 1.6 % faster over the 3-D
    // requests of the serving mix (up to 6 % on flat cubes), contiguous runs
    // 13 % faster on the 4-tile-deep cylindrical slab.
    if (p->variant == GCT_BP_LDS && img.dim == 3)
        p->round_robin = (img.z.num_samples + p->z_unroll - 1u) / p->z_unroll > 4u;
    if (!o.z_unroll && o.bp_kernel == GCT_BP_AUTO) {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune.find({geom.kind, img.x.num_samples, img.y.num_samples, img.dim == 3 ? img.z.num_samples : 1u,
                               p->V, p->nu, p->nv, o.storage, int(p->circular), int(p->setup_f64)});
        if (it != g_tune.end()) {
            p->z_unroll = it->second.z_unroll;
            p->variant = it->second.variant;
            p->view_block = it->second.view_block;
            p->lds_rows = it->second.lds_rows;
            p->lds_pairs = it->second.lds_pairs;
            p->round_robin = it->second.round_robin;
        }
    }
    return p;
}

FilterLaunch filter_args(gct_plan &p, const float *proj, uint32_t proj_view_offset, uint32_t first, uint32_t n,
                         uint32_t filt_view_offset)
{
    FilterLaunch a{};
    a.kind = p.geom.kind;
    a.nu = p.nu;
    a.nv = p.nv;
    a.L = p.L;
    a.first_view = first;
    a.num_views = n;
    a.du = p.geom.det_row.sampling_width;
    a.dv = is_1d(p.geom.kind) ? 1.0f : p.geom.det_col.sampling_width;
    a.proj = proj;
    a.proj_view_offset = proj_view_offset;
    a.frec = p.d_frec.as<FiltRec>();
    a.red = p.d_red.as<float>();
    a.H = p.d_H.as<float2>();
    a.tw = p.d_tw.as<float2>();
    // the filter kernel writes view v at buffer index v - filt_view_offset
    const size_t view_elems = size_t(p.rows_pv) * p.pitch;
    a.filt = static_cast<char *>(p.d_filt.p) -
             static_cast<ptrdiff_t>(size_t(filt_view_offset) * view_elems * storage_bytes(p.opts.storage));
    a.pitch = p.pitch;
    a.rows_pv = p.rows_pv;
    a.storage = p.opts.storage;
    return a;
}

BpLaunch bp_args(gct_plan &p, float *vol, uint32_t first, uint32_t n, bool accumulate, bool finish,
                 uint32_t filt_view_offset)
{
    BpLaunch a{};
    a.kind = p.geom.kind;
    a.storage = p.opts.storage;
    a.circular = p.circular;
    a.setup_f64 = p.setup_f64;
    a.nx = p.img.x.num_samples;
    a.ny = p.img.y.num_samples;
    a.nz = (p.img.dim == 3) ? p.img.z.num_samples : 1u;
    a.z0 = 0;
    a.zcount = a.nz;
    a.nu = p.nu;
    a.nv = p.nv;
    a.first_view = first;
    a.num_views = n;
    a.pitch = p.pitch;
    a.rows_pv = p.rows_pv;
    const size_t view_elems = size_t(p.rows_pv) * p.pitch;
    a.filt = static_cast<const char *>(p.d_filt.p) -
             static_cast<ptrdiff_t>(size_t(filt_view_offset) * view_elems * storage_bytes(p.opts.storage));
    a.rec = p.d_brec.as<BpRec>();
    a.rec64 = p.d_brec64.as<BpRec64>();
    a.vol = vol;
    a.inv_dgamma = p.inv_dgamma;
    a.scale = p.scale;
    a.zwin = p.spiral ? p.d_zwin.as<uint32_t>() : nullptr;
    a.accumulate = accumulate;
    a.finish = finish;
    a.z_unroll = p.z_unroll;
    a.variant = p.variant;
    a.view_block = p.view_block;
    a.lds_rows = p.lds_rows;
    a.lds_pairs = p.lds_pairs;
    a.rows_inside = p.rows_inside;
    a.narrow_fan = p.narrow_fan;
    a.box_inner = p.box_inner;
    a.round_robin = p.round_robin;
    return a;
}

namespace {

void check_range(const gct_plan &p, uint32_t first, uint32_t n)
{
    if (first > p.V || n > p.V - first) fail(GCT_ERR_INVALID_ARGUMENT, "view range outside the scan");
    if (n > p.filt_views) fail(GCT_ERR_INVALID_ARGUMENT, "view range larger than the plan's filtered buffer");
}

}  // namespace

FpLaunch fp_args(gct_plan &p, const float *vol, float *proj, float *pad)
{
    FpLaunch a{};
    a.kind = p.geom.kind;
    a.nx = p.img.x.num_samples;
    a.ny = p.img.y.num_samples;
    a.nz = (p.img.dim == 3) ? p.img.z.num_samples : 1u;
    if (uint64_t(a.nx) * a.ny * a.nz >= (uint64_t(1) << 32))
        fail(GCT_ERR_UNSUPPORTED, "forward projection needs a volume below 2^32 voxels (32-bit in-plane offsets)");
    a.dx = p.img.x.sampling_width;
    a.dy = p.img.y.sampling_width;
    a.dz = (p.img.dim == 3) ? p.img.z.sampling_width : 1.0f;
    a.x0 = static_cast<float>(double(p.img.mid_point[0]) - (double(a.nx) - 1.0) * a.dx / 2.0);
    a.y0 = static_cast<float>(double(p.img.mid_point[1]) - (double(a.ny) - 1.0) * a.dy / 2.0);
    a.z0 = (p.img.dim == 3) ? static_cast<float>(double(p.img.mid_point[2]) - (double(a.nz) - 1.0) * a.dz / 2.0) : 0.0f;
    a.nu = p.nu;
    a.nv = p.nv;
    a.nviews = p.V;
    a.du = p.geom.det_row.sampling_width;
    a.dv = is_1d(p.geom.kind) ? 1.0f : p.geom.det_col.sampling_width;
    a.views = p.d_views.as<gct_proj_geom>();
    a.vol = vol;
    a.pad = joseph_pad_elems(a.kind, a.nx, a.ny, a.nz) ? pad : nullptr;
    a.proj = proj;
    return a;
}

uint64_t plan_device_bytes(const gct_ct_geom &g, const gct_image_geom &img, const gct_recon_opts &o, uint32_t fviews)
{
    const uint64_t nu = g.det_row.num_samples, nv = is_1d(g.kind) ? 1u : g.det_col.num_samples;
    const uint64_t pitch = (nu + 3u + 15u) & ~uint64_t(15);
    const uint64_t V = g.num_proj;
    const uint64_t L = fft_len(uint32_t(nu));
    uint64_t b = (uint64_t(fviews ? fviews : V) * (nv + 3u) * pitch + filt_spare_elems(uint32_t(pitch))) *
                 storage_bytes(o.storage);
    b += V * (sizeof(BpRec) + sizeof(BpRec64) + sizeof(FiltRec) + sizeof(gct_proj_geom)) + V * nu * 4u + 2u * L * 8u;
    (void)img;
    return b;
}

uint64_t volume_bytes(const gct_image_geom &img)
{
    return uint64_t(img.x.num_samples) * img.y.num_samples * (img.dim == 3 ? img.z.num_samples : 1u) * sizeof(float);
}

}  // namespace gct

// ---- C ABI ------------------------------------------------------------------

extern "C" uint32_t gct_abi_version(void) { return GCT_AMD_ABI_VERSION; }

extern "C" const char *gct_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *gct_status_string(gct_status s)
{
    switch (s) {
    case GCT_OK: return "ok";
    case GCT_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GCT_ERR_RUNTIME: return "runtime error";
    case GCT_ERR_OUT_OF_MEMORY: return "out of memory";
    case GCT_ERR_NO_DEVICE: return "no device";
    case GCT_ERR_UNSUPPORTED: return "unsupported";
    case GCT_ERR_REJECTED: return "rejected";
    case GCT_ERR_NOT_READY: return "not ready";
    }
    return "unknown status";
}

extern "C" void gct_recon_opts_default(gct_recon_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->cut_off_freq = 1.0f;
    o->window = GCT_WINDOW_NONE;
    o->storage = GCT_STORE_F32;
    o->setup_precision = GCT_SETUP_AUTO;
    o->mem_mode = GCT_MEM_AUTO;
    o->device = 0;
}

extern "C" int gct_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

gct_recon_opts gct::opts_or_default(const gct_recon_opts *o)
{
    gct_recon_opts d;
    gct_recon_opts_default(&d);
    return o ? *o : d;
}

extern "C" gct_status gct_plan_create(const gct_ct_geom *geom, const gct_image_geom *img, const gct_recon_opts *opts,
                                      gct_plan **out)
{
    return guarded([&] {
        if (!geom || !img || !out) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (gct_device_count() == 0) fail(GCT_ERR_NO_DEVICE, "no HIP device");
        *out = make_plan(*geom, *img, opts_or_default(opts), 0).release();
    });
}

extern "C" void gct_plan_destroy(gct_plan *p)
{
    if (!p) return;
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(p->device);
    delete p;
    (void)hipSetDevice(prev);
}

extern "C" gct_status gct_plan_get_info(const gct_plan *p, gct_plan_info *info)
{
    return guarded([&] {
        if (!p || !info) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        std::memset(info, 0, sizeof(*info));
        info->fft_len = p->L;
        info->filt_pitch = p->pitch;
        info->filt_rows_per_view = p->rows_pv;
        info->z_unroll = p->z_unroll;
        info->filt_bytes = p->d_filt.bytes;
        info->device_bytes = p->d_filt.bytes + p->d_brec.bytes + p->d_brec64.bytes + p->d_frec.bytes + p->d_red.bytes +
                             p->d_H.bytes + p->d_tw.bytes + p->d_views.bytes;
        info->updates = uint64_t(p->img.x.num_samples) * p->img.y.num_samples *
                        (p->img.dim == 3 ? p->img.z.num_samples : 1u) * p->V;
        info->rot_ang_incr = p->rot.incr;
        info->rot_ang_range = p->rot.range;
        info->rot_dir = p->rot.dir;
        info->setup_f64 = p->setup_f64;
        info->circular = p->circular;
        info->storage = p->opts.storage;
        info->bp_kernel = p->variant;
        const int shape = lds_shape(p->z_unroll, p->view_block, p->lds_rows, p->lds_pairs);
        const bool lds = p->variant == GCT_BP_LDS && shape >= 0;
        info->view_block = lds ? kLdsShapes[shape].vb : 0u;
        info->lds_rows = lds ? kLdsShapes[shape].rows : 0u;
        info->lds_pairs = lds ? kLdsShapes[shape].pairs : 0u;
        info->rows_inside = p->rows_inside;
        info->round_robin = lds && p->round_robin;
        info->narrow_fan = p->narrow_fan;
        info->box_inner = p->box_inner;
    });
}

extern "C" gct_status gct_plan_filter(gct_plan *p, const float *proj, uint32_t first, uint32_t n, void *stream)
{
    return guarded([&] {
        if (!p || !proj) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        check_range(*p, first, n);
        DeviceGuard dg(p->device);
        launch_filter(filter_args(*p, proj, 0, first, n, p->filt_views < p->V ? first : 0), static_cast<hipStream_t>(stream));
    });
}

extern "C" gct_status gct_plan_backproject(gct_plan *p, float *vol, uint32_t first, uint32_t n, int accumulate,
                                           int finish, void *stream)
{
    return guarded([&] {
        if (!p || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        check_range(*p, first, n);
        DeviceGuard dg(p->device);
        launch_backproject(bp_args(*p, vol, first, n, accumulate != 0, finish != 0, p->filt_views < p->V ? first : 0),
                           static_cast<hipStream_t>(stream));
    });
}

extern "C" gct_status gct_plan_execute(gct_plan *p, const float *proj, float *vol, void *stream)
{
    return guarded([&] {
        if (!p || !proj || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (p->filt_views < p->V) fail(GCT_ERR_INVALID_ARGUMENT, "plan holds a partial filtered buffer");
        DeviceGuard dg(p->device);
        hipStream_t s = static_cast<hipStream_t>(stream);
        launch_filter(filter_args(*p, proj, 0, 0, p->V, 0), s);
        launch_backproject(bp_args(*p, vol, 0, p->V, false, true, 0), s);
    });
}

extern "C" gct_status gct_plan_load_filtered(gct_plan *p, const float *filt, void *stream)
{
    return guarded([&] {
        if (!p || !filt) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        DeviceGuard dg(p->device);
        launch_pack_filtered(filt, p->d_filt.p, p->opts.storage, p->nu, p->nv, p->filt_views, p->pitch, p->rows_pv,
                             static_cast<hipStream_t>(stream));
    });
}

extern "C" gct_status gct_plan_store_filtered(gct_plan *p, float *filt, void *stream)
{
    return guarded([&] {
        if (!p || !filt) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        DeviceGuard dg(p->device);
        launch_unpack_filtered(p->d_filt.p, filt, p->opts.storage, p->nu, p->nv, p->filt_views, p->pitch, p->rows_pv,
                               static_cast<hipStream_t>(stream));
    });
}

extern "C" gct_status gct_plan_forward_project(gct_plan *p, const float *vol, float *proj, void *stream)
{
    return guarded([&] {
        if (!p || !vol || !proj) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        DeviceGuard dg(p->device);
        // the padded-copy scratch is allocated on first use and kept with the plan
        const uint64_t pad = joseph_pad_elems(p->geom.kind, p->img.x.num_samples, p->img.y.num_samples,
                                              p->img.dim == 3 ? p->img.z.num_samples : 1u);
        if (pad && !p->d_vol_pad.p) p->d_vol_pad = DevBuf(pad * sizeof(float));
        const FpLaunch a = fp_args(*p, vol, proj, p->d_vol_pad.as<float>());
        launch_forward_project(a, static_cast<hipStream_t>(stream));
    });
}

extern "C" gct_status gct_required_memory(const gct_ct_geom *geom, const gct_image_geom *img,
                                          const gct_recon_opts *opts, uint64_t *bytes)
{
    return guarded([&] {
        if (!geom || !img || !bytes) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        validate_geom(*geom);
        validate_image(*geom, *img);
        const gct_recon_opts o = opts_or_default(opts);
        const uint64_t nv = is_1d(geom->kind) ? 1u : geom->det_col.num_samples;
        const uint64_t proj = uint64_t(geom->num_proj) * nv * geom->det_row.num_samples * sizeof(float);
        *bytes = proj + volume_bytes(*img) + plan_device_bytes(*geom, *img, o, 0);
    });
}

namespace {

// Streams of the library's own work.  Non-blocking: a blocking stream waits
// for, and holds up, the null stream, so copies on it could not overlap a
// caller's kernels on the null stream.
struct Stream {
    hipStream_t s = nullptr;
    Stream() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"); }
    ~Stream() { (void)hipStreamDestroy(s); }
    Stream(const Stream &) = delete;
    Stream &operator=(const Stream &) = delete;
};

struct Event {
    hipEvent_t e = nullptr;
    Event() { hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"); }
    ~Event() { (void)hipEventDestroy(e); }
    Event(const Event &) = delete;
    Event &operator=(const Event &) = delete;
};

// Two device staging buffers for chunks of the host projection stack: the
// copy of one chunk (on a stream of its own) overlaps the work queued on the
// other.  stage() returns the buffer the work must read; release() marks that
// work as queued, after which the buffer is refilled only once it is done.
class Stager {
public:
    explicit Stager(size_t bytes) : buf_{DevBuf(bytes), DevBuf(bytes)} {}
    ~Stager() { (void)hipStreamSynchronize(cs_.s); }
    const float *stage(const float *src, size_t bytes, hipStream_t s)
    {
        if (used_[idx_]) hip_check(hipStreamWaitEvent(cs_.s, done_[idx_].e, 0), "wait");
        hip_check(hipMemcpyAsync(buf_[idx_].p, src, bytes, hipMemcpyHostToDevice, cs_.s), "copy projection chunk");
        hip_check(hipEventRecord(copied_[idx_].e, cs_.s), "record");
        hip_check(hipStreamWaitEvent(s, copied_[idx_].e, 0), "wait");
        return buf_[idx_].as<float>();
    }
    void release(hipStream_t s)
    {
        hip_check(hipEventRecord(done_[idx_].e, s), "record");
        used_[idx_] = true;
        idx_ ^= 1u;
    }

private:
    DevBuf buf_[2];
    Stream cs_;
    Event copied_[2], done_[2];
    bool used_[2] = {false, false};
    uint32_t idx_ = 0;
};

// Views per chunk when the stack streams through: about 1/16 of free memory
// for a chunk's staging and filtered views.  The filter transforms detector
// rows in pairs (2k, 2k + 1) counted from the launch's first view, and a
// pair's rounding depends on both rows.  Every chunk must start at an even
// row, as the whole stack's launch does, so that chunked and whole-stack
// filtering pair (and round) alike: with an odd Nv, an even number of views
// per chunk.
uint32_t view_chunk(const gct_ct_geom &geom, const gct_recon_opts &o, size_t free_b, uint64_t view_bytes)
{
    const uint64_t nv = is_1d(geom.kind) ? 1u : geom.det_col.num_samples;
    uint32_t chunk = o.view_chunk;
    if (!chunk) {
        const uint64_t per = view_bytes + (nv + 3u) * ((geom.det_row.num_samples + 18u) & ~15u) * storage_bytes(o.storage);
        chunk = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(geom.num_proj, (free_b / 16) / per)));
    }
    chunk = std::min(chunk, geom.num_proj);
    if ((nv & 1u) && (chunk & 1u) && chunk < geom.num_proj) chunk = chunk > 1u ? chunk - 1u : 2u;
    return chunk;
}

// Filters and backprojects views [0, V) into `vol` (slices [z0, z0 + zcount)),
// chunk by chunk through the stager: each voxel adds its views in scan order.
void stream_views(gct_plan &plan, const float *proj, uint64_t view_bytes, uint32_t chunk, Stager &st, float *vol,
                  uint32_t z0, uint32_t zcount, hipStream_t s)
{
    for (uint32_t first = 0; first < plan.V; first += chunk) {
        const uint32_t n = std::min(chunk, plan.V - first);
        const float *d = st.stage(proj + size_t(first) * view_bytes / sizeof(float), size_t(n) * view_bytes, s);
        launch_filter(filter_args(plan, d, first, first, n, first), s);
        BpLaunch a = bp_args(plan, vol, first, n, first > 0, first + n >= plan.V, first);
        a.z0 = z0;
        a.zcount = zcount;
        launch_backproject(a, s);
        st.release(s);
    }
}

// Volumes beyond device memory: the volume is reconstructed in z slabs of
// `zslab` slices (a multiple of 32, so every backprojection tile, and the
// detector box it stages, is the whole volume's), two slab buffers in turn,
// the copy of slab k back to the host overlapping the backprojection of
// slab k + 1.  The filtered stack is built once and kept when it fits
// (`resident`); otherwise every slab streams the views through again.  A
// voxel's sum is the same either way as in the whole-volume paths.
void reconstruct_slabs(const gct_ct_geom &geom, const float *proj, const gct_image_geom &img, float *vol,
                       const gct_recon_opts &o, size_t free_b, uint64_t view_bytes)
{
    const uint32_t nz = img.z.num_samples;
    const uint64_t slice = uint64_t(img.x.num_samples) * img.y.num_samples;
    const uint32_t chunk = view_chunk(geom, o, free_b, view_bytes);
    const uint64_t stage_b = 2u * uint64_t(chunk) * view_bytes;
    // explicit view_chunk: the views stream per slab; otherwise the whole
    // filtered stack stays resident when it takes at most half of free memory
    const bool resident = !o.view_chunk && plan_device_bytes(geom, img, o, 0) + stage_b <= free_b / 2;
    const uint64_t fixed = plan_device_bytes(geom, img, o, resident ? 0u : chunk) + stage_b;
    uint32_t zslab = o.z_slab;
    if (!zslab) {
        const uint64_t avail = free_b > fixed ? (free_b - fixed) / 10u * 8u : 0u;
        zslab = static_cast<uint32_t>(std::min<uint64_t>(nz, avail / (2u * slice * sizeof(float))) & ~31u);
        if (!zslab) fail(GCT_ERR_OUT_OF_MEMORY, "device memory holds no 32-slice slab beside the filtered views");
    }
    zslab = std::min(zslab, nz);
    auto plan = make_plan(geom, img, o, resident ? geom.num_proj : chunk);
    Stream s, cs;
    Event bp_done[2], copied[2];
    Stager st(size_t(chunk) * view_bytes);
    if (resident) {
        for (uint32_t first = 0; first < plan->V; first += chunk) {
            const uint32_t n = std::min(chunk, plan->V - first);
            const float *d = st.stage(proj + size_t(first) * view_bytes / sizeof(float), size_t(n) * view_bytes, s.s);
            launch_filter(filter_args(*plan, d, first, first, n, 0), s.s);
            st.release(s.s);
        }
    }
    DevBuf slab[2] = {DevBuf(size_t(zslab) * slice * sizeof(float)), DevBuf(size_t(zslab) * slice * sizeof(float))};
    // The copy of slab k - 1 is queued after the backprojection of slab k, so
    // that the two overlap also when a copy to pageable memory holds the host.
    auto copy_back = [&](uint32_t k) {
        const uint32_t z0 = k * zslab, zc = std::min(zslab, nz - z0);
        hip_check(hipStreamWaitEvent(cs.s, bp_done[k & 1u].e, 0), "wait");
        hip_check(hipMemcpyAsync(vol + size_t(z0) * slice, slab[k & 1u].p, size_t(zc) * slice * sizeof(float),
                                 hipMemcpyDeviceToHost, cs.s),
                  "copy slab");
        hip_check(hipEventRecord(copied[k & 1u].e, cs.s), "record");
    };
    const uint32_t nslab = (nz + zslab - 1u) / zslab;
    for (uint32_t k = 0; k < nslab; ++k) {
        const uint32_t z0 = k * zslab, zc = std::min(zslab, nz - z0);
        // slab buffer k & 1 was last copied out by copy_back(k - 2); the
        // backprojection waits for that copy
        if (k >= 2) hip_check(hipStreamWaitEvent(s.s, copied[k & 1u].e, 0), "wait");
        float *dv = slab[k & 1u].as<float>();
        if (resident) {
            BpLaunch a = bp_args(*plan, dv, 0, plan->V, false, true, 0);
            a.z0 = z0;
            a.zcount = zc;
            launch_backproject(a, s.s);
        } else {
            stream_views(*plan, proj, view_bytes, chunk, st, dv, z0, zc, s.s);
        }
        hip_check(hipEventRecord(bp_done[k & 1u].e, s.s), "record");
        if (k >= 1) copy_back(k - 1);
    }
    copy_back(nslab - 1u);
    hip_check(hipStreamSynchronize(cs.s), "copy slabs");
    hip_check(hipStreamSynchronize(s.s), "reconstruct slabs");
}

// Views per chunk of a pipelined host-to-host execution: about an eighth of
// the scan, even when Nv is odd (the filter's row pairs, as in view_chunk).
uint32_t pipeline_chunk(const gct_plan &p, uint32_t requested)
{
    uint32_t c = requested ? std::min(requested, p.V) : std::max(1u, (p.V + 7u) / 8u);
    if ((p.nv & 1u) && (c & 1u) && c < p.V) c = c > 1u ? c - 1u : 2u;
    return c;
}

// Host projections -> device -> volume -> host, pipelined on `s`: views are
// uploaded chunk by chunk on a stream of their own, and each chunk is
// filtered and backprojected (accumulating) as soon as it is on the device,
// so the upload of chunk i + 1 overlaps the work on chunk i; the volume is
// copied back after the last chunk.  Each voxel adds its views in scan order,
// so the volume is the one-launch path's bit for bit.  Enqueued in chunk
// order: a copy from pageable memory that holds the host thread still lets
// the previous chunk's kernels run meanwhile.
void execute_host_pipelined(gct_plan &p, const float *proj_host, float *proj_dev, float *vol_dev, float *vol_host,
                            uint32_t chunk, hipStream_t s)
{
    const size_t view_elems = size_t(p.nv) * p.nu;
    const uint32_t nz = p.img.dim == 3 ? p.img.z.num_samples : 1u;
    const size_t vol_bytes = size_t(p.img.x.num_samples) * p.img.y.num_samples * nz * sizeof(float);
    const uint32_t nchunks = (p.V + chunk - 1u) / chunk;
    Stream cs;
    Event ready;
    std::vector<Event> copied(nchunks);
    // the uploads overwrite proj_dev: after the work already queued on s
    hip_check(hipEventRecord(ready.e, s), "record");
    hip_check(hipStreamWaitEvent(cs.s, ready.e, 0), "wait");
    for (uint32_t i = 0; i < nchunks; ++i) {
        const uint32_t first = i * chunk, n = std::min(chunk, p.V - first);
        hip_check(hipMemcpyAsync(proj_dev + first * view_elems, proj_host + first * view_elems,
                                 size_t(n) * view_elems * sizeof(float), hipMemcpyHostToDevice, cs.s),
                  "copy projection chunk");
        hip_check(hipEventRecord(copied[i].e, cs.s), "record");
        hip_check(hipStreamWaitEvent(s, copied[i].e, 0), "wait");
        launch_filter(filter_args(p, proj_dev, 0, first, n, 0), s);
        launch_backproject(bp_args(p, vol_dev, first, n, first > 0, first + n >= p.V, 0), s);
    }
    hip_check(hipMemcpyAsync(vol_host, vol_dev, vol_bytes, hipMemcpyDeviceToHost, s), "copy volume");
}

}  // namespace

extern "C" gct_status gct_plan_execute_host(gct_plan *p, const float *proj_host, float *proj_dev, float *vol_dev,
                                            float *vol_host, uint32_t chunk_views, void *stream)
{
    return guarded([&] {
        if (!p || !proj_host || !proj_dev || !vol_dev || !vol_host) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (p->filt_views < p->V) fail(GCT_ERR_INVALID_ARGUMENT, "plan holds a partial filtered buffer");
        DeviceGuard dg(p->device);
        execute_host_pipelined(*p, proj_host, proj_dev, vol_dev, vol_host, pipeline_chunk(*p, chunk_views),
                               static_cast<hipStream_t>(stream));
    });
}

// reconFDK (fdk_axial.cuh:69-89): every view resident when the estimate is
// below 75 % of device memory, otherwise views stream through in chunks.
// Both paths add each voxel's views in scan order, so they give the same
// volume bit for bit (the property fdk_axial_tests.cu:68-73 asserts).  A
// volume that does not fit beside a chunk's buffers (the reference has no
// path for it) is reconstructed in z slabs, with the same sums again.
extern "C" gct_status gct_reconstruct(const gct_ct_geom *geom, const float *proj, const gct_image_geom *img, float *vol,
                                      const gct_recon_opts *opts)
{
    return guarded([&] {
        if (!geom || !proj || !img || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (gct_device_count() == 0) fail(GCT_ERR_NO_DEVICE, "no HIP device");
        const gct_recon_opts o = opts_or_default(opts);
        validate_geom(*geom);
        validate_image(*geom, *img);
        if (o.mem_mode < GCT_MEM_AUTO || o.mem_mode > GCT_MEM_SLAB) fail(GCT_ERR_INVALID_ARGUMENT, "unknown mem_mode");
        if (o.z_slab % 32u) fail(GCT_ERR_INVALID_ARGUMENT, "z_slab must be a multiple of 32 slices");
        DeviceGuard dg(o.device);
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const uint64_t nv = is_1d(geom->kind) ? 1u : geom->det_col.num_samples;
        const uint64_t view_bytes = nv * geom->det_row.num_samples * sizeof(float);
        const uint64_t need = uint64_t(geom->num_proj) * view_bytes + volume_bytes(*img) + plan_device_bytes(*geom, *img, o, 0);
        int mode = o.mem_mode;
        if (mode == GCT_MEM_AUTO) {
            mode = double(need) / double(total_b) < 0.75 ? GCT_MEM_HIGH : GCT_MEM_LOW;
            if (mode == GCT_MEM_LOW && img->dim == 3) {
                const uint32_t c = view_chunk(*geom, o, free_b, view_bytes);
                const uint64_t low = volume_bytes(*img) + plan_device_bytes(*geom, *img, o, c) + 2u * uint64_t(c) * view_bytes;
                if (double(low) >= 0.9 * double(free_b)) mode = GCT_MEM_SLAB;
            }
        }
        if (mode == GCT_MEM_SLAB && img->dim == 3) {
            reconstruct_slabs(*geom, proj, *img, vol, o, free_b, view_bytes);
            return;
        }
        if (mode == GCT_MEM_SLAB) mode = GCT_MEM_LOW;   // a 2-D image is one slice
        const bool low = mode == GCT_MEM_LOW;
        const uint32_t chunk = low ? view_chunk(*geom, o, free_b, view_bytes) : geom->num_proj;
        auto plan = make_plan(*geom, *img, o, chunk);
        Stream s;
        DevBuf dvol(volume_bytes(*img));
        if (!low) {
            // every view resident; the upload is pipelined with the work
            DevBuf dproj(size_t(geom->num_proj) * view_bytes);
            execute_host_pipelined(*plan, proj, dproj.as<float>(), dvol.as<float>(), vol,
                                   pipeline_chunk(*plan, o.view_chunk), s.s);
            hip_check(hipStreamSynchronize(s.s), "reconstruct");
        } else {
            Stager st(size_t(chunk) * view_bytes);
            stream_views(*plan, proj, view_bytes, chunk, st, dvol.as<float>(), 0, plan->img.dim == 3 ? plan->img.z.num_samples : 1u,
                         s.s);
            hip_check(hipMemcpyAsync(vol, dvol.p, dvol.bytes, hipMemcpyDeviceToHost, s.s), "copy volume");
            hip_check(hipStreamSynchronize(s.s), "reconstruct");
        }
    });
}

// ---- autotuner ----------------------------------------------------------------

// reconFDK for SpiralCTGeom (fdk_spiral.cuh:28-94): every view is weighted
// and filtered once (the reference refilters each view for every slice), then
// one backprojection launch covers all slices with their own view windows.
extern "C" gct_status gct_reconstruct_spiral(const gct_ct_geom *geom, const float *proj, const gct_image_geom *img,
                                             float *vol, const gct_recon_opts *opts)
{
    return guarded([&] {
        if (!geom || !proj || !img || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (gct_device_count() == 0) fail(GCT_ERR_NO_DEVICE, "no HIP device");
        gct_recon_opts o = opts_or_default(opts);
        validate_geom(*geom);
        validate_image(*geom, *img);
        if (geom->kind != GCT_CYLINDRICAL && geom->kind != GCT_FLAT)
            fail(GCT_ERR_UNSUPPORTED, "spiral reconstruction is available for cylindrical and flat detectors");
        const std::vector<uint32_t> win = spiral_windows(*geom, *img);
        o.bp_kernel = GCT_BP_LDS;
        o.z_unroll = 0;
        o.view_block = 0;
        o.lds_rows = 0;
        o.lds_pairs = 0;
        DeviceGuard dg(o.device);
        auto plan = make_plan(*geom, *img, o, geom->num_proj, nullptr, false, &win);
        hipStream_t s = nullptr;
        hip_check(hipStreamCreate(&s), "hipStreamCreate");
        struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{s};
        const size_t view_bytes = size_t(geom->det_col.num_samples) * geom->det_row.num_samples * sizeof(float);
        DevBuf dvol(volume_bytes(*img));
        DevBuf dproj(size_t(geom->num_proj) * view_bytes);
        hip_check(hipMemcpyAsync(dproj.p, proj, dproj.bytes, hipMemcpyHostToDevice, s), "copy projections");
        launch_filter(filter_args(*plan, dproj.as<float>(), 0, 0, plan->V, 0), s);
        launch_backproject(bp_args(*plan, dvol.as<float>(), 0, plan->V, false, true, 0), s);
        hip_check(hipMemcpyAsync(vol, dvol.p, dvol.bytes, hipMemcpyDeviceToHost, s), "copy volume");
        hip_check(hipStreamSynchronize(s), "spiral reconstruction");
    });
}

extern "C" gct_status gct_plan_autotune(gct_plan *p, float *vol, uint32_t *best_zu, float *best_ms)
{
    return guarded([&] {
        if (!p || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        if (p->filt_views < p->V) fail(GCT_ERR_INVALID_ARGUMENT, "plan holds a partial filtered buffer");
        DeviceGuard dg(p->device);
        // The stream and events are released, and the plan keeps its previous
        // shape, whatever a timing launch throws.
        struct Res {
            hipStream_t s = nullptr;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            ~Res()
            {
                if (s) (void)hipStreamSynchronize(s);
                if (e0) (void)hipEventDestroy(e0);
                if (e1) (void)hipEventDestroy(e1);
                if (s) (void)hipStreamDestroy(s);
            }
        } res;
        struct Restore {
            gct_plan &p;
            uint32_t zu, vb, rows, pairs;
            int variant;
            bool rr, armed = true;
            ~Restore()
            {
                if (!armed) return;
                p.z_unroll = zu; p.view_block = vb; p.lds_rows = rows; p.lds_pairs = pairs;
                p.variant = variant; p.round_robin = rr;
            }
        } restore{*p, p->z_unroll, p->view_block, p->lds_rows, p->lds_pairs, p->variant, p->round_robin};
        hip_check(hipStreamCreate(&res.s), "hipStreamCreate");
        hip_check(hipEventCreate(&res.e0), "event");
        hip_check(hipEventCreate(&res.e1), "event");
        hipStream_t s = res.s;
        hipEvent_t e0 = res.e0, e1 = res.e1;
        // every LDS shape with its tiles dealt to the XCDs in contiguous runs and round robin
        struct Cand { int variant; uint32_t zu, vb, rows, pairs; bool rr; };
        std::vector<Cand> cands;
        for (const bool rr : {false, true})
            for (int si = 0; si < kNumLdsShapes; ++si) {
                // non-circular scans: flat panels have the one LDS shape (16, 2, 40, 40), others none
                if (!p->circular && (p->geom.kind != GCT_FLAT || si != 0)) continue;
                const LdsShape &sh = kLdsShapes[si];
                cands.push_back({GCT_BP_LDS, sh.zu, sh.vb, sh.rows, sh.pairs, rr});
            }
        cands.push_back({GCT_BP_DIRECT, 8u, 0u, 0u, 0u, false});
        cands.push_back({GCT_BP_DIRECT, 16u, 0u, 0u, 0u, false});
        // Two interleaved rounds (the second only for candidates within 10 % of
        // the best so far); a candidate keeps the better of its medians of
        // five, so drift in clocks or caches does not decide between close ones.
        std::vector<float> med(cands.size(), 1e30f);
        float best = 1e30f;
        size_t bi = 0;
        for (int round = 0; round < 2; ++round) {
            for (size_t ci = 0; ci < cands.size(); ++ci) {
                if (round == 1 && med[ci] > 1.1f * best) continue;
                const Cand &c = cands[ci];
                p->z_unroll = c.zu;
                p->variant = c.variant;
                p->view_block = c.vb;
                p->lds_rows = c.rows;
                p->lds_pairs = c.pairs;
                p->round_robin = c.rr;
                const BpLaunch a = bp_args(*p, vol, 0, p->V, false, true, 0);
                launch_backproject(a, s);   // warm-up
                float t[5];
                for (float &ms : t) {
                    hip_check(hipEventRecord(e0, s), "record");
                    launch_backproject(a, s);
                    hip_check(hipEventRecord(e1, s), "record");
                    hip_check(hipEventSynchronize(e1), "sync");
                    hip_check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
                }
                std::sort(t, t + 5);
                med[ci] = std::min(med[ci], t[2]);
                if (med[ci] < best) { best = med[ci]; bi = ci; }
            }
        }
        const Cand bc = cands[bi];
        restore.armed = false;
        p->variant = bc.variant;
        p->z_unroll = bc.zu;
        p->view_block = bc.vb;
        p->lds_rows = bc.rows;
        p->lds_pairs = bc.pairs;
        p->round_robin = bc.rr;
        const uint32_t bz = bc.zu;
        const int bv = bc.variant;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            g_tune[{p->geom.kind, p->img.x.num_samples, p->img.y.num_samples, p->img.dim == 3 ? p->img.z.num_samples : 1u,
                    p->V, p->nu, p->nv, p->opts.storage, int(p->circular), int(p->setup_f64)}] = Tuning{bz, bv, bc.vb, bc.rows, bc.pairs, bc.rr};
        }
        if (best_zu) *best_zu = bz;
        if (best_ms) *best_ms = best;
    });
}

