// Synthetic phantoms for benchmarks and tests: seeded random ellipses /
// ellipsoids, their exact line integrals along every detector ray of a scan
// (FP64, cast to FP32) with optional Poisson noise, and rasterisation at voxel
// centres.  Host code; no reference counterpart (the reference reads
// projections from disk, example_recons.cpp).
//
// Ray conventions are the adjoint of the backprojector's mapping
// (coord_trans_proj.cuh): parallel rays {t n + s d}, n = (-sin, cos),
// d = (cos, sin), t = (c - cu) du; fan/cone rays from the source through
// src + D ew + u eu (+ v ev), arc/cylindrical rays along cos(g) ew + sin(g) eu.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <thread>

namespace gct {
namespace {

struct Rng {
    uint64_t s;
    uint64_t next()
    {   // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Poisson sample: multiplication method below 10, otherwise Hormann's PTRS
// transformed rejection (W. Hormann, "The transformed rejection method for
// generating Poisson random variables", 1993).
double poisson(Rng &r, double lam)
{
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double L = std::exp(-lam);
        double p = 1.0;
        int k = 0;
        do { ++k; p *= r.uniform(); } while (p > L);
        return double(k - 1);
    }
    const double slam = std::sqrt(lam), loglam = std::log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = r.uniform() - 0.5, V = r.uniform();
        const double us = 0.5 - std::fabs(U);
        const double k = std::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (std::log(V) + std::log(invalpha) - std::log(a / (us * us) + b) <= -lam + k * loglam - std::lgamma(k + 1.0))
            return k;
    }
}

struct Ell {
    double c[3], ax[3], R[3][3], dens;   // R rows: body axes in world coordinates
};

std::vector<Ell> unpack(const double *p, uint32_t n, int dim)
{
    std::vector<Ell> out(n);
    for (uint32_t i = 0; i < n; ++i) {
        const double *q = p + 10 * size_t(i);
        Ell &e = out[i];
        for (int k = 0; k < 3; ++k) { e.c[k] = q[k]; e.ax[k] = q[3 + k]; }
        e.dens = q[9];
        if (dim == 2) {
            const double ca = std::cos(q[6]), sa = std::sin(q[6]);
            const double R[3][3] = {{ca, sa, 0}, {-sa, ca, 0}, {0, 0, 1}};
            std::copy(&R[0][0], &R[0][0] + 9, &e.R[0][0]);
            e.c[2] = 0.0;
            e.ax[2] = 1.0;
        } else {
            // Euler z-y-z: M = Rz(a) Ry(b) Rz(g); body axes are M's columns, R = M^T
            const double a = q[6], b = q[7], g = q[8];
            const double ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b), cg = std::cos(g), sg = std::sin(g);
            const double M[3][3] = {{ca * cb * cg - sa * sg, -ca * cb * sg - sa * cg, ca * sb},
                                    {sa * cb * cg + ca * sg, -sa * cb * sg + ca * cg, sa * sb},
                                    {-sb * cg, sb * sg, cb}};
            for (int i2 = 0; i2 < 3; ++i2)
                for (int j = 0; j < 3; ++j) e.R[i2][j] = M[j][i2];
        }
    }
    return out;
}

// Sum of density x chord length along the line {o + t d}.
double line_integral(const std::vector<Ell> &ph, int dim, const double o[3], const double d[3])
{
    double acc = 0.0;
    const double dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + (dim == 3 ? d[2] * d[2] : 0.0));
    for (const Ell &e : ph) {
        double p[3], q[3];
        for (int i = 0; i < 3; ++i) {
            double pp = 0.0, qq = 0.0;
            for (int j = 0; j < dim; ++j) {
                pp += e.R[i][j] * (o[j] - e.c[j]);
                qq += e.R[i][j] * d[j];
            }
            p[i] = pp / e.ax[i];
            q[i] = qq / e.ax[i];
        }
        double A = 0.0, B = 0.0, C = -1.0;
        for (int i = 0; i < dim; ++i) { A += q[i] * q[i]; B += 2.0 * p[i] * q[i]; C += p[i] * p[i]; }
        const double disc = B * B - 4.0 * A * C;
        if (disc > 0.0 && A > 0.0) acc += e.dens * dl * std::sqrt(disc) / A;
    }
    return acc;
}

void ray_of(const gct_ct_geom &g, uint32_t v, uint32_t r, uint32_t c, double o[3], double d[3])
{
    const gct_proj_geom &p = g.proj_geom[v];
    const double uc = (double(c) - p.det_row_center) * g.det_row.sampling_width;
    const double vr = is_1d(g.kind) ? 0.0 : (double(r) - p.det_col_center) * g.det_col.sampling_width;
    const double D = p.src_det_dist;
    const double eu[3] = {p.det_row_dir[0], p.det_row_dir[1], p.det_row_dir[2]};
    const double ev[3] = {p.det_col_dir[0], p.det_col_dir[1], p.det_col_dir[2]};
    const double ew[3] = {p.iso_ray_dir[0], p.iso_ray_dir[1], p.iso_ray_dir[2]};
    for (int i = 0; i < 3; ++i) { o[i] = 0.0; d[i] = 0.0; }
    switch (g.kind) {
    case GCT_PARALLEL_1D:
    case GCT_PARALLEL_2D:
    case GCT_CONE_PARALLEL: {
        const double th = p.rot_ang;
        o[0] = -uc * std::sin(th); o[1] = uc * std::cos(th);
        d[0] = std::cos(th); d[1] = std::sin(th);
        if (g.kind == GCT_PARALLEL_2D) o[2] = -vr;
        if (g.kind == GCT_CONE_PARALLEL) o[2] = -vr * double(p.src_iso_dist) / D;
        break;
    }
    case GCT_FAN_LINE_1D:
        o[0] = p.src_pos[0]; o[1] = p.src_pos[1];
        d[0] = D * (-eu[1]) + uc * eu[0];
        d[1] = D * eu[0] + uc * eu[1];
        break;
    case GCT_FAN_ARC_1D:
        o[0] = p.src_pos[0]; o[1] = p.src_pos[1];
        d[0] = std::cos(uc) * ew[0] + std::sin(uc) * eu[0];
        d[1] = std::cos(uc) * ew[1] + std::sin(uc) * eu[1];
        break;
    case GCT_CYLINDRICAL:
        for (int i = 0; i < 3; ++i) {
            o[i] = p.src_pos[i];
            d[i] = D * (std::cos(uc) * ew[i] + std::sin(uc) * eu[i]) + vr * ev[i];
        }
        break;
    default: {
        const double w[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0]};
        for (int i = 0; i < 3; ++i) { o[i] = p.src_pos[i]; d[i] = D * w[i] + uc * eu[i] + vr * ev[i]; }
        break;
    }
    }
}

}  // namespace
}  // namespace gct

using namespace gct;

extern "C" gct_status gct_phantom_random(uint64_t seed, uint32_t count, int dim, double radius, double axis_lo,
                                         double axis_hi, double dens_lo, double dens_hi, double *params)
{
    return guarded([&] {
        if (!params || (dim != 2 && dim != 3) || !(radius > 0.0)) fail(GCT_ERR_INVALID_ARGUMENT, "invalid phantom request");
        Rng r{seed};
        const double pi = std::numbers::pi;
        for (uint32_t i = 0; i < count; ++i) {
            double *q = params + 10 * size_t(i);
            // centre uniform in the disk / ball of radius 0.8 R
            double c[3] = {0, 0, 0};
            for (;;) {
                double s2 = 0.0;
                for (int k = 0; k < dim; ++k) { c[k] = (2.0 * r.uniform() - 1.0); s2 += c[k] * c[k]; }
                if (s2 <= 1.0) break;
            }
            for (int k = 0; k < 3; ++k) q[k] = 0.8 * radius * c[k];
            for (int k = 0; k < 3; ++k) q[3 + k] = (axis_lo + (axis_hi - axis_lo) * r.uniform()) * radius;
            if (dim == 2) {
                q[2] = 0.0; q[5] = 1.0;
                q[6] = pi * r.uniform(); q[7] = 0.0; q[8] = 0.0;
            } else {
                // uniform random rotation: z-y-z Euler with cos(beta) uniform
                q[6] = 2.0 * pi * r.uniform();
                q[7] = std::acos(2.0 * r.uniform() - 1.0);
                q[8] = 2.0 * pi * r.uniform();
            }
            q[9] = dens_lo + (dens_hi - dens_lo) * r.uniform();
        }
    });
}

extern "C" gct_status gct_phantom_project(const gct_ct_geom *geom, int dim, const double *params, uint32_t count,
                                          double i0, uint64_t noise_seed, float *proj, int nthreads)
{
    return guarded([&] {
        if (!geom || !params || !proj) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        validate_geom(*geom);
        if ((dim == 2) != is_1d(geom->kind)) fail(GCT_ERR_INVALID_ARGUMENT, "phantom dimension does not match the geometry");
        const std::vector<Ell> ph = unpack(params, count, dim);
        const uint32_t nv = is_1d(geom->kind) ? 1u : geom->det_col.num_samples, nu = geom->det_row.num_samples;
        const size_t rows = size_t(geom->num_proj) * nv;
        int nt = nthreads > 0 ? nthreads : int(std::max(1u, std::thread::hardware_concurrency()));
        nt = std::min<int>(nt, 64);
        auto work = [&](int t) {
            for (size_t row = size_t(t); row < rows; row += size_t(nt)) {
                const uint32_t v = uint32_t(row / nv), r = uint32_t(row % nv);
                Rng rng{noise_seed * 0x100000001B3ull + row * 0x9E3779B97F4A7C15ull + 1};
                for (uint32_t c = 0; c < nu; ++c) {
                    double o[3], d[3];
                    ray_of(*geom, v, r, c, o, d);
                    double p = line_integral(ph, dim, o, d);
                    if (i0 > 0.0) {
                        const double n = poisson(rng, i0 * std::exp(-p));
                        p = -std::log(std::max(n, 1.0) / i0);
                    }
                    proj[row * nu + c] = static_cast<float>(p);
                }
            }
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work, t);
        work(0);
        for (auto &x : th) x.join();
    });
}

extern "C" gct_status gct_phantom_rasterize(const gct_image_geom *img, const double *params, uint32_t count, float *vol)
{
    return guarded([&] {
        if (!img || !params || !vol) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        const int dim = int(img->dim);
        const std::vector<Ell> ph = unpack(params, count, dim);
        const uint32_t nx = img->x.num_samples, ny = img->y.num_samples, nz = dim == 3 ? img->z.num_samples : 1u;
        const double x0 = img->mid_point[0] - (nx - 1.0) * img->x.sampling_width / 2.0;
        const double y0 = img->mid_point[1] - (ny - 1.0) * img->y.sampling_width / 2.0;
        const double z0 = dim == 3 ? img->mid_point[2] - (nz - 1.0) * img->z.sampling_width / 2.0 : 0.0;
        for (uint32_t z = 0; z < nz; ++z)
            for (uint32_t y = 0; y < ny; ++y)
                for (uint32_t x = 0; x < nx; ++x) {
                    const double P[3] = {x0 + x * double(img->x.sampling_width), y0 + y * double(img->y.sampling_width),
                                         dim == 3 ? z0 + z * double(img->z.sampling_width) : 0.0};
                    double acc = 0.0;
                    for (const Ell &e : ph) {
                        double s = 0.0;
                        for (int i = 0; i < dim; ++i) {
                            double q = 0.0;
                            for (int j = 0; j < dim; ++j) q += e.R[i][j] * (P[j] - e.c[j]);
                            q /= e.ax[i];
                            s += q * q;
                        }
                        if (s <= 1.0) acc += e.dens;
                    }
                    vol[(size_t(z) * ny + y) * nx + x] = static_cast<float>(acc);
                }
    });
}
