// Phantom arithmetic shared by the host generator (phantom.cpp) and the
// device generator (kernels.hip): seeded splitmix64 stream, Poisson sampler,
// ellipse/ellipsoid chord lengths and the detector ray of every pixel.  All
// FP64.  No reference counterpart (the reference reads projections from disk,
// example_recons.cpp).
//
// Ray conventions are the adjoint of the backprojector's mapping
// (coord_trans_proj.cuh): parallel rays {t n + s d}, n = (-sin, cos),
// d = (cos, sin), t = (c - cu) du; fan/cone rays from the source through
// src + D ew + u eu (+ v ev), arc/cylindrical rays along cos(g) ew + sin(g) eu.
#pragma once

#include "gct_amd.h"

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define GCT_HD __host__ __device__
#else
#define GCT_HD
#endif

namespace gct {
namespace phantom {

struct Rng {
    uint64_t s;
    GCT_HD uint64_t next()
    {   // splitmix64
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    GCT_HD double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Seed of the noise stream of detector row `row` (one stream per row, drawn
// along the columns in order).
GCT_HD inline uint64_t row_seed(uint64_t noise_seed, uint64_t row)
{
    return noise_seed * 0x100000001B3ull + row * 0x9E3779B97F4A7C15ull + 1;
}

// Poisson sample: multiplication method below 10, otherwise Hormann's PTRS
// transformed rejection (W. Hormann, "The transformed rejection method for
// generating Poisson random variables", 1993).
GCT_HD inline double poisson(Rng &r, double lam)
{
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        const double L = ::exp(-lam);
        double p = 1.0;
        int k = 0;
        do { ++k; p *= r.uniform(); } while (p > L);
        return double(k - 1);
    }
    const double slam = ::sqrt(lam), loglam = ::log(lam);
    const double b = 0.931 + 2.53 * slam, a = -0.059 + 0.02483 * b;
    const double invalpha = 1.1239 + 1.1328 / (b - 3.4), vr = 0.9277 - 3.6224 / (b - 2.0);
    for (;;) {
        const double U = r.uniform() - 0.5, V = r.uniform();
        const double us = 0.5 - ::fabs(U);
        const double k = ::floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (::log(V) + ::log(invalpha) - ::log(a / (us * us) + b) <= -lam + k * loglam - ::lgamma(k + 1.0))
            return k;
    }
}

// Noisy line integral -ln(max(Poisson(I0 e^-p), 1) / I0).
GCT_HD inline double add_noise(Rng &r, double p, double i0)
{
    const double n = poisson(r, i0 * ::exp(-p));
    return -::log((n > 1.0 ? n : 1.0) / i0);
}

struct Ell {
    double c[3], ax[3], R[3][3], dens;   // R rows: body axes in world coordinates
};

// Ellipse (dim 2) or ellipsoid (dim 3) from its 10 parameters
// {cx, cy, cz, ax, ay, az, a, b, g, density}; 2-D uses angle a only.
GCT_HD inline Ell unpack_one(const double *q, int dim)
{
    Ell e;
    for (int k = 0; k < 3; ++k) { e.c[k] = q[k]; e.ax[k] = q[3 + k]; }
    e.dens = q[9];
    if (dim == 2) {
        const double ca = ::cos(q[6]), sa = ::sin(q[6]);
        e.R[0][0] = ca;  e.R[0][1] = sa; e.R[0][2] = 0.0;
        e.R[1][0] = -sa; e.R[1][1] = ca; e.R[1][2] = 0.0;
        e.R[2][0] = 0.0; e.R[2][1] = 0.0; e.R[2][2] = 1.0;
        e.c[2] = 0.0;
        e.ax[2] = 1.0;
    } else {
        // Euler z-y-z: M = Rz(a) Ry(b) Rz(g); body axes are M's columns, R = M^T
        const double a = q[6], b = q[7], g = q[8];
        const double ca = ::cos(a), sa = ::sin(a), cb = ::cos(b), sb = ::sin(b), cg = ::cos(g), sg = ::sin(g);
        const double M[3][3] = {{ca * cb * cg - sa * sg, -ca * cb * sg - sa * cg, ca * sb},
                                {sa * cb * cg + ca * sg, -sa * cb * sg + ca * cg, sa * sb},
                                {-sb * cg, sb * sg, cb}};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) e.R[i][j] = M[j][i];
    }
    return e;
}

// Sum of density x chord length along the line {o + t d}.
GCT_HD inline double line_integral(const Ell *ph, uint32_t n, int dim, const double o[3], const double d[3])
{
    double acc = 0.0;
    const double dl = ::sqrt(d[0] * d[0] + d[1] * d[1] + (dim == 3 ? d[2] * d[2] : 0.0));
    for (uint32_t k = 0; k < n; ++k) {
        const Ell &e = ph[k];
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
        if (disc > 0.0 && A > 0.0) acc += e.dens * dl * ::sqrt(disc) / A;
    }
    return acc;
}

// Density at point P (sum over the shapes that contain it).
GCT_HD inline double density_at(const Ell *ph, uint32_t n, int dim, const double P[3])
{
    double acc = 0.0;
    for (uint32_t k = 0; k < n; ++k) {
        const Ell &e = ph[k];
        double s = 0.0;
        for (int i = 0; i < dim; ++i) {
            double q = 0.0;
            for (int j = 0; j < dim; ++j) q += e.R[i][j] * (P[j] - e.c[j]);
            q /= e.ax[i];
            s += q * q;
        }
        if (s <= 1.0) acc += e.dens;
    }
    return acc;
}

GCT_HD inline bool kind_is_1d(int kind)
{
    return kind == GCT_PARALLEL_1D || kind == GCT_FAN_ARC_1D || kind == GCT_FAN_LINE_1D;
}

// Ray through detector pixel (r, c) of view p.
GCT_HD inline void ray_of(int kind, const gct_proj_geom &p, double du, double dv, uint32_t r, uint32_t c, double o[3],
                          double d[3])
{
    const double uc = (double(c) - p.det_row_center) * du;
    const double vr = kind_is_1d(kind) ? 0.0 : (double(r) - p.det_col_center) * dv;
    const double D = p.src_det_dist;
    const double eu[3] = {p.det_row_dir[0], p.det_row_dir[1], p.det_row_dir[2]};
    const double ev[3] = {p.det_col_dir[0], p.det_col_dir[1], p.det_col_dir[2]};
    const double ew[3] = {p.iso_ray_dir[0], p.iso_ray_dir[1], p.iso_ray_dir[2]};
    for (int i = 0; i < 3; ++i) { o[i] = 0.0; d[i] = 0.0; }
    switch (kind) {
    case GCT_PARALLEL_1D:
    case GCT_PARALLEL_2D:
    case GCT_CONE_PARALLEL: {
        const double th = p.rot_ang;
        o[0] = -uc * ::sin(th); o[1] = uc * ::cos(th);
        d[0] = ::cos(th); d[1] = ::sin(th);
        if (kind == GCT_PARALLEL_2D) o[2] = -vr;
        if (kind == GCT_CONE_PARALLEL) o[2] = -vr * double(p.src_iso_dist) / D;
        break;
    }
    case GCT_FAN_LINE_1D:
        o[0] = p.src_pos[0]; o[1] = p.src_pos[1];
        d[0] = D * (-eu[1]) + uc * eu[0];
        d[1] = D * eu[0] + uc * eu[1];
        break;
    case GCT_FAN_ARC_1D:
        o[0] = p.src_pos[0]; o[1] = p.src_pos[1];
        d[0] = ::cos(uc) * ew[0] + ::sin(uc) * eu[0];
        d[1] = ::cos(uc) * ew[1] + ::sin(uc) * eu[1];
        break;
    case GCT_CYLINDRICAL:
        for (int i = 0; i < 3; ++i) {
            o[i] = p.src_pos[i];
            d[i] = D * (::cos(uc) * ew[i] + ::sin(uc) * eu[i]) + vr * ev[i];
        }
        break;
    default: {
        const double w[3] = {eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0]};
        for (int i = 0; i < 3; ++i) { o[i] = p.src_pos[i]; d[i] = D * w[i] + uc * eu[i] + vr * ev[i]; }
        break;
    }
    }
}

}  // namespace phantom
}  // namespace gct
