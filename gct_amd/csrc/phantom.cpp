// Synthetic phantoms for benchmarks and tests: seeded random ellipses /
// ellipsoids, their exact line integrals along every detector ray of a scan
// (FP64, cast to FP32) with optional Poisson noise, and rasterisation at voxel
// centres.  Host code; the arithmetic is shared with the device generator
// (phantom_math.hpp).  No reference counterpart (the reference reads
// projections from disk, example_recons.cpp).
#include "common.hpp"
#include "phantom_math.hpp"

#include <algorithm>
#include <cmath>
#include <numbers>
#include <thread>
#include <vector>

namespace gct {
namespace {

using phantom::Ell;
using phantom::Rng;

std::vector<Ell> unpack(const double *p, uint32_t n, int dim)
{
    std::vector<Ell> out(n);
    for (uint32_t i = 0; i < n; ++i) out[i] = phantom::unpack_one(p + 10 * size_t(i), dim);
    return out;
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
                Rng rng{phantom::row_seed(noise_seed, row)};
                for (uint32_t c = 0; c < nu; ++c) {
                    double o[3], d[3];
                    phantom::ray_of(geom->kind, geom->proj_geom[v], geom->det_row.sampling_width,
                                    geom->det_col.sampling_width, r, c, o, d);
                    double p = phantom::line_integral(ph.data(), count, dim, o, d);
                    if (i0 > 0.0) p = phantom::add_noise(rng, p, i0);
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
                    vol[(size_t(z) * ny + y) * nx + x] = static_cast<float>(phantom::density_at(ph.data(), count, dim, P));
                }
    });
}
