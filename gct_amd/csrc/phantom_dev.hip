// Device twins of the synthetic-phantom generators: the same FP64 arithmetic
// as phantom.cpp (phantom_math.hpp), run on the GPU so that benchmark inputs
// of serving and batch workloads (up to 10^9 rays) are made in milliseconds.
// No reference counterpart.
//
// Noise-free projections use one thread per detector pixel.  With Poisson
// noise each detector row is one thread that draws its row's stream column by
// column, exactly as the host generator does, so both produce the same draws.
#include "common.hpp"
#include "phantom_math.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace gct {
namespace {

using phantom::Ell;

constexpr uint32_t kMaxShapes = 64;

struct ProjectLaunch {
    const gct_proj_geom *views;
    const Ell *shapes;
    float *out;
    uint64_t pixels, rows, noise_seed;
    double du, dv, i0;
    uint32_t nu, nv, count;
    int kind, dim;
};

__device__ inline double pixel_integral(const ProjectLaunch &a, const Ell *sh, uint64_t row, uint32_t c)
{
    const uint32_t v = static_cast<uint32_t>(row / a.nv), r = static_cast<uint32_t>(row % a.nv);
    double o[3], d[3];
    phantom::ray_of(a.kind, a.views[v], a.du, a.dv, r, c, o, d);
    return phantom::line_integral(sh, a.count, a.dim, o, d);
}

__global__ __launch_bounds__(256) void phantom_project_kernel(ProjectLaunch a)
{
    __shared__ Ell sh[kMaxShapes];
    for (uint32_t i = threadIdx.x; i < a.count; i += blockDim.x) sh[i] = a.shapes[i];
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t px = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; px < a.pixels; px += stride)
        a.out[px] = static_cast<float>(pixel_integral(a, sh, px / a.nu, static_cast<uint32_t>(px % a.nu)));
}

__global__ __launch_bounds__(64) void phantom_project_noisy_kernel(ProjectLaunch a)
{
    __shared__ Ell sh[kMaxShapes];
    for (uint32_t i = threadIdx.x; i < a.count; i += blockDim.x) sh[i] = a.shapes[i];
    __syncthreads();
    const uint64_t row = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= a.rows) return;
    phantom::Rng rng{phantom::row_seed(a.noise_seed, row)};
    for (uint32_t c = 0; c < a.nu; ++c)
        a.out[row * a.nu + c] = static_cast<float>(phantom::add_noise(rng, pixel_integral(a, sh, row, c), a.i0));
}

struct RasterLaunch {
    const Ell *shapes;
    float *out;
    double x0, y0, z0, dx, dy, dz;
    uint64_t voxels;
    uint32_t nx, ny, count;
    int dim;
};

__global__ __launch_bounds__(256) void phantom_rasterize_kernel(RasterLaunch a)
{
    __shared__ Ell sh[kMaxShapes];
    for (uint32_t i = threadIdx.x; i < a.count; i += blockDim.x) sh[i] = a.shapes[i];
    __syncthreads();
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.voxels; i += stride) {
        const uint32_t x = static_cast<uint32_t>(i % a.nx);
        const uint64_t yz = i / a.nx;
        const uint32_t y = static_cast<uint32_t>(yz % a.ny), z = static_cast<uint32_t>(yz / a.ny);
        const double P[3] = {a.x0 + x * a.dx, a.y0 + y * a.dy, a.dim == 3 ? a.z0 + z * a.dz : 0.0};
        a.out[i] = static_cast<float>(phantom::density_at(sh, a.count, a.dim, P));
    }
}

// Device copy of a host array, freed when the owner leaves scope (after the
// stream has been synchronised).
template <class T>
struct Staged {
    T *p = nullptr;
    Staged(const T *host, size_t n, hipStream_t s)
    {
        if (n == 0) return;
        hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
        hip_check(hipMemcpyAsync(p, host, n * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync");
    }
    ~Staged() { if (p) (void)hipFree(p); }
    Staged(const Staged &) = delete;
    Staged &operator=(const Staged &) = delete;
};

std::vector<Ell> unpack_all(const double *params, uint32_t count, int dim)
{
    if (count > kMaxShapes) fail(GCT_ERR_INVALID_ARGUMENT, "at most 64 shapes per phantom on the device");
    std::vector<Ell> e(count);
    for (uint32_t i = 0; i < count; ++i) e[i] = phantom::unpack_one(params + 10 * size_t(i), dim);
    return e;
}

uint32_t grid_for(uint64_t items, uint32_t block)
{
    const uint64_t g = (items + block - 1) / block;
    return static_cast<uint32_t>(g < 65536u ? (g ? g : 1u) : 65536u);
}

}  // namespace
}  // namespace gct

using namespace gct;

extern "C" gct_status gct_phantom_project_device(const gct_ct_geom *geom, int dim, const double *params, uint32_t count,
                                                 double i0, uint64_t noise_seed, float *proj_dev, void *stream)
{
    return guarded([&] {
        if (!geom || !params || !proj_dev) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        validate_geom(*geom);
        if ((dim == 2) != is_1d(geom->kind)) fail(GCT_ERR_INVALID_ARGUMENT, "phantom dimension does not match the geometry");
        if (!(i0 >= 0.0)) fail(GCT_ERR_INVALID_ARGUMENT, "I0 must be >= 0");
        const std::vector<Ell> shapes = unpack_all(params, count, dim);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        Staged<gct_proj_geom> views(geom->proj_geom, geom->num_proj, s);
        Staged<Ell> sh(shapes.data(), shapes.size(), s);
        ProjectLaunch a{};
        a.views = views.p;
        a.shapes = sh.p;
        a.out = proj_dev;
        a.nu = geom->det_row.num_samples;
        a.nv = is_1d(geom->kind) ? 1u : geom->det_col.num_samples;
        a.rows = static_cast<uint64_t>(geom->num_proj) * a.nv;
        a.pixels = a.rows * a.nu;
        a.noise_seed = noise_seed;
        a.du = geom->det_row.sampling_width;
        a.dv = geom->det_col.sampling_width;
        a.i0 = i0;
        a.count = count;
        a.kind = geom->kind;
        a.dim = dim;
        if (i0 > 0.0)
            hipLaunchKernelGGL(phantom_project_noisy_kernel, dim3(static_cast<uint32_t>((a.rows + 63) / 64)), dim3(64), 0, s, a);
        else
            hipLaunchKernelGGL(phantom_project_kernel, dim3(grid_for(a.pixels, 256)), dim3(256), 0, s, a);
        hip_check(hipGetLastError(), "phantom projection launch");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

extern "C" gct_status gct_phantom_rasterize_device(const gct_image_geom *img, const double *params, uint32_t count,
                                                   float *vol_dev, void *stream)
{
    return guarded([&] {
        if (!img || !params || !vol_dev) fail(GCT_ERR_INVALID_ARGUMENT, "null argument");
        const int dim = static_cast<int>(img->dim);
        if (dim != 2 && dim != 3) fail(GCT_ERR_INVALID_ARGUMENT, "image dimension must be 2 or 3");
        const std::vector<Ell> shapes = unpack_all(params, count, dim);
        const hipStream_t s = static_cast<hipStream_t>(stream);
        Staged<Ell> sh(shapes.data(), shapes.size(), s);
        RasterLaunch a{};
        a.shapes = sh.p;
        a.out = vol_dev;
        a.nx = img->x.num_samples;
        a.ny = img->y.num_samples;
        const uint32_t nz = dim == 3 ? img->z.num_samples : 1u;
        a.dx = img->x.sampling_width;
        a.dy = img->y.sampling_width;
        a.dz = dim == 3 ? img->z.sampling_width : 0.0;
        a.x0 = img->mid_point[0] - (a.nx - 1.0) * a.dx / 2.0;
        a.y0 = img->mid_point[1] - (a.ny - 1.0) * a.dy / 2.0;
        a.z0 = dim == 3 ? img->mid_point[2] - (nz - 1.0) * a.dz / 2.0 : 0.0;
        a.voxels = static_cast<uint64_t>(a.nx) * a.ny * nz;
        a.count = count;
        a.dim = dim;
        hipLaunchKernelGGL(phantom_rasterize_kernel, dim3(grid_for(a.voxels, 256)), dim3(256), 0, s, a);
        hip_check(hipGetLastError(), "phantom rasterisation launch");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}
