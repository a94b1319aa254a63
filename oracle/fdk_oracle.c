/*
 * CPU oracle for the FDK reconstruction path -- TEST INFRASTRUCTURE ONLY.
 * The product never links or loads this file; see oracle.h.
 *
 * Built with -ffp-contract=off so FP32 expressions round as written.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAX_TAPS 8192
/* std::numbers::pi_v<float> and the reference's pi() = 3.141592654f round to
 * the same float. */
#define OR_PI_F 3.14159274101257324f

/* ---- a small parallel-for over pthreads ---------------------------------- */

typedef void (*or_task_fn)(void *ctx, size_t idx);
typedef struct { or_task_fn fn; void *ctx; size_t n; size_t next; pthread_mutex_t mu; } or_pool;

static void *or_worker(void *arg)
{
    or_pool *p = (or_pool *)arg;
    for (;;) {
        pthread_mutex_lock(&p->mu);
        size_t i = p->next++;
        pthread_mutex_unlock(&p->mu);
        if (i >= p->n) break;
        p->fn(p->ctx, i);
    }
    return NULL;
}

static void or_parallel_for(size_t n, int nthreads, or_task_fn fn, void *ctx)
{
    if (nthreads <= 1 || n < 2) {
        for (size_t i = 0; i < n; ++i) fn(ctx, i);
        return;
    }
    if (nthreads > 256) nthreads = 256;
    or_pool p = { fn, ctx, n, 0 };
    pthread_mutex_init(&p.mu, NULL);
    pthread_t th[256];
    int started = 0;
    for (int t = 0; t < nthreads; ++t)
        if (pthread_create(&th[t], NULL, or_worker, &p) == 0) ++started;
    if (started == 0) or_worker(&p);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    pthread_mutex_destroy(&p.mu);
}

/* ---- rotation analysis (src/cpp/geom/rotation_geom.cpp, utils.h) --------- */

/* Utils::wrapFrom0To2PI(Radians), utils.h:338-348 */
float or_wrap_0_2pi(float a)
{
    const float pi2 = 2.0f * OR_PI_F;
    while (a < 0.0f) a += pi2;
    return (a >= pi2) ? (float)remainder((double)a, (double)pi2) : a;
}

/* Utils::quadrantIndex(Radians), utils.h:370-389 */
unsigned or_quadrant(float ang)
{
    float a = or_wrap_0_2pi(ang);
    const float pi = OR_PI_F;
    if (a >= 0.0f && a < pi / 2.0f) return 1u;
    if (a >= pi / 2.0f && a < pi) return 2u;
    if (a >= pi && a < (3.0f / 2.0f) * pi) return 3u;
    return 4u;
}

/* Detail::rotAng, rotation_geom.h:32-36 and rotation_geom.cpp:255-295 */
static float or_rot_ang(const or_scan *s, unsigned v)
{
    const or_view *g = &s->views[v];
    if (s->kind == OR_PAR1D || s->kind == OR_PAR2D || s->kind == OR_CONEPAR)
        return or_wrap_0_2pi(g->rot_ang);
    return or_wrap_0_2pi(atan2f(g->src_pos[1], g->src_pos[0]));
}

/* Detail::rotDir, rotation_geom.cpp:7-43; 0 = invalid */
static int or_rot_dir(const float *ang, unsigned n)
{
    if (n <= 1u) return 0;
    float a = or_wrap_0_2pi(ang[0]), b = or_wrap_0_2pi(ang[1]);
    unsigned qa = or_quadrant(a), qb = or_quadrant(b);
    if (qa == 1u && qb == 4u) return -1;
    if (qa == 4u && qb == 1u) return 1;
    float diff = b - a;
    if (diff > 0.0f) return 1;
    if (diff < 0.0f) return -1;
    return 0;
}

int or_rotation_geom(const or_scan *s, or_rot *out, float *rot_ang, float *rel_ang)
{
    unsigned n = s->nviews;
    if (n == 0) return -1;
    float *ang = (float *)malloc(sizeof(float) * n);
    float *rel = (float *)malloc(sizeof(float) * n);
    for (unsigned v = 0; v < n; ++v) ang[v] = or_rot_ang(s, v);
    int dir = or_rot_dir(ang, n);
    if (dir == 0) { free(ang); free(rel); return -1; }
    /* relRotAngles, rotation_geom.cpp:45-74 */
    float first = ang[0];
    for (unsigned v = 0; v < n; ++v)
        rel[v] = or_wrap_0_2pi(dir < 0 ? first - ang[v] : ang[v] - first);
    /* rotAngIncr, :76-91 */
    float sum = 0.0f;
    for (unsigned v = 0; v + 1 < n; ++v) sum = sum + or_wrap_0_2pi(rel[v + 1] - rel[v]);
    float incr = sum / (float)(n - 1);
    /* rotAngRange, :93-104 */
    float mx = rel[0];
    for (unsigned v = 1; v < n; ++v) if (rel[v] > mx) mx = rel[v];
    float range = incr + mx;
    /* numProjPerRot, :106-134 */
    unsigned nppr = n;
    if (rel[0] != 0.0f) {
        nppr = 0;  /* the reference returns nullopt; getRotationGeom would throw on .value() */
    } else {
        for (unsigned v = 1; v < n; ++v) {
            float r = rel[v];
            if (fabsf(r) < 0.000001f) r = 0.0f;
            if (r == 0.0f) { nppr = v; break; }
        }
    }
    if (out) { out->dir = dir; out->incr = incr; out->range = range; out->nppr = nppr; }
    if (rot_ang) memcpy(rot_ang, ang, sizeof(float) * n);
    if (rel_ang) memcpy(rel_ang, rel, sizeof(float) * n);
    free(ang); free(rel);
    return 0;
}

static int or_rotation_f32(const or_scan *s, int *dir, float *incr, float *range, float *rel)
{
    or_rot r;
    if (or_rotation_geom(s, &r, NULL, rel) != 0) return -1;
    *dir = r.dir; *incr = r.incr; *range = r.range;
    return 0;
}

/* The same analysis in FP64 (accuracy form). */
static double or_wrap_d(double a)
{
    const double pi2 = 2.0 * M_PI;
    while (a < 0.0) a += pi2;
    return (a >= pi2) ? remainder(a, pi2) : a;
}

static unsigned or_quadrant_d(double ang)
{
    double a = or_wrap_d(ang);
    if (a >= 0.0 && a < M_PI / 2.0) return 1u;
    if (a >= M_PI / 2.0 && a < M_PI) return 2u;
    if (a >= M_PI && a < 1.5 * M_PI) return 3u;
    return 4u;
}

int or_rotation_geom_f64(const or_scan *s, int *dir_out, double *incr_out, double *range_out, double *rel_out)
{
    unsigned n = s->nviews;
    if (n <= 1) return -1;
    double *ang = (double *)malloc(sizeof(double) * n);
    for (unsigned v = 0; v < n; ++v) {
        const or_view *g = &s->views[v];
        if (s->kind == OR_PAR1D || s->kind == OR_PAR2D || s->kind == OR_CONEPAR)
            ang[v] = or_wrap_d((double)g->rot_ang);
        else
            ang[v] = or_wrap_d(atan2((double)g->src_pos[1], (double)g->src_pos[0]));
    }
    double a = ang[0], b = ang[1];
    unsigned qa = or_quadrant_d(a), qb = or_quadrant_d(b);
    int dir;
    if (qa == 1u && qb == 4u) dir = -1;
    else if (qa == 4u && qb == 1u) dir = 1;
    else if (b - a > 0.0) dir = 1;
    else if (b - a < 0.0) dir = -1;
    else { free(ang); return -1; }
    double mx = 0.0, sum = 0.0, prev = 0.0;
    for (unsigned v = 0; v < n; ++v) {
        double r = or_wrap_d(dir < 0 ? ang[0] - ang[v] : ang[v] - ang[0]);
        if (rel_out) rel_out[v] = r;
        if (v > 0) sum += or_wrap_d(r - prev);
        if (r > mx) mx = r;
        prev = r;
    }
    double incr = sum / (double)(n - 1);
    *dir_out = dir; *incr_out = incr; *range_out = incr + mx;
    free(ang);
    return 0;
}

static int or_rotation_f64(const or_scan *s, int *dir, double *incr, double *range, double *rel)
{
    return or_rotation_geom_f64(s, dir, incr, range, rel);
}

/* ---- ramp-filter taps (src/cpp/recon/ramp_filter.{h,cpp}, proj_filt_kernel.cpp) */

static double or_M(double t)
{
    return (t == 0.0) ? 0.5 : (sin(t) / t + (cos(t) - 1.0) / (t * t));
}

void or_ramp_taps(int arc, unsigned n, float width, float cutoff, int window, float *taps32, double *taps64)
{
    double a = 1.0;                                   /* RampFilter::apodFact, ramp_filter.cpp:5-20 */
    if (window == OR_WIN_HANN) a = 0.5;
    else if (window == OR_WIN_HAMMING) a = 0.54;
    double c = (double)cutoff, sw = (double)width, pi = M_PI;
    int half = (int)n / 2;
    for (unsigned k = 0; k < n; ++k) {
        int m = (int)k - half;                          /* ramp_filter.h:36-38 */
        double nd = (double)m, h;
        if (window == OR_WIN_SHEPP_LOGAN) {
            /* Shepp & Logan (1974) kernel, in the reference's tap scaling
             * (Ram-Lak h(0) = 1/(4 tau)); an extension with no reference parity. */
            h = -2.0 / (pi * pi * sw * (4.0 * nd * nd - 1.0));
        } else {
            h = ((c * c) / (2.0 * sw)) * (a * or_M(c * pi * nd) +
                                         ((1 - a) / 2.0) * or_M(pi * c * nd + pi) +
                                         ((1 - a) / 2.0) * or_M(pi * c * nd - pi));
        }
        float hf = (float)h;
        if (arc && m != 0) {
            /* proj_filt_kernel.cpp:19-45: weight the FP32 tap by (n dg / sin n dg)^2 in FP64 */
            double beta = nd * sw;
            double wgt = pow(beta / sin(beta), 2);
            if (taps32) taps32[k] = (float)(wgt * (double)hf);
            if (taps64) taps64[k] = wgt * h;
        } else {
            if (taps32) taps32[k] = hf;
            if (taps64) taps64[k] = h;
        }
    }
}

/* ---- FP64 FFT used to evaluate the convolution exactly ------------------- */

static void or_fft(double *re, double *im, unsigned L, int inverse)
{
    for (unsigned i = 1, j = 0; i < L; ++i) {
        unsigned bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (unsigned len = 2; len <= L; len <<= 1) {
        double ang = 2.0 * M_PI / (double)len * (inverse ? 1.0 : -1.0);
        for (unsigned i = 0; i < L; i += len) {
            for (unsigned j = 0; j < len / 2; ++j) {
                double wr = cos(ang * j), wi = sin(ang * j);
                unsigned a = i + j, b = i + j + len / 2;
                double xr = re[b] * wr - im[b] * wi;
                double xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr; im[b] = im[a] - xi;
                re[a] += xr; im[a] += xi;
            }
        }
    }
}

static unsigned or_next_pow2(unsigned n)
{
    unsigned v = 1;
    while (v < n) v *= 2;
    return v;
}

typedef struct {
    unsigned n, L, nrows;
    const double *Hre, *Him;
    const float *in32; float *out32;
    const double *in64; double *out64;
} or_fctx;

/* Two rows per complex FFT: z = a + i b, IFFT(FFT(z) H) = (a*h) + i (b*h)
 * because h is real. */
static void or_filter_task(void *ctx, size_t pair)
{
    or_fctx *c = (or_fctx *)ctx;
    unsigned L = c->L, n = c->n;
    size_t ra = 2 * pair, rb = 2 * pair + 1;
    double *re = (double *)calloc(L, sizeof(double));
    double *im = (double *)calloc(L, sizeof(double));
    for (unsigned j = 0; j < n; ++j) {
        re[j] = c->in32 ? (double)c->in32[ra * n + j] : c->in64[ra * n + j];
        if (rb < c->nrows) im[j] = c->in32 ? (double)c->in32[rb * n + j] : c->in64[rb * n + j];
    }
    or_fft(re, im, L, 0);
    for (unsigned k = 0; k < L; ++k) {
        double xr = re[k] * c->Hre[k] - im[k] * c->Him[k];
        double xi = re[k] * c->Him[k] + im[k] * c->Hre[k];
        re[k] = xr; im[k] = xi;
    }
    or_fft(re, im, L, 1);
    for (unsigned j = 0; j < n; ++j) {
        double ya = re[j] / (double)L, yb = im[j] / (double)L;
        if (c->out32) {
            c->out32[ra * n + j] = (float)ya;
            if (rb < c->nrows) c->out32[rb * n + j] = (float)yb;
        } else {
            c->out64[ra * n + j] = ya;
            if (rb < c->nrows) c->out64[rb * n + j] = yb;
        }
    }
    free(re); free(im);
}

static void or_filter_common(unsigned nrows, unsigned n, const double *taps,
                             const float *in32, float *out32, const double *in64, double *out64, int nthreads)
{
    unsigned L = or_next_pow2(2u * n);                  /* ramp_filtering_fft.cuh:36-48 */
    double *hre = (double *)calloc(L, sizeof(double));
    double *him = (double *)calloc(L, sizeof(double));
    int half = (int)n / 2;
    /* tap n = k - floor(N/2) placed at index n mod L: the circular shift that
     * replaces the reference's extraction offset (ramp_filtering_fft.cu:56-60) */
    for (unsigned k = 0; k < n; ++k) {
        int m = (int)k - half;
        hre[(unsigned)((m + (int)L) % (int)L)] = taps[k];
    }
    or_fft(hre, him, L, 0);
    or_fctx c = { n, L, nrows, hre, him, in32, out32, in64, out64 };
    or_parallel_for((nrows + 1) / 2, nthreads, or_filter_task, &c);
    free(hre); free(him);
}

void or_filter_rows_f32(unsigned nrows, unsigned n, const float *taps, const float *in, float *out, int nthreads)
{
    double *t = (double *)malloc(sizeof(double) * n);
    for (unsigned k = 0; k < n; ++k) t[k] = (double)taps[k];
    or_filter_common(nrows, n, t, in, out, NULL, NULL, nthreads);
    free(t);
}

void or_filter_rows_f64(unsigned nrows, unsigned n, const double *taps, const double *in, double *out, int nthreads)
{
    or_filter_common(nrows, n, taps, NULL, NULL, in, out, nthreads);
}

/* O(N^2) restatement of the same convolution, used only to pin the FFT path. */
void or_filter_rows_direct_f64(unsigned nrows, unsigned n, const double *taps, const double *in, double *out)
{
    int half = (int)n / 2;
    for (unsigned r = 0; r < nrows; ++r)
        for (unsigned c = 0; c < n; ++c) {
            double acc = 0.0;
            for (unsigned j = 0; j < n; ++j) {
                int k = (int)c - (int)j + half;
                if (k >= 0 && k < (int)n) acc += in[(size_t)r * n + j] * taps[k];
            }
            out[(size_t)r * n + c] = acc;
        }
}

/* ---- precision-generic stages -------------------------------------------- */

/* ---- spiral view windows (spiral_ct_geom.h:30-95, fdk_spiral.cuh:53-70) ----
 * SpiralCTGeom's checks: projections per rotation != 0, source z sorted
 * (non-decreasing or non-increasing), pitch |z(0) - z(nppr)| / 2 pi > 0.
 * Slice s sits at leftEndPointPos(mz, nz, dz) + s dz (FP32); its centre view is
 * the first view with the smallest |src z - slice z| (sliceProjIndex); the
 * window [mid - nppr/2, + nppr) in unsigned arithmetic is used only when
 * first > 0 and last < V (fdk_spiral.cuh:68), otherwise the slice stays 0. */
int or_spiral_windows(const or_scan *s, const or_image *img, unsigned *win)
{
    or_rot r;
    unsigned V = s->nviews;
    if (img->dim != 3 || or_rotation_geom(s, &r, NULL, NULL) != 0) return -1;
    if (r.nppr == 0 || r.nppr >= V) return -1;
    int asc = 1, desc = 1;
    for (unsigned v = 1; v < V; ++v) {
        if (s->views[v].src_pos[2] < s->views[v - 1].src_pos[2]) asc = 0;
        if (s->views[v].src_pos[2] > s->views[v - 1].src_pos[2]) desc = 0;
    }
    if (!asc && !desc) return -1;
    float pitch = fabsf(s->views[0].src_pos[2] - s->views[r.nppr].src_pos[2]) / (2.0f * OR_PI_F);
    if (!(pitch > 0.0f)) return -1;
    float first_pos = img->mz - (((float)img->nz - 1.0f) * img->dz / 2.0f);
    for (unsigned z = 0; z < img->nz; ++z) {
        float sp = first_pos + (float)z * img->dz;
        unsigned mid = 0;
        float best = fabsf(s->views[0].src_pos[2] - sp);
        for (unsigned v = 1; v < V; ++v) {
            float d = fabsf(s->views[v].src_pos[2] - sp);
            if (d < best) { best = d; mid = v; }
        }
        unsigned first = mid - r.nppr / 2u;
        unsigned last = first + r.nppr;
        if (first > 0u && last < V) { win[2 * z] = first; win[2 * z + 1] = last; }
        else { win[2 * z] = 0u; win[2 * z + 1] = 0u; }
    }
    return 0;
}

#define REAL float
#define SFX(x) x##_f32
#define RSQRT sqrtf
#define RCOS cosf
#define RSIN sinf
#define RATAN2 atan2f
#define RFLOOR floorf
#define RPOW powf
#define PI_R 3.141592654f
#define OR_FILTER_CALL(x) x
#define OR_TAPS taps32
#include "oracle_impl.inc"
#undef REAL
#undef SFX
#undef RSQRT
#undef RCOS
#undef RSIN
#undef RATAN2
#undef RFLOOR
#undef RPOW
#undef PI_R
#undef OR_TAPS

#define REAL double
#define SFX(x) x##_f64
#define RSQRT sqrt
#define RCOS cos
#define RSIN sin
#define RATAN2 atan2
#define RFLOOR floor
#define RPOW pow
#define PI_R M_PI
#define OR_TAPS taps64
#include "oracle_impl.inc"

/* ---- Joseph forward projection (FP64) ------------------------------------ */

/* Bilinear in-plane sample of slice/plane data with zero outside. */
static double or_bilin(const double *vol, const or_image *im, int axis, long plane, double fa, double fb)
{
    /* axis is the driving axis; (fa, fb) are fractional indices of the two others
     * in increasing axis order. */
    long n[3] = { (long)im->nx, (long)im->ny, (long)((im->dim == 3) ? im->nz : 1u) };
    int oa = (axis == 0) ? 1 : 0, ob = (axis == 2) ? 1 : 2;
    double ia = floor(fa), ib = floor(fb);
    double wa = fa - ia, wb = fb - ib;
    long a0 = (long)ia, b0 = (long)ib;
    double acc = 0.0;
    for (int da = 0; da < 2; ++da)
        for (int db = 0; db < 2; ++db) {
            long a = a0 + da, b = b0 + db;
            if (a < 0 || b < 0 || a >= n[oa] || b >= n[ob]) continue;
            long idx[3];
            idx[axis] = plane; idx[oa] = a; idx[ob] = b;
            double w = (da ? wa : 1.0 - wa) * (db ? wb : 1.0 - wb);
            acc += w * vol[((size_t)idx[2] * n[1] + idx[1]) * n[0] + idx[0]];
        }
    return acc;
}

static double or_bilin2(const double *img, const or_image *im, int axis, long plane, double f)
{
    long n[2] = { (long)im->nx, (long)im->ny };
    int o = 1 - axis;
    double fi = floor(f), w = f - fi;
    long i0 = (long)fi;
    double acc = 0.0;
    for (int d = 0; d < 2; ++d) {
        long i = i0 + d;
        if (i < 0 || i >= n[o]) continue;
        long idx[2];
        idx[axis] = plane; idx[o] = i;
        acc += (d ? w : 1.0 - w) * img[(size_t)idx[1] * n[0] + idx[0]];
    }
    return acc;
}

/* Line integral of the image along {o + t d}; rays from a source only count t > 0. */
static double or_joseph_ray(const or_image *im, const double *vol, const double o[3], const double d[3], int from_source)
{
    int dim = (int)im->dim;
    double first[3] = {
        (double)im->mx - ((double)im->nx - 1.0) * (double)im->dx / 2.0,
        (double)im->my - ((double)im->ny - 1.0) * (double)im->dy / 2.0,
        (dim == 3) ? (double)im->mz - ((double)im->nz - 1.0) * (double)im->dz / 2.0 : 0.0 };
    double sp[3] = { im->dx, im->dy, im->dz };
    long n[3] = { (long)im->nx, (long)im->ny, (long)((dim == 3) ? im->nz : 1u) };
    int axis = 0;
    double best = -1.0;
    for (int a = 0; a < dim; ++a) {
        double q = fabs(d[a]) / sp[a];
        if (q > best) { best = q; axis = a; }
    }
    if (!(fabs(d[axis]) > 0.0)) return 0.0;
    double len = sqrt(d[0] * d[0] + d[1] * d[1] + (dim == 3 ? d[2] * d[2] : 0.0));
    double step = sp[axis] * len / fabs(d[axis]);
    double acc = 0.0;
    for (long i = 0; i < n[axis]; ++i) {
        double pa = first[axis] + (double)i * sp[axis];
        double t = (pa - o[axis]) / d[axis];
        if (from_source && t <= 0.0) continue;
        double q[3] = { o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2] };
        if (dim == 2) {
            int ot = 1 - axis;
            acc += or_bilin2(vol, im, axis, i, (q[ot] - first[ot]) / sp[ot]);
        } else {
            int oa = (axis == 0) ? 1 : 0, ob = (axis == 2) ? 1 : 2;
            acc += or_bilin(vol, im, axis, i, (q[oa] - first[oa]) / sp[oa], (q[ob] - first[ob]) / sp[ob]);
        }
    }
    return acc * step;
}

typedef struct { const or_scan *s; const or_image *im; const double *vol; double *proj; } or_jctx;

static void or_joseph_task(void *ctx, size_t idx)
{
    or_jctx *c = (or_jctx *)ctx;
    const or_scan *s = c->s;
    unsigned nv = (s->kind <= OR_LINE1D) ? 1u : s->nv;
    unsigned v = (unsigned)(idx / nv), r = (unsigned)(idx % nv);
    const or_view *g = &s->views[v];
    double cu = g->det_row_center, cv = g->det_col_center, D = g->src_det_dist;
    double s3[3] = { g->src_pos[0], g->src_pos[1], g->src_pos[2] };
    double eu[3] = { g->det_row_dir[0], g->det_row_dir[1], g->det_row_dir[2] };
    double ev[3] = { g->det_col_dir[0], g->det_col_dir[1], g->det_col_dir[2] };
    double ew[3] = { g->iso_ray_dir[0], g->iso_ray_dir[1], g->iso_ray_dir[2] };
    double vr = ((double)r - cv) * (double)s->dv;
    for (unsigned col = 0; col < s->nu; ++col) {
        double uc = ((double)col - cu) * (double)s->du;
        double o[3] = { 0, 0, 0 }, d[3] = { 0, 0, 0 };
        int from_src = 1;
        switch (s->kind) {
        case OR_PAR1D: case OR_PAR2D: case OR_CONEPAR: {
            double th = (double)g->rot_ang;
            o[0] = -uc * sin(th); o[1] = uc * cos(th);
            d[0] = cos(th); d[1] = sin(th);
            if (s->kind == OR_PAR2D) o[2] = -vr;
            if (s->kind == OR_CONEPAR) o[2] = -vr * (double)g->src_iso_dist / D;
            from_src = 0;
            break;
        }
        case OR_LINE1D: {
            double w0 = -eu[1], w1 = eu[0];           /* rot90(eu) */
            o[0] = s3[0]; o[1] = s3[1];
            d[0] = D * w0 + uc * eu[0]; d[1] = D * w1 + uc * eu[1];
            break;
        }
        case OR_ARC1D:
            o[0] = s3[0]; o[1] = s3[1];
            d[0] = cos(uc) * ew[0] + sin(uc) * eu[0];
            d[1] = cos(uc) * ew[1] + sin(uc) * eu[1];
            break;
        case OR_CYL:
            for (int a = 0; a < 3; ++a) {
                o[a] = s3[a];
                d[a] = D * (cos(uc) * ew[a] + sin(uc) * eu[a]) + vr * ev[a];
            }
            break;
        case OR_FLAT: {
            double w3[3] = { eu[1] * ev[2] - eu[2] * ev[1], eu[2] * ev[0] - eu[0] * ev[2], eu[0] * ev[1] - eu[1] * ev[0] };
            for (int a = 0; a < 3; ++a) { o[a] = s3[a]; d[a] = D * w3[a] + uc * eu[a] + vr * ev[a]; }
            break;
        }
        }
        c->proj[((size_t)v * nv + r) * s->nu + col] = or_joseph_ray(c->im, c->vol, o, d, from_src);
    }
}

int or_joseph_f64(const or_scan *s, const or_image *im, const double *vol, double *proj, int nthreads)
{
    unsigned nv = (s->kind <= OR_LINE1D) ? 1u : s->nv;
    if ((s->kind <= OR_LINE1D) != (im->dim == 2)) return -1;
    or_jctx c = { s, im, vol, proj };
    or_parallel_for((size_t)s->nviews * nv, nthreads, or_joseph_task, &c);
    return 0;
}
