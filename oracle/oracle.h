/*
 * CPU oracle for the FDK reconstruction path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke()
 * function of __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it.  The product (libgct_amd.so) never links, loads or calls it.
 *
 * It restates the reference's FDK path (satidev/gct, src/cuda/recon + src/cpp)
 * on the CPU in two forms:
 *   *_f32  -- FP32 arithmetic in the reference's operation order (parity form);
 *   *_f64  -- the same discrete algorithm in FP64 (accuracy oracle).
 * Every function cites the reference file:line it follows.  The whole file is
 * compiled with -ffp-contract=off so the FP32 form rounds where the reference's
 * source says it rounds.
 */
#ifndef GCT_ORACLE_H
#define GCT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* One view; identical layout to gct_proj_geom of the product ABI (union of the
 * seven ProjGeom<> specialisations, src/cpp/geom/proj_geom.h:21-103). */
typedef struct or_view {
    float rot_ang;
    float src_det_dist;
    float src_iso_dist;
    float det_row_center;
    float det_col_center;
    float src_pos[3];
    float det_row_dir[3];
    float det_col_dir[3];
    float iso_ray_dir[3];
} or_view;

enum { OR_PAR1D = 0, OR_ARC1D = 1, OR_LINE1D = 2, OR_PAR2D = 3, OR_CONEPAR = 4, OR_CYL = 5, OR_FLAT = 6 };
enum { OR_WIN_NONE = 0, OR_WIN_HANN = 1, OR_WIN_HAMMING = 2, OR_WIN_SHEPP_LOGAN = 3 };

typedef struct or_scan {
    int kind;
    unsigned nu;       /* DetRow sampler: columns of a projection */
    float du;
    unsigned nv;       /* DetCol sampler: rows of a projection (1 for 1-D kinds) */
    float dv;
    unsigned nviews;
    const or_view *views;
} or_scan;

typedef struct or_image {
    unsigned dim;      /* 2 or 3 */
    unsigned nx, ny, nz;
    float dx, dy, dz;
    float mx, my, mz;
} or_image;

typedef struct or_rot {
    int dir;           /* +1 CW, -1 CCW (rotation_geom.h:10-14) */
    float incr;
    float range;
    unsigned nppr;
} or_rot;

/* Rotation analysis (rotation_geom.cpp:7-253).  Returns 0, or -1 for an
 * invalid trajectory (the reference returns std::nullopt and CTGeom throws). */
int or_rotation_geom(const or_scan *scan, or_rot *out, float *rot_ang, float *rel_ang);
int or_rotation_geom_f64(const or_scan *scan, int *dir, double *incr, double *range, double *rel_ang);

/* Wrap / quadrant helpers (utils.h:322-389), exported for the known-answer tests. */
float or_wrap_0_2pi(float a);
unsigned or_quadrant(float a);

/* Ramp-filter taps (ramp_filter.h:29-56, proj_filt_kernel.cpp:5-45).
 * arc != 0 applies the (n dgamma / sin n dgamma)^2 weighting. */
void or_ramp_taps(int arc, unsigned n, float width, float cutoff, int window,
                  float *taps32, double *taps64);

/* Pre-weights: cosine [V][Nv][Nu] (cos_weighting.cuh:16-153) and redundancy
 * [V][Nu] (parker_weighting.cuh:19-200). */
void or_weights_f32(const or_scan *scan, float *cosw, float *red);
void or_weights_f64(const or_scan *scan, double *cosw, double *red);

/* Row-wise truncated linear convolution q[c] = sum_j p[j] h(c - j), the
 * operation ramp_filtering_fft.cuh:208-219 performs with cuFFT. */
void or_filter_rows_f32(unsigned nrows, unsigned n, const float *taps, const float *in, float *out, int nthreads);
void or_filter_rows_f64(unsigned nrows, unsigned n, const double *taps, const double *in, double *out, int nthreads);
void or_filter_rows_direct_f64(unsigned nrows, unsigned n, const double *taps, const double *in, double *out);

/* Voxel-driven backprojection of filtered projections [V][Nv][Nu] into the
 * slices [z0, z1) of vol (back_proj.cuh:16-104, coord_trans_proj.cuh).
 * vol holds nz (3-D) or 1 (2-D) slices laid out x fastest.  scale != 0
 * multiplies by the angular increment afterwards (fdk_axial.cuh:196-199). */
int or_bp_f32(const or_scan *scan, const or_image *img, const float *filt, float *vol,
              unsigned z0, unsigned z1, int scale, int nthreads);
int or_bp_f64(const or_scan *scan, const or_image *img, const double *filt, double *vol,
              unsigned z0, unsigned z1, int scale, int nthreads);

/* Whole FDK (fdk_axial.cuh:152-205): cosine -> redundancy -> ramp -> BP -> x dtheta. */
int or_fdk_f32(const or_scan *scan, const or_image *img, float cutoff, int window,
               const float *proj, float *vol, unsigned z0, unsigned z1, int nthreads);
int or_fdk_f64(const or_scan *scan, const or_image *img, float cutoff, int window,
               const float *proj, double *vol, unsigned z0, unsigned z1, int nthreads);

/* Spiral FDK (fdk_spiral.cuh:28-94): per-slice view windows [win[2z], win[2z+1])
 * (empty when the reference skips the slice); -1 for a geometry SpiralCTGeom
 * rejects (spiral_ct_geom.h:30-53). */
int or_spiral_windows(const or_scan *scan, const or_image *img, unsigned *win);
/* Slices [z0, z1) of the spiral reconstruction: cosine weight, Ram-Lak ramp,
 * backprojection of each slice's window, x dtheta. */
int or_fdk_spiral_f32(const or_scan *scan, const or_image *img, const float *proj, float *vol,
                      unsigned z0, unsigned z1, int nthreads);
int or_fdk_spiral_f64(const or_scan *scan, const or_image *img, const float *proj, double *vol,
                      unsigned z0, unsigned z1, int nthreads);

/* Joseph forward projection (no reference counterpart: the reference's
 * forward_proj.cuh is dead code).  Rays follow the backprojector's conventions. */
int or_joseph_f64(const or_scan *scan, const or_image *img, const double *vol, double *proj, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
