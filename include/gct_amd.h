/*
 * gct_amd.h -- C ABI of the MI355X (gfx950) FDK reconstruction library.
 *
 * Plain C types only: pointers, sizes, PODs.  No exceptions cross this
 * boundary: every entry point returns a gct_status, and gct_last_error()
 * returns the calling thread's last message.  The C++ drop-in API in
 * include/gct/ (namespace GCT) is a header-only layer over these calls that
 * restores the reference's exception behaviour.
 *
 * Data layouts are those of the reference (bit-compatible host buffers):
 *   projections  [num_proj][det_col.num_samples][det_row.num_samples], float,
 *                detector column (u) fastest (src/cpp/data/projection_stack.h:14,
 *                src/cpp/data/proj_meta_info.cpp:25-30);
 *   image/volume [z][y][x], float, x fastest (src/cuda/recon/back_proj.cuh:76).
 */
#ifndef GCT_AMD_H
#define GCT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCT_AMD_ABI_VERSION 2

/* Status codes.  The reference throws std::invalid_argument / std::runtime_error
 * (src/cuda/utils/check_error.cu:6-36, src/cpp/geom/ct_geom.h:26-32); the C++
 * wrapper maps GCT_ERR_INVALID_ARGUMENT to the former and the rest to the latter. */
typedef enum gct_status {
    GCT_OK = 0,
    GCT_ERR_INVALID_ARGUMENT = 1,
    GCT_ERR_RUNTIME = 2,          /* HIP failure, invalid source trajectory */
    GCT_ERR_OUT_OF_MEMORY = 3,
    GCT_ERR_NO_DEVICE = 4,
    GCT_ERR_UNSUPPORTED = 5,
    GCT_ERR_REJECTED = 6,         /* serving: request larger than the HBM budget */
    GCT_ERR_NOT_READY = 7         /* serving: ticket still in flight */
} gct_status;

/* The seven axial geometries of recon_wrappers.cuh:16-43. */
typedef enum gct_geom_kind {
    GCT_PARALLEL_1D = 0,   /* AxialCTGeom<DetParallel> */
    GCT_FAN_ARC_1D = 1,    /* AxialCTGeom<DetFanArc> */
    GCT_FAN_LINE_1D = 2,   /* AxialCTGeom<DetFanLine> */
    GCT_PARALLEL_2D = 3,   /* AxialCTGeom<DetParallel, DetParallel> */
    GCT_CONE_PARALLEL = 4, /* AxialCTGeom<DetParallel, DetFanLine> (rebinned) */
    GCT_CYLINDRICAL = 5,   /* AxialCTGeom<DetFanArc, DetFanLine> */
    GCT_FLAT = 6           /* AxialCTGeom<DetFanLine, DetFanLine> */
} gct_geom_kind;

/* Sampler1D (src/cpp/geom/sampling_info.h:8-17): mm, or radians for DetFanArc. */
typedef struct gct_sampler {
    uint32_t num_samples;
    float sampling_width;
} gct_sampler;

/* One view: the union of the seven ProjGeom<> specialisations
 * (src/cpp/geom/proj_geom.h:21-103).  Fields a geometry does not have are
 * ignored.  Vectors are (x, y, z); 1-D fan geometries ignore z. */
typedef struct gct_proj_geom {
    float rot_ang;          /* rad: parallel, cone-parallel */
    float src_det_dist;     /* mm: fan, cone, cone-parallel */
    float src_iso_dist;     /* mm: cone-parallel only (others use |src_pos|, ct_geom_dev.cu:84,145) */
    float det_row_center;   /* pixels (u) */
    float det_col_center;   /* pixels (v); 2-D detectors */
    float src_pos[3];
    float det_row_dir[3];
    float det_col_dir[3];
    float iso_ray_dir[3];   /* arc and cylindrical detectors */
} gct_proj_geom;

/* AxialCTGeom<DetRow, DetCol...> (src/cpp/geom/ct_geom.h:16-53). */
typedef struct gct_ct_geom {
    int32_t kind;                    /* gct_geom_kind */
    gct_sampler det_row;             /* along a detector row: projection columns (u) */
    gct_sampler det_col;             /* along a detector column: projection rows (v); ignored for 1-D */
    uint32_t num_proj;
    const gct_proj_geom *proj_geom;  /* num_proj entries, caller-owned */
} gct_ct_geom;

/* ReconImgGeom<2|3> (src/cpp/geom/recon_img_geom.h:15-30). */
typedef struct gct_image_geom {
    uint32_t dim;                    /* 2 for the 1-D detector kinds, 3 otherwise */
    gct_sampler x, y, z;             /* z ignored when dim == 2 */
    float mid_point[3];
} gct_image_geom;

/* RotationGeom (src/cpp/geom/rotation_geom.h:16-24). */
typedef struct gct_rotation_geom {
    int32_t dir;                     /* +1 CW, -1 CCW */
    float rot_ang_incr;
    float rot_ang_range;
    uint32_t num_proj_per_rot;
} gct_rotation_geom;

/* ApodFunc (src/cpp/io/parameters/filt_parameters.h:7-18) plus Shepp-Logan. */
typedef enum gct_window {
    GCT_WINDOW_NONE = 0,
    GCT_WINDOW_HANN = 1,
    GCT_WINDOW_HAMMING = 2,
    GCT_WINDOW_SHEPP_LOGAN = 3       /* extension: not in the reference */
} gct_window;

/* Storage of the filtered projections read by backprojection (accumulation is FP32). */
typedef enum gct_storage {
    GCT_STORE_F32 = 0,
    GCT_STORE_F16 = 1,
    GCT_STORE_BF16 = 2
} gct_storage;

/* Precision of the per-thread voxel-to-detector setup in backprojection. */
typedef enum gct_setup_precision {
    GCT_SETUP_AUTO = 0,              /* FP64 for volumes >= 512^3 voxels, else FP32 */
    GCT_SETUP_F32 = 1,
    GCT_SETUP_F64 = 2
} gct_setup_precision;

typedef enum gct_mem_mode {
    GCT_MEM_AUTO = 0,                /* reconFDK's 75 % of HBM rule (fdk_axial.cuh:85-88) */
    GCT_MEM_HIGH = 1,                /* all views resident */
    GCT_MEM_LOW = 2,                 /* views streamed in chunks */
    GCT_MEM_SLAB = 3                 /* volume built in z slabs of z_slab slices (volumes beyond device
                                      * memory; AUTO picks it for a 3-D image when the volume and a
                                      * chunk's buffers exceed 90 % of free memory).  The filtered
                                      * stack is built once and kept when it takes at most half of free
                                      * memory and view_chunk is 0; otherwise the views stream through
                                      * again for every slab.  Same volume, bit for bit, as HIGH/LOW */
} gct_mem_mode;

/* Backprojection kernel family. */
typedef enum gct_bp_kernel {
    GCT_BP_AUTO = 0,                 /* tuned choice, else LDS-staged (3-D) / direct (2-D images) */
    GCT_BP_DIRECT = 1,               /* taps read from HBM/L2 through the vector cache */
    GCT_BP_LDS = 2                   /* per-tile detector footprint staged in LDS */
} gct_bp_kernel;

typedef struct gct_recon_opts {
    float cut_off_freq;              /* FiltParams::cut_off_freq, default 1 */
    int32_t window;                  /* gct_window, default NONE (the reference's FiltParams{}) */
    int32_t storage;                 /* gct_storage, default F32 */
    int32_t setup_precision;         /* gct_setup_precision, default AUTO */
    int32_t mem_mode;                /* gct_mem_mode, default AUTO */
    int32_t device;                  /* HIP device ordinal, default 0 (fdk_axial.cuh:73) */
    uint32_t z_unroll;               /* BP voxels per thread along z; 0 = tuned */
    uint32_t view_chunk;             /* LOW and SLAB modes: views per chunk; HIGH: views per upload chunk of
                                        the pipelined copy (gct_plan_execute_host); 0 = derived */
    int32_t bp_kernel;               /* gct_bp_kernel, default AUTO */
    /* LDS backprojection shape (z_unroll, view_block, lds_rows, lds_pairs): one of
     * (16, 2, 40, 40), (16, 2, 32, 40), (16, 2, 40, 48), (16, 3, 28, 34), (16, 4, 24, 28), (32, 3, 48, 34),
     * (32, 4, 36, 28);
     * a zero field matches any value (the first matching shape is taken); all four zero = the
     * shape the autotuner found for this scan shape, else the first of (32, 4, 36, 28),
     * (32, 3, 48, 34), (16, 4, 24, 28), (16, 3, 28, 34), (16, 2, 32, 40) whose staging area holds
     * 99 % of a sample of the scan's (tile, view) boxes, else (16, 2, 32, 40) if it holds
     * within 0.5 % as many of them as (16, 2, 40, 40), else (16, 2, 40, 40) */
    uint32_t view_block;             /* views staged per block */
    uint32_t lds_rows;               /* detector rows staged per view */
    uint32_t lds_pairs;              /* (value, difference) pairs staged per row */
    uint32_t z_slab;                 /* SLAB mode slices per slab, a multiple of 32; 0 = derived */
} gct_recon_opts;

/* ---- library ----------------------------------------------------------- */

uint32_t gct_abi_version(void);
const char *gct_last_error(void);
const char *gct_status_string(gct_status s);
void gct_recon_opts_default(gct_recon_opts *opts);
/* Number of visible HIP devices (0 on a host without a GPU). */
int gct_device_count(void);

/* ---- geometry (host, no GPU needed) ------------------------------------ */

/* Circular-trajectory parameters: the union of Parallel1D/FanBeam/Parallel2D/
 * ConeBeamProjGeomParams and RotationParams (src/cpp/io/parameters/proj_geom_params.h:16-59). */
typedef struct gct_circular_params {
    int32_t kind;                    /* gct_geom_kind */
    int32_t dir;                     /* +1 CW (default), -1 CCW */
    float init_ang;                  /* rad */
    float ang_incr;                  /* rad */
    uint32_t num_proj;
    float hel_pitch;                 /* must be 0 for axial scans */
    float det_row_center;
    float det_col_center;
    float src_det_dist;
    float src_iso_dist;
} gct_circular_params;

/* ProjGeomFactory<...>::create for every view (proj_geom_factory_params.cpp:8-108,
 * proj_geom_cb_factory.cpp:7-48, proj_geom_fb_factory.cpp:6-45).
 * out holds params->num_proj entries. */
gct_status gct_geom_create_circular(const gct_circular_params *params, gct_proj_geom *out);

/* getRotationGeom (rotation_geom.h:90-115).  rot_ang and rel_rot_ang are
 * optional outputs of num_proj floats.  GCT_ERR_RUNTIME for a stationary or
 * ambiguous trajectory (the case in which CTGeom's constructor throws). */
gct_status gct_rotation_geom_compute(const gct_ct_geom *geom, gct_rotation_geom *out,
                             float *rot_ang, float *rel_rot_ang);

/* Recon::isValidAxialRecon (src/cpp/recon/check_validity.h:66-75): projection
 * sizes against the detector and view count.  Returns GCT_OK or
 * GCT_ERR_INVALID_ARGUMENT with the reason in gct_last_error(). */
gct_status gct_check_axial_recon(const gct_ct_geom *geom, uint32_t proj_cols,
                                 uint32_t proj_rows, uint32_t num_proj);

/* RampFilter::kernel / projFiltKernel (ramp_filter.h:29-56, proj_filt_kernel.cpp:5-45):
 * det_row.num_samples taps of the spatial ramp kernel for geom's row detector. */
gct_status gct_ramp_kernel(const gct_ct_geom *geom, float cut_off_freq, int32_t window, float *taps);

/* Flat-panel projection matrix P = K[R | -R s] / P[2][3]
 * (ProjectionMatrix(pg, du, dv), projection_matrix.cpp:66-96), row-major 3x4, FP64. */
gct_status gct_projection_matrix(const gct_proj_geom *pg, float pix_width, float pix_height, double *p3x4);

/* Decomposition of a 3x4 projection matrix into principal point, focal
 * length, rotation and source (ProjectionMatrix(mat), projection_matrix.cpp:4-65),
 * returned as a flat-panel view (det_row/col_center, src_det_dist = focal_u *
 * pix_width, src_pos, det_row_dir, det_col_dir) -- proj_geom_factory_proj_mat.cpp:7-26. */
gct_status gct_projection_matrix_decompose(const double *p3x4, float pix_width,
                                           gct_proj_geom *out, double *focal_uv);

/* ---- one-shot reconstruction ------------------------------------------- */

/* Recon::CUDA::reconstruct x7 / reconFDK (recon_wrappers.cuh:16-43,
 * fdk_axial.cuh:69-89): host projections in, host image out; synchronous.
 * opts may be NULL (defaults). */
gct_status gct_reconstruct(const gct_ct_geom *geom, const float *proj,
                           const gct_image_geom *img, float *vol, const gct_recon_opts *opts);

/* Memory the one-shot path needs with every view resident
 * (Detail::reqMemoryForRecon, fdk_axial.cuh:53-66, for this library's buffers). */
gct_status gct_required_memory(const gct_ct_geom *geom, const gct_image_geom *img,
                               const gct_recon_opts *opts, uint64_t *bytes);

/* ---- plans: device-resident work on caller streams --------------------- */

typedef struct gct_plan gct_plan;

typedef struct gct_plan_info {
    uint32_t fft_len;                /* nextPow2(N + floor(N/2)), at least 8: the shortest FFT whose circular
                                        convolution leaves the N kept outputs unwrapped (the reference's
                                        nextPow2(2 N), ramp_filtering_fft.cuh:36-48, gives the same values) */
    uint32_t filt_pitch;             /* elements per padded filtered row */
    uint32_t filt_rows_per_view;     /* Nv + 3 (one zero row above, two below) */
    uint32_t z_unroll;
    uint64_t filt_bytes;             /* padded filtered-projection buffer */
    uint64_t device_bytes;           /* everything the plan holds on the device */
    uint64_t updates;                /* voxels x views */
    float rot_ang_incr;
    float rot_ang_range;
    int32_t rot_dir;
    int32_t setup_f64;
    int32_t circular;                /* 1 when every view keeps the row axis along z.  0 (e.g. views from
                                        projection matrices with a tilted detector): flat panels backproject
                                        with the LDS kernel of that case, shape (16, 2, 40, 40), or the
                                        direct kernel; other kinds with the direct kernel */
    int32_t storage;
    int32_t bp_kernel;               /* gct_bp_kernel in use */
    uint32_t view_block;             /* LDS BP shape in use (0 for the direct kernel): views per block, */
    uint32_t lds_rows;               /*   staged rows per view, */
    uint32_t lds_pairs;              /*   pairs per staged row */
    int32_t rows_inside;             /* 1 when every voxel projects inside the detector rows in every
                                        view: the LDS backprojection then runs without its row clamp */
    int32_t round_robin;             /* 1 when the LDS backprojection deals its tiles to the XCDs round
                                        robin instead of in contiguous runs (the autotuner's choice, else
                                        round robin for volumes more than 4 tiles deep) */
    int32_t narrow_fan;              /* arc and cylindrical detectors, FP32 set-up: 1 when every voxel lies
                                        in front of the source in every view and the detector spans at most
                                        +-0.6 rad: the kernels take the fan angle without octant reduction */
    int32_t box_inner;               /* 1 when the LDS backprojection may leave the row clamp out for a view
                                        whose staged box the detector's edges did not cut (every voxel of a
                                        tile then projects inside the box's rows) */
} gct_plan_info;

/* Builds per-view tables (FP64 on the host, stored FP32), the redundancy
 * table and the filter spectrum on the device, and the padded filtered-
 * projection buffer.  Synchronous. */
gct_status gct_plan_create(const gct_ct_geom *geom, const gct_image_geom *img,
                           const gct_recon_opts *opts, gct_plan **out);
void gct_plan_destroy(gct_plan *plan);
gct_status gct_plan_get_info(const gct_plan *plan, gct_plan_info *info);

/* Stage 1+2 (fused): cosine and redundancy weighting and ramp filtering of the
 * device projections into the plan's padded buffer (fdk_axial.cuh:170-178).
 * first_view/num_views select a contiguous range of views (all: 0, num_proj).
 * Rows are transformed in pairs counted from first_view; a range filters to
 * the same bits as the whole stack when first_view x Nv is even. */
gct_status gct_plan_filter(gct_plan *plan, const float *proj_dev, uint32_t first_view,
                           uint32_t num_views, void *stream);

/* Stage 3: backprojection of views [first_view, first_view + num_views) from the
 * plan's filtered buffer.  accumulate != 0 continues sums already in vol_dev;
 * finish != 0 applies the x dtheta scale (fdk_axial.cuh:189-199). */
gct_status gct_plan_backproject(gct_plan *plan, float *vol_dev, uint32_t first_view,
                                uint32_t num_views, int accumulate, int finish, void *stream);

/* All stages, all views: proj_dev [V][Nv][Nu] -> vol_dev. */
gct_status gct_plan_execute(gct_plan *plan, const float *proj_dev, float *vol_dev, void *stream);

/* All stages from host projections to a host volume, as reconFDKHighMem
 * (fdk_axial.cuh:152-205) with its copies: the views are uploaded in chunks of
 * chunk_views (0: about an eighth of the scan) on a stream of the library's,
 * into proj_dev, and each chunk is filtered and backprojected into vol_dev on
 * `stream` as soon as it is there, so the upload of one chunk overlaps the
 * work on the previous one; vol_dev is then copied to vol_host.  Same volume
 * as gct_plan_execute, bit for bit.  Asynchronous on `stream` for pinned host
 * memory (keep both host buffers alive until it completes); pageable memory
 * works too, with the copies holding the calling thread. */
gct_status gct_plan_execute_host(gct_plan *plan, const float *proj_host, float *proj_dev, float *vol_dev,
                                 float *vol_host, uint32_t chunk_views, void *stream);

/* Copies filtered projections [V][Nv][Nu] (FP32) into / out of the plan's
 * padded buffer, to test the stages separately. */
gct_status gct_plan_load_filtered(gct_plan *plan, const float *filt_dev, void *stream);
gct_status gct_plan_store_filtered(gct_plan *plan, float *filt_dev, void *stream);

/* Joseph ray-driven forward projection of vol_dev into proj_dev [V][Nv][Nu]
 * along the rays the backprojector's mapping implies.  No reference
 * counterpart (the reference's forward_proj.cuh is never compiled).  3-D
 * kinds trace blocks of rays through volume bricks staged in LDS; the
 * environment variable GCT_JOSEPH=per_ray (read once per process) selects
 * the one-ray-per-thread kernel instead, which gives the same bits. */
gct_status gct_plan_forward_project(gct_plan *plan, const float *vol_dev, float *proj_dev, void *stream);

/* ---- batches: many uneven scans in one graph launch --------------------- */

typedef struct gct_batch gct_batch;

/* plans must stay alive while the batch exists; they are not copied.  Plans of
 * any kind, storage and set-up precision mix: scans that share a specialised
 * kernel (kind, storage, FP64 set-up, tile type) form a group. */
gct_status gct_batch_create(gct_plan *const *plans, uint32_t num_plans, gct_batch **out);
void gct_batch_destroy(gct_batch *batch);
/* Filters every scan, then backprojects each group in one persistent launch
 * (groups run side by side).  use_graph != 0 captures the sequence once into
 * a hipGraph and replays it while the stream and buffers stay the same; the
 * graph and the plain path keep separate device work tables.  Volumes equal
 * each plan's own gct_plan_execute bit for bit.  A batch, like its plans,
 * must not execute on two streams at once. */
gct_status gct_batch_execute(gct_batch *batch, const float *const *proj_dev, float *const *vol_dev,
                             int use_graph, void *stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc: gct_batch_create + gct_batch_execute(use_graph = 0) +
 * gct_batch_destroy in one call that allocates nothing (no hipMalloc,
 * hipFree, host staging or stream/event creation).  The work table and work
 * counters live in a fixed array of the library on each device (about 160 KiB:
 * 256 B of counters and GCT_NOALLOC_MAX_SCANS table entries), held until the
 * process ends; the table reaches it through kernel arguments.  plans[i]
 * reconstructs proj_dev[i] into vol_dev[i], as the plan's own execute would,
 * bit for bit.  At most GCT_NOALLOC_MAX_SCANS scans and 64 kernel groups
 * (GCT_ERR_INVALID_ARGUMENT / GCT_ERR_UNSUPPORTED beyond).  Every call on one
 * device shares that device's array: calls that overlap on one device
 * (several host threads, or streams that run concurrently) race and must be
 * serialised by the caller.  Groups run one after another on `stream`. */
#define GCT_NOALLOC_MAX_SCANS 1024
gct_status gct_batch_execute_noalloc(gct_plan *const *plans, uint32_t num_plans, const float *const *proj_dev,
                                     float *const *vol_dev, void *stream);

/* ---- autotuner ----------------------------------------------------------- */

/* Times backprojection for the plan's shape over the candidate kernels --
 * LDS-staged with z-unroll 16 (view blocks 2, 3, 4) or 32 (view blocks 1, 2),
 * direct with z-unroll 8 or 16 -- and stores the fastest in the plan (and a
 * process-wide cache keyed by shape; gct_plan_get_info reports the choice).
 * Uses vol_dev as scratch: its contents are overwritten.  On an error the plan
 * keeps the kernel shape it had. */
gct_status gct_plan_autotune(gct_plan *plan, float *vol_dev, uint32_t *best_z_unroll, float *best_ms);

/* ---- serving: continuous batching of reconstruction requests ------------ */

typedef struct gct_server gct_server;

typedef struct gct_server_opts {
    int32_t device;
    uint64_t hbm_budget_bytes;       /* 0: 80 % of free device memory */
    uint32_t max_inflight;           /* requests on the device at once; 0: 8 */
    uint32_t num_streams;            /* HIP streams work is spread over; 0: 4 */
    /* Plans kept for reuse after their request completes, keyed by the whole
     * geometry, image grid and options (repeated protocols skip the set-up:
     * per-view tables, redundancy table, filter spectrum, kernel shape).  Their
     * device memory counts against hbm_budget_bytes and the least recently
     * used idle plan is released when a request needs the room.  0: 16;
     * UINT32_MAX: none. */
    uint32_t plan_cache_entries;
    /* Admission is first fit in arrival order; a request passed over this
     * many times (later requests admitted while it did not fit) blocks every
     * later request until it is admitted.  0: 4. */
    uint32_t max_bypass;
} gct_server_opts;

typedef struct gct_server_stats {
    uint64_t submitted, completed, failed, rejected;
    uint64_t updates;                /* voxel x view updates completed */
    uint64_t peak_device_bytes;
    uint32_t peak_inflight;
    double busy_seconds;             /* first submit -> last completion */
    uint64_t plan_cache_hits, plan_cache_misses;
} gct_server_stats;

gct_status gct_server_create(const gct_server_opts *opts, gct_server **out);
/* Destroys after draining every admitted request. */
void gct_server_destroy(gct_server *server);

/* Queues one request.  geom/img/opts and their arrays are copied; proj_host
 * must stay valid and vol_host writable until the ticket completes.
 * reproj_err (optional) receives the RMS of (Joseph reprojection of the
 * result - projections) / RMS(projections) when reproject != 0. */
gct_status gct_server_submit(gct_server *server, const gct_ct_geom *geom, const gct_image_geom *img,
                             const gct_recon_opts *opts, const float *proj_host, float *vol_host,
                             int reproject, uint64_t *ticket);
/* Blocks until the ticket completes; returns the request's own status. */
gct_status gct_server_wait(gct_server *server, uint64_t ticket, double *latency_ms, double *reproj_err);
/* GCT_ERR_NOT_READY while in flight. */
gct_status gct_server_poll(gct_server *server, uint64_t ticket, double *latency_ms, double *reproj_err);
gct_status gct_server_get_stats(gct_server *server, gct_server_stats *stats);

/* ---- synthetic phantoms (benchmark inputs) ------------------------------- */

/* Random ellipses (2-D) or ellipsoids (3-D) with seeded parameters, as
 * [n][10] doubles: centre(3), semi-axes(3), rotation angles (3: Euler z-y-z),
 * density.  radius is the field-of-view radius in mm. */
gct_status gct_phantom_random(uint64_t seed, uint32_t count, int dim, double radius,
                              double axis_lo, double axis_hi, double dens_lo, double dens_hi,
                              double *params);
/* Exact line integrals (FP64, cast to FP32) of a phantom along every detector
 * ray of geom; i0 > 0 adds Poisson noise: p = -ln(max(Poisson(i0 e^-p), 1) / i0). */
gct_status gct_phantom_project(const gct_ct_geom *geom, int dim, const double *params, uint32_t count,
                               double i0, uint64_t noise_seed, float *proj, int nthreads);
/* Phantom rasterised at voxel centres into img (used by reprojection tests). */
gct_status gct_phantom_rasterize(const gct_image_geom *img, const double *params, uint32_t count, float *vol);
/* Device twins of the two calls above: the same FP64 arithmetic (and the same
 * per-row noise streams) on the current device, written to device memory and
 * enqueued on `stream` (hipStream_t, NULL = null stream); they return when the
 * work is done.  At most 64 shapes. */
gct_status gct_phantom_project_device(const gct_ct_geom *geom, int dim, const double *params, uint32_t count,
                                      double i0, uint64_t noise_seed, float *proj_dev, void *stream);
gct_status gct_phantom_rasterize_device(const gct_image_geom *img, const double *params, uint32_t count,
                                        float *vol_dev, void *stream);

/* ---- spiral FDK --------------------------------------------------------- */

/* SpiralCTGeom::firstAndLastProjIndices for every slice of img
 * (spiral_ct_geom.h:30-95, fdk_spiral.cuh:53-70): win[2z], win[2z+1] =
 * [first, last) views of slice z, or 0, 0 when the reference skips the slice.
 * GCT_ERR_RUNTIME for a geometry SpiralCTGeom rejects (zero projections per
 * rotation, unsorted source z, non-positive pitch). */
gct_status gct_spiral_windows(const gct_ct_geom *geom, const gct_image_geom *img, uint32_t *win);
/* Recon::CUDA::reconstruct(SpiralCTGeom<DetFanArc, DetFanLine>, ...)
 * (recon_wrappers.cu:58-63 -> fdk_spiral.cuh:28-94): cosine weighting and the
 * Ram-Lak ramp (no redundancy weighting), each slice backprojected from its
 * own window, x rot_ang_incr.  Cylindrical (the reference's case) and flat
 * detectors; host buffers in and out, synchronous.  opts: storage, device. */
gct_status gct_reconstruct_spiral(const gct_ct_geom *geom, const float *proj, const gct_image_geom *img,
                                  float *vol, const gct_recon_opts *opts);

/* ---- array files: MetaImage, Dennerlein, raw (the reference's src/cpp/io) -- */

/* ArrayElemType (metadata.h:9-21). */
typedef enum gct_elem_type {
    GCT_T_UINT8 = 0, GCT_T_INT8, GCT_T_UINT16, GCT_T_INT16, GCT_T_UINT32, GCT_T_INT32,
    GCT_T_UINT64, GCT_T_INT64, GCT_T_FLOAT, GCT_T_DOUBLE
} gct_elem_type;

typedef enum gct_file_format {
    GCT_FILE_AUTO = 0,   /* by extension: .mha/.mhd MetaImage, .den Dennerlein, else raw */
    GCT_FILE_META = 1,   /* MetaImgFileReader (meta_img_reader.cpp:10-165) */
    GCT_FILE_DEN = 2,    /* DenFileReader (den_file_reader.cpp:7-61) */
    GCT_FILE_RAW = 3     /* RawDataFileReader (raw_data_file_reader.h:13-33): caller's metadata */
} gct_file_format;

/* ArrayMetadata (metadata.h:23-29): rows x cols per slice, slices per file. */
typedef struct gct_array_metadata {
    uint32_t num_rows;
    uint32_t num_cols;
    uint32_t num_slices;
    int32_t type;                    /* gct_elem_type */
} gct_array_metadata;

/* ArrayFileReader::readMetadata.  raw_md: the metadata of a raw file (else NULL). */
gct_status gct_file_read_metadata(const char *filename, int32_t format, const gct_array_metadata *raw_md,
                                  gct_array_metadata *md);
/* ArrayFileReader::readData<T> (array_file_reader.h:66-72): the elements of the
 * file -- or, with series != 0, of every file in its directory tree with the
 * same extension and size, in name order (getAllFiles, file_utils.cpp:7-28) --
 * converted to out_type by static_cast.  *count receives the element count;
 * out == NULL only queries it.  GCT_ERR_RUNTIME for the reference's
 * runtime_error cases (missing file or data file, size mismatch, bad
 * ElementType). */
gct_status gct_file_read(const char *filename, int32_t format, int series, const gct_array_metadata *raw_md,
                         int32_t out_type, void *out, uint64_t capacity, uint64_t *count);
/* "MET_FLOAT" etc. (MetaImage::typeStr, meta_img_file_writer.cpp:11-69); NULL if unknown. */
const char *gct_meta_type_name(int32_t type);
/* ArrayFileWriter::write for MetaImage (meta_img_file_writer.h:101-152): the
 * reference's header lines; .mha holds the data after the header, any other
 * extension writes the data to <stem>.raw beside it.  num_elems must equal
 * the product of dim_size. */
gct_status gct_meta_write(const char *filename, uint32_t ndims, const uint32_t *dim_size, const float *spacing,
                          const float *offset, int32_t type, const void *data, uint64_t num_elems);
/* ArrayFileWriter::write for Dennerlein files (den_file_writer.h:14-60). */
gct_status gct_den_write(const char *filename, uint32_t num_rows, uint32_t num_cols, uint32_t num_slices,
                         int32_t type, const void *data, uint64_t num_elems);
/* ArrayFileWriter::write for headerless files (raw_data_file_writer.h:13-36). */
gct_status gct_raw_write(const char *filename, int32_t type, const void *data, uint64_t num_elems);

#ifdef __cplusplus
}
#endif
#endif /* GCT_AMD_H */
