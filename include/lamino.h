/*
 * lamino.h -- C ABI of the MI355X (gfx950) vector-laminography library.
 *
 * Each entry point replaces one function of the reference (lbl-camera/lamino)
 * on its reconstruction hot path; the comment on each cites the interface it
 * stands in for.  Conventions:
 *   - volumes are row-major float32 [n1][n2][n3] (dims_t, include/dtypes.h:31-84);
 *     dims[] = {n1, n2, n3};
 *   - projections are float32 [nproj][nrows][ncols];
 *   - complex arrays are interleaved float32 pairs (std::complex<float>);
 *   - unless the name says "host", array arguments are DEVICE pointers and
 *     the work is ordered on the context's stream;
 *   - every call returns LAM_OK or an error code; lam_last_error() gives the
 *     message, and the C++ layer (include/tomocam/) rethrows the exception
 *     the reference would throw.
 */
#ifndef LAMINO_H
#define LAMINO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    LAM_OK = 0,
    LAM_ERR_ARG = 1,      /* std::invalid_argument */
    LAM_ERR_HIP = 2,      /* HIP runtime failure */
    LAM_ERR_NOMEM = 3,    /* std::bad_alloc (include/fft.h:69) */
    LAM_ERR_NUFFT = 4,    /* "Error in finufft_*" (include/nufft.h:48,63) */
    LAM_ERR_SHAPE = 5,    /* split-Bregman x0/yT mismatch (src/bregman.cpp:44-48) */
    LAM_ERR_TYPE = 6,     /* NUFFT type not in {1,2} (include/finufft_plan_cache.h:63-64) */
    LAM_ERR_DIVZERO = 7,  /* "Division by zero" (include/array.h:163-165) */
    LAM_ERR_STATE = 8,    /* job / handle used in the wrong state */
    LAM_ERR_BUDGET = 9    /* request exceeds the device memory budget */
};
typedef int lam_status;

typedef struct lam_ctx lam_ctx;    /* device, stream, plan cache, scratch */
typedef struct lam_grid lam_grid;  /* polar grid: per-projection plane geometry */
typedef struct lam_toep lam_toep;  /* Toeplitz normal operator (sum over grids) */

/* Message of the last failed call on this thread. */
const char *lam_last_error(void);

/* ------------------------------------------------------------------ context */
lam_status lam_ctx_create(int device, lam_ctx **out);
lam_status lam_ctx_destroy(lam_ctx *ctx);
/* Use an external hipStream_t (NULL restores the library's own stream). */
lam_status lam_ctx_set_stream(lam_ctx *ctx, void *hip_stream);
void *lam_ctx_stream(lam_ctx *ctx);
lam_status lam_ctx_sync(lam_ctx *ctx);

/* Device memory helpers (stream-ordered copies on the context's stream). */
lam_status lam_malloc(lam_ctx *ctx, size_t bytes, void **ptr);
lam_status lam_free(lam_ctx *ctx, void *ptr);
lam_status lam_upload(lam_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
lam_status lam_download(lam_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
lam_status lam_memset(lam_ctx *ctx, void *dst_dev, int value, size_t bytes);
lam_status lam_copy_device(lam_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* -------------------------------------------------------------------- grids */
/* PolarGrid<T>(angles, nrows, ncols, gamma), include/polar_grid.h:44-75.
 * Points are generated analytically on the device; nothing of size P*R*C is stored. */
lam_status lam_grid_polar(lam_ctx *ctx, const double *theta_host, int64_t nproj, int64_t nrows, int64_t ncols,
                          double gamma, lam_grid **out);
/* Per-projection generalisation (alpha_p, gamma_p): the reference ctor is alpha = theta,
 * gamma_p = gamma.  Used for the laminography scan (alpha = tilt, gamma = offset + rotation). */
lam_status lam_grid_polar_pp(lam_ctx *ctx, const double *alpha_host, const double *gamma_host, int64_t nproj,
                             int64_t nrows, int64_t ncols, lam_grid **out);
lam_status lam_grid_destroy(lam_grid *grid);
lam_status lam_grid_shape(const lam_grid *grid, int64_t *nproj, int64_t *nrows, int64_t *ncols);
/* Materialise PolarGrid's public x, y, z fields (host float arrays of P*R*C),
 * computed as the reference computes them (include/polar_grid.h:61-78). */
lam_status lam_grid_coords_host(const lam_grid *grid, float *x, float *y, float *z);

/* --------------------------------------------------------------------- NUFFT */
/* nufft::nufft3d1 (include/nufft.h:37-50): f_k = sum_p c_p exp(+i k.x_p);
 * c: complex [P][R][C], f: complex volume dims.  eps <= 0 selects the default (1e-5). */
lam_status lam_nufft3d1(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const void *c, void *f);
/* nufft::nufft3d2 (include/nufft.h:52-65): c_p = sum_k f_k exp(-i k.x_p). */
lam_status lam_nufft3d2(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, void *c, const void *f);

/* --------------------------------------------------------------- projectors */
/* forward (include/projection.h:52-54, src/projection.cpp:41-77) */
lam_status lam_forward(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *m0,
                       const float *m1, const float *m2, float *proj);
/* adjoint (include/projection.h:64-66, src/projection.cpp:89-126) */
lam_status lam_adjoint(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *proj,
                       float *m0, float *m1, float *m2);
/* adjoint with accumulation into m (the yT += adjoint(...) loop of src/mbir_*.cpp) */
lam_status lam_adjoint_accumulate(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps,
                                  const float *proj, float *m0, float *m1, float *m2);
/* The datasets loop of the MBIR drivers in one call (src/mbir3.cpp:66-83,
 * src/mbir_sb.cpp:60-78: one forward / adjoint per gamma dataset of one volume):
 * the de-apodisation and the fine-grid 3-D transform run once for all grids.
 * lam_forward_multi: proj[k] = forward(m, grids[k]).  lam_adjoint_multi:
 * m = sum_k adjoint(proj[k], grids[k]) (accumulate: m +=); the grids share the
 * detector size. */
lam_status lam_forward_multi(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3], double eps,
                             const float *m0, const float *m1, const float *m2, float *const *proj);
lam_status lam_adjoint_multi(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3], double eps,
                             const float *const *proj, float *m0, float *m1, float *m2, int accumulate);
/* sysmat (include/tomocam.h:50-52, src/gradient.cpp:33-89), NUFFT form */
lam_status lam_sysmat(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, const float *const x[3],
                      float *const y[3]);

/* ---------------------------------------------------- Toeplitz normal operator */
/* Builds the exact (closed-form) Toeplitz kernels of sum_g sysmat_g for the
 * given grids.  keep_per_grid != 0 additionally keeps each grid's kernels for
 * lam_toep_quad (the per-dataset residual of src/gradient.cpp:115-127). */
lam_status lam_toep_create(lam_ctx *ctx, lam_grid *const *grids, int ngrids, const int64_t dims[3],
                           int keep_per_grid, lam_toep **out);
lam_status lam_toep_destroy(lam_toep *op);
/* y = sum_g sysmat_g(x) (+ mu * laplacian(x) per component when mu != 0) */
lam_status lam_toep_apply(lam_toep *op, const float *const x[3], float *const y[3], float mu);
/* q[g*3+i] = x_i . sysmat_g(x)_i for every grid g (host output, FP64). */
lam_status lam_toep_quad(lam_toep *op, const float *const x[3], double *q_host);

/* ---------------------------------------------------------------- spectral */
/* fft::fft2 (sign -1) / fft::ifft2 (sign +1), include/fft.h:34-114: batched,
 * unnormalised 2-D complex transform over (n2, n3) for each of n1 slices.
 * shifted != 0 wraps it as ifftshift2(fft(fftshift2(x))) (src/projection.cpp:73-75). */
lam_status lam_fft2(lam_ctx *ctx, const int64_t dims[3], int sign, int shifted, const void *in, void *out);

/* ------------------------------------------------------ stencils, elementwise */
/* opt::RampPreconditioner<T>(dims).apply(in) (include/precond.h:33-87) */
lam_status lam_precond(lam_ctx *ctx, const int64_t dims[3], const float *in, float *out);
/* opt::grad_u (src/finitediff.cpp:7-47) */
lam_status lam_grad_u(lam_ctx *ctx, const int64_t dims[3], const float *u, float *g0, float *g1, float *g2);
/* opt::divergence (src/finitediff.cpp:53-71) */
lam_status lam_divergence(lam_ctx *ctx, const int64_t dims[3], const float *v0, const float *v1, const float *v2,
                          float *out);
/* opt::laplacian = divergence(grad_u(u)) (include/bregman.h:33-36), fused */
lam_status lam_laplacian(lam_ctx *ctx, const int64_t dims[3], const float *u, float *out);
/* opt::qggmrf (src/qggmrf.cpp:59-95): g += prior gradient, in place */
lam_status lam_qggmrf(lam_ctx *ctx, const int64_t dims[3], float *g, const float *x, float sigma, float p);
/* array::dot / norm2 / max (include/array_ops.h:63-93), FP64 accumulation, fixed order */
lam_status lam_dot(lam_ctx *ctx, int64_t n, const float *a, const float *b, double *out_host);
lam_status lam_max(lam_ctx *ctx, int64_t n, const float *a, float *out_host);
/* lam_dot without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under hipGraph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same kernels, launch
 * shape and result as lam_dot; the FP64 partial sums live in a static per-device
 * array of 1088 doubles (8.5 KiB, held until the process ends) instead of context
 * memory.  n in [0, 2^40], else LAM_ERR_ARG.  Calls that overlap on one device
 * share that array and race: serialise them (one stream, or wait in between). */
lam_status lam_dot_noalloc(lam_ctx *ctx, int64_t n, const float *a, const float *b, double *out_host);

/* ------------------------------------------------------------------- solvers */
/* Operator callback: y = A(x) on device arrays, stream-ordered on the context stream. */
typedef lam_status (*lam_op_fn)(void *user, const float *const x[3], float *const y[3]);
/* Loss callback (host scalar). */
typedef lam_status (*lam_loss_fn)(void *user, const float *const x[3], double *loss);

/* opt::cgsolver (include/optimize.h:89-91, src/conjgrad.cpp:36-117); FP64 scalars. */
lam_status lam_cgsolver(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, const float *const b[3],
                        const float *const x0[3], int64_t max_iters, double tol, float *const x[3],
                        int64_t *iters_done);
/* opt::split_bregman (include/optimize.h:74-76, src/bregman.cpp:38-134) */
lam_status lam_split_bregman(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, const float *const yT[3],
                             const float *const x0[3], double lambda, double mu, int64_t outer_max,
                             int64_t inner_max, double tol, double xtol, float *const x[3]);
/* opt::nagopt, vector form (include/optimize.h:107-112, src/nagopt.cpp:115-209);
 * x is updated in place as in the reference. */
lam_status lam_nagopt(lam_ctx *ctx, const int64_t dims[3], lam_op_fn grad, lam_loss_fn loss, void *user,
                      float *const x[3], int64_t max_iters, double lipschitz, double tol, double xtol,
                      int64_t max_inner);
/* opt::lipschitz (include/optimize.h:140-142, src/lipschitz.cpp:30-53): FP64 power iteration. */
lam_status lam_lipschitz(lam_ctx *ctx, const int64_t dims[3], lam_op_fn A, void *user, int64_t max_iters,
                         double tol, uint64_t seed, double *L);

/* Operator adapter over a Toeplitz operator for the solvers above:
 * user = lam_toep*, y = sum_g sysmat_g(x).  (No counterpart in the reference,
 * whose operators are std::function closures: src/mbir_cg.cpp:96-105.) */
lam_status lam_toep_op(void *user, const float *const x[3], float *const y[3]);

/* Spreader/interpolator autotuner (no reference counterpart: FINUFFT picks its
 * bins internally).  Times candidate tile shapes (z, y, x) of this grid's
 * three-component spread + interpolation on the fine grid of dims and keeps
 * the fastest for every later call with the same fine grid and detector, in
 * this process.  tile_out: the chosen shape; best_ms: its spread + interp time.
 * Grids whose projections share the tilt use the row-factorised projectors,
 * whose tiling follows from the fine grid: tile_out is then (Tz, W, Tx), the
 * (z, x) tile and the kernel width, and best_ms their measured time. */
lam_status lam_autotune_bins(lam_ctx *ctx, lam_grid *grid, const int64_t dims[3], double eps, int tile_out[3],
                             double *best_ms);

/* ---------------------------------------------------------- instrumentation */
/* HIP-event timing of the library's own kernels on the context stream.
 * Enabling clears earlier records.  (No reference counterpart; the reference
 * only has a wall-clock Timer, include/timer.h:31-64.) */
lam_status lam_timing_enable(lam_ctx *ctx, int on);
/* Restrict the timing to the scopes of one kernel name (NULL or "": all).
 * Each timed scope adds two event records to the stream (about 5 us per
 * scope on MI355X), so a throughput measurement times only the kernel it
 * reports on. */
lam_status lam_timing_only(lam_ctx *ctx, const char *name);
/* Total device milliseconds and launch count recorded under a kernel name:
 * "interp", "spread", "bin_points", "fft_axis", "deapod_pad", "crop_deapod",
 * "toep_build", "toep_mul". */
lam_status lam_timing_query(lam_ctx *ctx, const char *name, double *total_ms, int64_t *count);

/* ---------------------------------------------------------- reconstruction jobs */
typedef struct {
    int64_t nproj, nrows, ncols;  /* projection stack dims */
    const float *proj;            /* HOST [nproj][nrows][ncols] */
    const double *alpha;          /* HOST per projection (the reference's angles) */
    const double *gamma;          /* HOST per projection, or NULL to use gamma_const */
    double gamma_const;
} lam_dataset;

typedef struct {
    int method;             /* 1 MBIR1 (CG), 2 MBIR2 (split-Bregman TV), 3 MBIR3 (NAG + qGGMRF) */
    int64_t max_iters;      /* ReconParams::maxIters */
    int64_t inner_iters;    /* ReconParams::innerIters */
    double lambda, mu, sigma, p, tol, xtol;   /* MBIR3 with mu <= 0: mu is the normal operator's
                                                  Lipschitz constant, estimated on the device */
    double pad_factor;      /* ReconParams::PAD_FACTOR (1.4142) */
    double eps;             /* NUFFT tolerance for the backprojection (<= 0: default) */
} lam_recon_params;

typedef struct {
    double setup_s;         /* normalisation, padding, backprojection, Toeplitz build */
    double solve_s;         /* solver loop */
    int64_t normal_ops;     /* normal-operator applications */
    int64_t cg_iters;       /* inner CG iterations executed */
    int64_t outer_iters;    /* outer iterations executed */
    double lipschitz;       /* MBIR3: the Lipschitz constant passed to nagopt (mu, or the estimate) */
} lam_recon_stats;

/* MBIR1/2/3 (include/tomocam.h:101-154; src/mbir_cg.cpp, src/mbir_sb.cpp, src/mbir3.cpp).
 * out3_host receives the 3 cropped components [3][n1][n2][n3]. */
lam_status lam_mbir(lam_ctx *ctx, const lam_dataset *datasets, int ndatasets, const int64_t recon_dims[3],
                    const lam_recon_params *params, float *out3_host, lam_recon_stats *stats);

/* Upper bound of the device memory one lam_mbir job of these shapes allocates. */
lam_status lam_mbir_footprint(const lam_dataset *datasets, int ndatasets, const int64_t recon_dims[3],
                              const lam_recon_params *params, size_t *bytes);

/* ------------------------------------------------------- batched / served jobs
 * (No counterpart in the reference, which runs one job per process,
 * src/recon.cpp:65-73.)  A server owns `slots` worker threads on one device,
 * each with its own stream, and admits submitted jobs in arrival order as a
 * slot frees and the HBM budget allows (lam_mbir_footprint per job).  Jobs on
 * different slots overlap on the device; each job keeps its iterate, Bregman
 * auxiliaries and Toeplitz kernels resident until it finishes.  A job's
 * result does not depend on what runs beside it.  Each worker's stream should
 * have a hardware queue of its own: run a serving process with
 * GPU_MAX_HW_QUEUES >= slots + 1 (HIP's default, 4, lets two workers share a
 * queue, and their work then runs one after the other). */
typedef struct lam_server lam_server;

enum {
    LAM_SERVE_PLAIN = 0,   /* the lam_mbir solver loop (host reads per stopping test) */
    LAM_SERVE_ASYNC = 1,   /* solver scalars and stopping tests on the device */
    LAM_SERVE_GRAPH = 2    /* ASYNC, with each iteration replayed from a captured hipGraph */
};

typedef struct {
    int64_t id;
    int status;                /* LAM_OK or the job's error code */
    int slot;                  /* worker that ran it */
    double submit_s;           /* seconds since the server was created */
    double start_s;            /* admitted */
    double end_s;              /* result written to out3_host */
    size_t footprint_est;      /* admission estimate (bytes) */
    size_t footprint_peak;     /* measured peak of the job's device buffers (bytes) */
    lam_recon_stats stats;
} lam_job_info;

typedef struct {
    int64_t submitted, completed, failed, running, queued;
    size_t budget, in_use, peak_in_use;
    int slots, mode;
} lam_server_stats;

lam_status lam_server_create(int device, size_t hbm_budget, int slots, int mode, lam_server **out);
/* Finishes every queued job, then stops the workers. */
lam_status lam_server_destroy(lam_server *server);
/* Queues a job; projections and angles are copied, out3_host must stay valid
 * until lam_server_wait returns for the job.  Rejects (LAM_ERR_BUDGET) a job
 * whose footprint exceeds the whole budget. */
lam_status lam_server_submit(lam_server *server, const lam_dataset *datasets, int ndatasets,
                             const int64_t recon_dims[3], const lam_recon_params *params, float *out3_host,
                             int64_t *job_id);
/* File output of a served job (src/recon.cpp:76-110): the result leaves the
 * device as array::transpose(., {1, 2, 0}) when transpose120 is set (the recon
 * CLI's order), goes to pinned memory, and a writer thread writes the files
 * while the worker goes on with the next job; the job finishes when they are
 * written.  Formats: tiff::write3 (<path without .tif*>{0,1,2}.tiff), VTI
 * (<path>.vti), tiff::write_vectors (path itself). */
enum { LAM_OUT_TIFF = 1, LAM_OUT_VTI = 2, LAM_OUT_TIFF_VECTORS = 4 };
typedef struct {
    const char *path;
    int formats;          /* LAM_OUT_* flags */
    int transpose120;
} lam_output;
/* lam_server_submit with file output; out3_host (may be NULL) also receives the
 * written result.  The strings are copied. */
lam_status lam_server_submit_output(lam_server *server, const lam_dataset *datasets, int ndatasets,
                                    const int64_t recon_dims[3], const lam_recon_params *params,
                                    const lam_output *output, float *out3_host, int64_t *job_id);
/* Blocks until the job has finished; its error, if any, is the return value.
 * A job is waited on once: its record is released then (a second wait on the
 * same id is LAM_ERR_ARG, unknown job id). */
lam_status lam_server_wait(lam_server *server, int64_t job_id, lam_job_info *info);
lam_status lam_server_query(lam_server *server, lam_server_stats *stats);

/* ------------------------------------------------------- request ingest and output
 * Host-side readers and writers of the reference's file formats (include/tiff.h,
 * include/write_vti.h, include/config.h), for turning files into lam_dataset
 * requests and results into files.  Paths are UTF-8; arrays are host memory. */

/* Stack dims (pages, rows, cols) tiff::read returns (include/tiff.h:31-60): an
 * even image width drops the last row and an even height the last column. */
lam_status lam_tiff_info(const char *path, int64_t dims[3]);
/* tiff::read (include/tiff.h:31-85): the first dims[1] rows and dims[2] columns
 * of every page of 32-bit samples, taken as float; dims from lam_tiff_info. */
lam_status lam_tiff_read(const char *path, float *out_host, const int64_t dims[3]);
/* tiff::write (include/tiff.h:87-126): one float page per slice, one strip each. */
lam_status lam_tiff_write(const char *path, const float *data_host, const int64_t dims[3]);
/* tiff::write_vectors (include/tiff.h:128-174): 3 float samples per pixel, LZW. */
lam_status lam_tiff_write_vectors(const char *path, const float *c0, const float *c1, const float *c2,
                                  const int64_t dims[3]);
/* tiff::write3 (include/tiff.h:176-188): files <name without .tif*>{0,1,2}.tiff. */
lam_status lam_tiff_write3(const char *name, const float *c0, const float *c1, const float *c2,
                           const int64_t dims[3]);
/* vti::write_vectors (include/write_vti.h:36-111): <path with extension .vti>,
 * appended raw Float64 vectors, origin at the volume centre; field_name NULL:
 * "magnetization". */
lam_status lam_vti_write_vectors(const char *path, const float *c0, const float *c1, const float *c2,
                                 const int64_t dims[3], const char *field_name);
/* read_angles_file<float> (include/config.h:61-83): out NULL queries the count. */
lam_status lam_read_angles(const char *path, double *out, int64_t cap, int64_t *n);
/* The forward-simulation CLI's computation (src/forward.cpp:123-183): the three
 * components as read from TIFF ([d0][d1][d2] host arrays) are transposed
 * {2, 0, 1}, padded by sqrt(2) - 1 (pad3d), projected on
 * PolarGrid(angles, padded(rows), padded(cols), gamma), and the projections are
 * cropped back to rows x cols: proj_host [nangles][rows][cols].  eps <= 0: 1e-5. */
lam_status lam_forward_simulate(lam_ctx *ctx, const float *m0, const float *m1, const float *m2,
                                const int64_t dims[3], const double *angles, int64_t nangles, double gamma,
                                int64_t rows, int64_t cols, double eps, float *proj_host);
/* opt::demag (include/demag.h:13-79): the demagnetising field
 * H = F^-1[-k (k . F m) / |k|^2] of a device volume (DC zeroed), k in numpy's
 * fftfreq order on every axis (on an even last axis the reference's r2c
 * transform takes +n/2 for the Nyquist plane; that plane is where they differ). */
lam_status lam_demag(lam_ctx *ctx, const int64_t dims[3], const float *m0, const float *m1, const float *m2,
                     float *h0, float *h1, float *h2);
/* mask_infs_nans (include/mask.h:22-32) in place on n device floats. */
lam_status lam_mask_infs_nans(lam_ctx *ctx, int64_t n, float *data);
/* array::transpose(a, {1, 2, 0}) (include/array_ops.h:117-137), the output step of
 * src/recon.cpp:76-78: in [n1][n2][n3] -> out [n2][n3][n1], device memory. */
lam_status lam_transpose120(lam_ctx *ctx, const int64_t dims[3], const float *in, float *out);

#ifdef __cplusplus
}
#endif

#endif /* LAMINO_H */
