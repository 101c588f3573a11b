/*
 * dwhip: C-ABI of the MI355X (gfx950) scaled-heavy-ball Richardson-Lucy
 * (SHB) deconvolution engine.
 *
 * The engine is a drop-in for deconwolf's CPU method `deconvolve_shb`
 * (src/method_shb.c:77-424), whose signature is the plugin type
 * `dw_function` (src/dw.h:89-91).  A deconwolf method module
 * (deconwolf_amd/method_shb_hip.c) maps `dw_opts` onto `dwhip_params` and
 * calls dwhip_shb_run().  Plain pointers and sizes only: no HIP, torch or
 * deconwolf types appear here.
 *
 * Arrays are x-fastest: element (m, n, p) of an M x N x P volume sits at
 * m + n*M + p*M*N (src/dw.c:79-85).
 */
#ifndef DWHIP_H
#define DWHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define DWHIP_OK 0
#define DWHIP_ERR_ARG (-1)       /* invalid argument */
#define DWHIP_ERR_SIZE (-2)      /* transform size not supported */
#define DWHIP_ERR_HIP (-3)       /* HIP runtime error */
#define DWHIP_ERR_NOMEM (-4)     /* device allocation failed / over budget */
#define DWHIP_ERR_PSF (-5)       /* PSF maximum not movable to the origin (method_shb.c:199-203) */
#define DWHIP_ERR_NODEVICE (-6)  /* no usable gfx950 device */
#define DWHIP_ERR_STATE (-7)     /* call not valid in the object's state */
#define DWHIP_ERR_SCRATCH (-8)   /* job larger than DWHIP_SCRATCH_BYTES (no-allocation entry) */

/* ---- option values, numerically equal to deconwolf's enums ---------- */
/* dw_iter_type (src/dw.h:79-83) */
#define DWHIP_ITER_ABS 0
#define DWHIP_ITER_REL 1
#define DWHIP_ITER_FIXED 2
/* dw_metric (src/dw.h:74-77) */
#define DWHIP_METRIC_MSE 0
#define DWHIP_METRIC_IDIV 1
/* deconvolution method, the values of dw_method (src/dw.h:52-60:
 * DW_METHOD_RL = 0, DW_METHOD_ID = 1, DW_METHOD_SHB = 2): plain
 * Richardson-Lucy of method_rl.c or the scaled heavy-ball loop of
 * method_shb.c.  DW_METHOD_ID has no GPU path (DWHIP_ERR_ARG). */
#define DWHIP_METHOD_RL 0
#define DWHIP_METHOD_SHB 2
/* dw_start_condition (src/dw.h:64-72) */
#define DWHIP_START_FLAT 0
#define DWHIP_START_LP 1
#define DWHIP_START_IDENTITY 2

/*
 * The dw_opts fields deconvolve_shb reads (src/dw.h:104-211), resolved to
 * plain values by the caller.
 */
typedef struct {
    int niter;           /* nIter: iterations for DWHIP_ITER_FIXED */
    int iter_type;       /* iter_type */
    int maxiter;         /* maxiter: cap for REL / ABS */
    float err_rel;       /* err_rel */
    float err_abs;       /* err_abs */
    float bg;            /* bg, used when bg_auto == 0 */
    int bg_auto;         /* bg_auto: bg = max(min(im), 1e-2) (method_shb.c:108-117) */
    int border_quality;  /* borderQuality: 0, 1 or 2 (method_shb.c:137-153) */
    int positivity;      /* positivity: clamp x to bg after each update */
    float alphamax;      /* alphamax: cap of the heavy-ball weight */
    int metric;          /* DWHIP_METRIC_* */
    int start;           /* DWHIP_START_*; LP passes the smoothed image as start_image */
    const float *start_image; /* M*N*P start for IDENTITY/LP, NULL = im (same memory space as im) */
    int64_t work[3];     /* transform size; zeros = automatic (see dwhip_work_size) */
    int verbosity;       /* 0 silent; > 0 prints sizes and timings to stdout */
    int method;          /* DWHIP_METHOD_*: read by the engine and the tiled path;
                            dwhip_shb_run* and dwhip_rl_run* set it themselves */
    int planning;        /* DWHIP_PLAN_*: fftw3_planning (src/dw.h:201) */
    /* Per-iteration host hook, for deconwolf's --iterdump, --tsv and --fulldump
     * (method_shb.c:205-209,302-322,380,407-411); NULL for none.  With a hook every
     * pass is launched directly and the host waits for each iteration.  After
     * iteration k (k = 0, 1, ...) hook(hook_user, k, e_k, x, nx, ny, nz) is called
     * with e_k the iteration's error and x the iterate x_k (the guess at the top of
     * iteration k), cropped to M*N*P in host memory, when hook_every > 0 and
     * k % hook_every == 0, else x = NULL.  With hook_full != 0 two calls follow the
     * loop: k = -1 with the returned iterate at the transform size (work[0..2]
     * extents) and k = -2 with the PSF inserted at the transform size and shifted
     * so its maximum is at the origin (the volume deconvolve_shb transforms). */
    void (*hook)(void *user, int iter, float error, const float *x, int64_t nx, int64_t ny, int64_t nz);
    void *hook_user;
    int hook_every;
    int hook_full;
} dwhip_params;

/* Transform planning, the analogue of deconwolf's fftw3_planning flag
 * (FFTW_ESTIMATE / FFTW_MEASURE, src/dw.c:400,1068): RULE plans by rule
 * (and uses a plan measured before, from the plan cache); MEASURE also
 * measures the plans of a new transform size on its first use
 * (dwhip_plan_tune) and caches the result. */
#define DWHIP_PLAN_RULE 0
#define DWHIP_PLAN_MEASURE 1

/* Results besides the output volume. */
typedef struct {
    float bg;              /* background level used (written back to s->bg by the caller) */
    int iterations;        /* iterations run */
    int64_t work[3];       /* transform size used */
    int64_t ref_work[3];   /* deconwolf's work size for these inputs */
    double setup_ms;       /* OTF, weights and start guess */
    double iter_ms;        /* all iterations */
    float *err_trace;      /* caller buffer (host), one error per iteration, or NULL */
    int err_trace_len;     /* its length */
} dwhip_stats;

/* Fills p with deconwolf's defaults (dw_opts_new, src/dw.c:340-413) and
 * niter iterations in FIXED mode. */
void dwhip_params_default(dwhip_params *p, int niter);

/* Text for a status code. */
const char *dwhip_strerror(int code);

/* Selects device 0 and checks it is usable.  Called lazily by the entry
 * points; returns DWHIP_OK or DWHIP_ERR_NODEVICE. */
int dwhip_init(void);

/*
 * Transform size for an M*N*P image and a pM*pN*pP PSF.  ref_work receives
 * deconwolf's work size (method_shb.c:137-153); work receives the size the
 * engine transforms: for border_quality 2 the smallest size >= ref_work
 * with an even x extent whose halves and the y, z extents factor into 2, 3,
 * 5 and 7 (the image region is unchanged by that padding); for quality 0
 * and 1 the reference size itself, run by the fused passes when it factors
 * that way, else by the any-size path (odd extents and any prime factor,
 * through Bluestein's form; extents up to about 4800).
 * Returns DWHIP_OK or DWHIP_ERR_SIZE (an extent beyond both).
 */
int dwhip_work_size(int64_t M, int64_t N, int64_t P, int64_t pM, int64_t pN, int64_t pP,
                    int border_quality, int64_t *work, int64_t *ref_work);

/*
 * Measured planning (the analogue of FFTW's planner and wisdom file,
 * src/fft.c:135-172,206-319).  dwhip_plan_tune times, axis by axis, the
 * radix orders and tile shapes the kernels can run for the transform size
 * work (as dwhip_work_size returns it) against the rule planner's plan,
 * keeps a candidate only when it is more than 3 % faster, and stores the
 * plan in the process's plan cache and its file; retune != 0 measures a
 * size already cached.  tuned_ms / rule_ms (may be NULL) receive the
 * summed pass times of the chosen and of the rule's plan.  The cache file
 * is $DWHIP_PLAN_CACHE, else $HOME/.config/deconwolf/dwhip_plans_gfx950.txt
 * (where deconwolf keeps its FFTW wisdom) when $HOME/.config exists;
 * dwhip_plan_cache_path sets another ("" for none, NULL for the default)
 * and loads it.  dwhip_plan_get writes the plan a job of size work uses:
 * radices[12] (x-row, y and z forward radices, 4 each, 0-padded) and
 * tiles[3] (x rows, y and z columns per workgroup); returns 1 for a
 * measured plan, 0 for the rule's, or an error.  Sizes the fused passes
 * cannot run give DWHIP_ERR_SIZE (the any-size path has one plan).
 */
int dwhip_plan_tune(const int64_t *work, int retune, double *tuned_ms, double *rule_ms);
int dwhip_plan_cache_path(const char *path);
int dwhip_plan_get(const int64_t *work, int *radices, int *tiles);

/* Device bytes one job of these sizes needs. */
size_t dwhip_job_bytes(const int64_t *work, int64_t M, int64_t N, int64_t P, int64_t pM,
                       int64_t pN, int64_t pP, int border_quality, int niter);

/*
 * SHB deconvolution, host buffers: the drop-in for deconvolve_shb
 * (src/method_shb.c:77-424, called through dw_function at src/dw.c:2841 and
 * src/dw.c:2137).
 *   im   M*N*P image with deconwolf's offset already added (borrowed)
 *   psf  pM*pN*pP PSF, normalised to sum 1 and autocropped (borrowed; the
 *        adapter frees it, as deconvolve_shb does at method_shb.c:191)
 *   out  M*N*P output; receives the iterate deconvolve_shb returns: with
 *        n iterations run, x_{n-1} (swap-back at method_shb.c:385-390);
 *        with niter == 0 in FIXED mode a copy of im (method_shb.c:101-105)
 * Returns a DWHIP_* status.
 */
int dwhip_shb_run(const float *im, int64_t M, int64_t N, int64_t P, const float *psf, int64_t pM,
                  int64_t pN, int64_t pP, const dwhip_params *params, float *out,
                  dwhip_stats *stats);

/*
 * Same, with im, psf, start_image and out in device memory, work queued on
 * `stream` (a hipStream_t, NULL = default stream); returns after the
 * result is in out.  No counterpart in deconwolf: it is the entry for
 * callers whose volumes are already resident on the GPU.
 */
int dwhip_shb_run_device(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                         int64_t pM, int64_t pN, int64_t pP, const dwhip_params *params,
                         float *out, dwhip_stats *stats, void *stream);

/*
 * dwhip_shb_run_device without any device allocation.  No counterpart in
 * deconwolf: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The job's state lives in a static device block of
 * DWHIP_SCRATCH_BYTES per GPU, held until the process ends; a job needing
 * more (dwhip_job_bytes) returns DWHIP_ERR_SCRATCH before anything runs
 * (256 MiB holds, for instance, a 128x128x32 image with a 127x127x63 PSF,
 * transformed as 256x256x96).  Every call on a GPU shares that block:
 * calls that overlap on one GPU race.  setup_ms and iter_ms read 0.
 */
#define DWHIP_SCRATCH_BYTES ((size_t)256 << 20)
int dwhip_shb_run_device_noalloc(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                                 int64_t pM, int64_t pN, int64_t pP, const dwhip_params *params,
                                 float *out, dwhip_stats *stats, void *stream);

/*
 * Richardson-Lucy deconvolution, host buffers: the drop-in for
 * deconvolve_rl (src/method_rl.c:105-400), the other dw_function of the
 * same shape.  Same arguments as dwhip_shb_run; out receives the last
 * iterate x_n (no swap-back); positivity and alphamax are not read, x is
 * clamped to bg when bg > 0 (method_rl.c:357-366); the ratio outside the
 * image is 1e-6 and g/y <= 0 gives bg (method_rl.c:45-65).
 */
int dwhip_rl_run(const float *im, int64_t M, int64_t N, int64_t P, const float *psf, int64_t pM,
                 int64_t pN, int64_t pP, const dwhip_params *params, float *out,
                 dwhip_stats *stats);

/* Same on device memory and a caller stream, as dwhip_shb_run_device. */
int dwhip_rl_run_device(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                        int64_t pM, int64_t pN, int64_t pP, const dwhip_params *params,
                        float *out, dwhip_stats *stats, void *stream);

/*
 * Output path of dw_run (src/dw.c:2862-2904) on the device.
 *
 * dwhip_output_stats_device: largest value of v (NaN skipped), from which
 * the 16-bit scaling is (2^16-1)/max (scaling_for_u16, src/dw.c:98-103;
 * 65535/max for tiled runs, src/dw.c:2166-2172), and the number of voxels
 * equal to bg (get_nbg, src/dw.c:2406-2418, logged as the percentage of
 * pixels at the background level).  Synchronises `stream`.
 *
 * dwhip_to_u16_device: out[i] = the uint16 that deconwolf's writer `mode`
 * stores for v[i] with this scaling:
 *   DWHIP_U16_ROUND  TIFF writer fim_tiff_write_opt (src/fim_tiff.c:1138-1150)
 *   DWHIP_U16_TRUNC  .npy writer fim_imwrite_u16 (src/fim.c:5142-5147) and the
 *                    tiled TIFF writer (src/fim_tiff.c:819-822)
 *   DWHIP_U16_RINT   tiled .npy writer raw_to_npio (src/dw.c:187-190)
 * Values the C conversions leave undefined (outside [0, 65536), NaN)
 * saturate to [0, 65535], NaN to 0.  Queued on `stream`.
 */
#define DWHIP_U16_ROUND 0
#define DWHIP_U16_TRUNC 1
#define DWHIP_U16_RINT 2
int dwhip_output_stats_device(const float *v, int64_t n, float bg, float *vmax, int64_t *n_bg,
                              void *stream);
int dwhip_to_u16_device(const float *v, int64_t n, float scaling, int mode, uint16_t *out,
                        void *stream);

/*
 * Born-Wolf widefield PSF on the device, the model of deconwolf's dw_bw
 * (src/dw_bwpsf.c:538-775, src/bw_gsl.c:79-157): out receives N*N*P floats
 * (x fastest), centred at ((N-1)/2, (N-1)/2, (P-1)/2), not normalised, for
 * wavelength lambda_nm, numerical aperture NA, immersion index ni and pixel
 * sizes dxy_nm (lateral) and dz_nm (axial).  The radial profile
 * |int_0^1 J0(k0 NA r rho) exp(-i k0 NA^2 z rho^2 / (2 ni)) rho drho|^2 is
 * sampled at 17 radii per pixel by Gauss-Legendre quadrature; each pixel is
 * the mean over its square (8 x 8 Gauss-Legendre nodes) of the Lanczos-5
 * interpolation of that profile, 0 beyond dw_bw's radmax; octant and z
 * symmetry as dw_bw.  dw_bw integrates adaptively with GSL to 1e-6.
 * Synchronises `stream`.
 */
int dwhip_bw_psf_device(float *out, int64_t N, int64_t P, double lambda_nm, double NA, double ni,
                        double dxy_nm, double dz_nm, void *stream);

/*
 * FFT check entry (device memory): unnormalised 3-D R2C of an nx*ny*nz
 * real volume (nx even) into (nx/2+1)*ny*nz complex, interleaved, the
 * layout of fftwf_plan_dft_r2c_3d(nz, ny, nx) that src/fft.c:236-290
 * plans; inverse != 0 runs the C2R direction instead (spec -> vol).
 * No counterpart in deconwolf's method API; it exposes the transforms the
 * loop uses so that they can be checked alone.
 */
int dwhip_fft3_device(float *vol, float *spec, int64_t nx, int64_t ny, int64_t nz, int inverse,
                      void *stream);

/*
 * Batched and served execution.  No counterpart in deconwolf, which runs
 * one job (or one tile, src/dw.c:2093-2153) at a time; this engine is for
 * callers with many jobs: tiles of one image, fields and channels of an
 * acquisition, or a request stream.
 *
 * An engine owns `hbm_budget` bytes of device memory.  Submitted jobs
 * (device pointers; the caller keeps im, psf, start image and out alive
 * until the job is reported done) wait in a FIFO queue and are admitted
 * while their state fits in the budget and fewer than `max_active` run.
 * Every dwhip_engine_step() admits what fits, sets the new jobs up, runs
 * SHB iterations of every active job (one launch per pass for all of them;
 * one round, or up to rounds_per_step while every active job is FIXED), and
 * completes the jobs that reached their iteration count or stopping rule:
 * their result is in `out` when the step returns and their memory is free
 * for the next admissions.  Results equal those of
 * dwhip_shb_run_device for the same job, whatever else runs beside it.
 * Not thread-safe: one host thread drives an engine.
 */
typedef struct dwhip_engine dwhip_engine;

/* admission order of queued jobs */
#define DWHIP_ORDER_FIFO 0      /* arrival order */
#define DWHIP_ORDER_SHORTEST 1  /* least work (job voxels x iterations) first */

typedef struct {
    size_t hbm_budget;  /* device bytes for job state (allocated at creation) */
    int max_active;     /* jobs iterating together */
    int order;          /* DWHIP_ORDER_*; default FIFO */
    int active_mvox;    /* cap on the transform voxels (millions) of the jobs iterating
                           together, 0 = none; a job larger than the cap runs alone */
    int groups;         /* 1 (default) or 2: the active jobs split in two groups of about
                           equal voxels whose rounds run on two streams, so that passes
                           of one group overlap those of the other */
    int rounds_per_step; /* rounds one dwhip_engine_step may run (default 8, 0 reads as 1):
                           while every active job is FIXED, a step runs rounds up to the
                           next completion, all but the last replayed from one captured
                           HIP graph of the round, with one host synchronisation */
    int aging_ms;       /* DWHIP_ORDER_SHORTEST: a job that has waited aging_ms or more
                           goes before every younger job (oldest first), and younger
                           jobs do not overtake it while it waits for room, so large
                           jobs are not passed over for ever; 0 = least work first only */
    int reserved[1];
} dwhip_engine_config;

void dwhip_engine_config_default(dwhip_engine_config *cfg);
int dwhip_engine_create(const dwhip_engine_config *cfg, void *stream, dwhip_engine **engine);
void dwhip_engine_destroy(dwhip_engine *engine);
/* Queues a job; `tag` identifies it in results.  DWHIP_ERR_NOMEM when the
 * job alone exceeds the budget, DWHIP_ERR_SIZE / DWHIP_ERR_ARG as for
 * dwhip_shb_run. */
int dwhip_engine_submit(dwhip_engine *engine, const float *im, int64_t M, int64_t N, int64_t P,
                        const float *psf, int64_t pM, int64_t pN, int64_t pP,
                        const dwhip_params *params, float *out, int64_t tag);
/* One round; writes up to max_done tags of jobs completed in this round. */
int dwhip_engine_step(dwhip_engine *engine, int64_t *done_tags, int max_done, int *n_done);
int dwhip_engine_status(const dwhip_engine *engine, int *queued, int *active, size_t *bytes_in_use);
/* Stats of a completed job (bg, iterations, sizes, error trace); release
 * != 0 forgets the job afterwards.  DWHIP_ERR_STATE if not completed. */
int dwhip_engine_result(dwhip_engine *engine, int64_t tag, dwhip_stats *stats, int release);

/*
 * XY tiling, deconwolf's tiling.c and the tile loop of deconvolve_tiles
 * (src/dw.c:2000-2229), with the image and tiles in device memory instead
 * of raw temporary files.  pos is a tile's core box and xpos the box grown
 * by the overlap and clamped to the image (inclusive bounds x0 x1 y0 y1 z0
 * z1); xsize = xpos extents.
 */
typedef struct {
    int64_t pos[6];
    int64_t xpos[6];
    int64_t xsize[3];
} dwhip_tile;

/* tiling_create (src/tiling.c:70-130): ntiles receives the count; tiles
 * (may be NULL to ask for the count) receives them, x-divisions outer. */
int dwhip_tiling_create(int64_t M, int64_t N, int64_t P, int64_t max_size, int64_t overlap,
                        dwhip_tile *tiles, int max_tiles, int *ntiles);
/* Copies tile t's grown box out of im (M*N*P) into out (xsize volume) and
 * adds offset when offset > 0 (src/dw.c:2113,2131-2134). */
int dwhip_tile_extract_device(const float *im, int64_t M, int64_t N, int64_t P, const dwhip_tile *t,
                              float offset, float *out, void *stream);
/* Blends deconvolved tiles (outs[i]: tile i's method output, offset still
 * added) into out (M*N*P): tile outputs get -offset and are clipped at 0
 * when offset > 0 (src/dw.c:2141-2145), then out = sum over tiles, in
 * order, of tile_getWeight / tiling_getWeights times the tile
 * (src/tiling.c:197-223,353-410). */
int dwhip_tiles_blend_device(const dwhip_tile *tiles, const float *const *outs, int ntiles, int64_t M,
                             int64_t N, int64_t P, float offset, float *out, void *stream);

/*
 * Kernel timing for measurement.  While enabled, every pass launch is
 * bracketed by HIP events on the stream it runs on and its algorithmic
 * bytes (each operand read or written once) are added up per kernel.
 * dwhip_profile_enable(1) clears the totals.  Kernel ids 0..15 name the
 * X-pass modes (x_psf, x_ones, x_weights, x_update, x_ratio, x_load,
 * x_store) and column modes (y_fwd, y_inv, z_otf, z_conv, z_convj, z_fwd,
 * z_inv, y_inv_img, y_fwd_img).  No counterpart in deconwolf (its timing is the --times flag,
 * src/dw.c:2233-2319).
 */
int dwhip_profile_enable(int on);
int dwhip_profile_get(int kernel, const char **name, long long *launches, double *total_ms,
                      double *bytes);

#ifdef __cplusplus
}
#endif
#endif
