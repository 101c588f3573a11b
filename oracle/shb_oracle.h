/*
 * SHB oracle: a CPU restatement of deconwolf's scaled-heavy-ball loop
 * (src/method_shb.c:77-424) and of its Richardson-Lucy method
 * (src/method_rl.c:28-400).  TEST INFRASTRUCTURE ONLY: used by tests/,
 * the smoke check and the CPU-baseline leg of the benchmark as the
 * checker, never as the product.
 */
#ifndef DWHIP_ORACLE_SHB_H
#define DWHIP_ORACLE_SHB_H
#include <stdint.h>

/* dw_iter_type / dw_metric / dw_start_condition values (src/dw.h:64-83) */
enum { ORACLE_ITER_ABS = 0, ORACLE_ITER_REL = 1, ORACLE_ITER_FIXED = 2 };
enum { ORACLE_METRIC_MSE = 0, ORACLE_METRIC_IDIV = 1 };
enum { ORACLE_START_FLAT = 0, ORACLE_START_LP = 1, ORACLE_START_IDENTITY = 2 };

/* The dw_opts fields deconvolve_shb reads (src/dw.h:104-211). */
typedef struct {
    int niter;          /* nIter */
    int iter_type;      /* iter_type */
    int maxiter;        /* maxiter */
    float err_rel;      /* err_rel */
    float err_abs;      /* err_abs */
    float bg;           /* bg, used when bg_auto == 0 */
    int bg_auto;        /* bg_auto */
    int border_quality; /* borderQuality */
    int positivity;     /* positivity */
    float alphamax;     /* alphamax */
    int metric;         /* metric */
    int start;          /* start_condition */
    const void *start_image; /* M*N*P start for IDENTITY/LP (NULL: im) */
    int64_t work[3];    /* transform size; 0 = reference work size */
    int nthreads;       /* OpenMP threads; 0 = default */
} oracle_shb_opts;

int oracle_shb_f32(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                   int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o, float *out,
                   float *err_trace, double *stats);
int oracle_shb_f64(const double *im, int64_t M, int64_t N, int64_t P, const double *psf,
                   int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o, double *out,
                   float *err_trace, double *stats);
/* Richardson-Lucy (src/method_rl.c:105-400) on the same options; positivity
 * and alphamax are not read. */
int oracle_rl_f32(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                  int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o, float *out,
                  float *err_trace, double *stats);
int oracle_rl_f64(const double *im, int64_t M, int64_t N, int64_t P, const double *psf,
                  int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o, double *out,
                  float *err_trace, double *stats);
int oracle_shb_setup_f32(const float *im, int64_t M, int64_t N, int64_t P, const float *psf,
                         int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o,
                         float *cK, float *W, float *x0, double *stats);
int oracle_shb_setup_f64(const double *im, int64_t M, int64_t N, int64_t P, const double *psf,
                         int64_t pM, int64_t pN, int64_t pP, const oracle_shb_opts *o,
                         double *cK, double *W, double *x0, double *stats);
int oracle_num_threads(void);
#endif
