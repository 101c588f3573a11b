/*
 * deconwolf method module: the scaled-heavy-ball loop of deconvolve_shb
 * (src/method_shb.c:77-424) and the Richardson-Lucy loop of deconvolve_rl
 * (src/method_rl.c:105-400) run on an MI355X by libdwhip.
 *
 * Contract, as for deconvolve_shb and deconvolve_rl:
 *   - im is borrowed, already has the offset added (src/dw.c:2835-2838);
 *   - psf is owned and freed here (method_shb.c:103,191; method_rl.c:221),
 *     normalised to sum 1 and autocropped by dw_run;
 *   - the return value is a fresh M*N*P buffer from fim_malloc, released
 *     by the caller with fim_free;
 *   - s->bg receives the resolved background when s->bg_auto is set
 *     (method_shb.c:108-117);
 *   - the log gets the job size and one line per iteration, like
 *     deconvolve_shb and dw_iterator_show (src/dw.c:321-333).
 * Errors: any libdwhip error, a transform size it cannot run included,
 * prints the reason (and the job size) and exits with EXIT_FAILURE, as
 * deconwolf does on fatal errors (method_shb.c:199-203).  Nothing here
 * calls deconwolf's own CPU methods.
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "method_shb_hip.h"
#include "dw_util.h"
#include "fim.h"
#include "fim_tiff.h"
#include "dwhip.h"
#include "dwhip_adapter.h"

/* The side outputs deconvolve_shb writes from inside its loop, fed by
 * libdwhip's per-iteration hook:
 *   --iterdump k: the guess at the top of every k-th iteration, cropped,
 *     offset removed, as itd%05d.tif (method_shb.c:302-322; the reference
 *     then projects im instead of the dumped copy, :309, a no-op here);
 *   --tsv: iteration, seconds since the start, the iteration's error and
 *     KL = 0 (benchmark_write, dw.c:1339-1370; s->ref, the only source of a
 *     KL term, is set by no deconwolf caller);
 *   --fulldump: fulldump.tif, the returned guess, and fullPSF.tif, the
 *     shifted PSF, both at the transform size (method_shb.c:205-209,
 *     :407-411; with border quality 2 that size is the padded one, so the
 *     band outside the image differs from deconwolf's dump). */
typedef struct {
    dw_opts *s;
    ftif_t *ftif;
} dump_ctx;

static void on_iteration(void *user, int iter, float error, const float *x, int64_t nx, int64_t ny, int64_t nz)
{
    dump_ctx *c = (dump_ctx *)user;
    dw_opts *s = c->s;
    if (iter >= 0) {
        if (x != NULL && s->iterdump > 0) {
            float *temp = fim_copy(x, nx * ny * nz);
            if (s->offset > 0)
                fim_add_scalar(temp, nx * ny * nz, -s->offset);
            char *outname = gen_iterdump_name(s, iter);
            if (s->outFormat == 32)
                fim_tiff_write_float(c->ftif, outname, temp, NULL, nx, ny, nz);
            else
                fim_tiff_write(c->ftif, outname, temp, NULL, nx, ny, nz);
            free(outname);
            fim_free(temp);
        }
        if (s->tsv != NULL) {
            struct timespec now;
            dw_gettime(&now);
            fprintf(s->tsv, "%d\t%f\t%f\t%f\n", iter, clockdiff(&now, &s->tstart), (double)error, 0.0);
        }
    } else if (iter == -1) {
        printf("Dumping to fulldump.tif\n");
        fim_tiff_write(c->ftif, "fulldump.tif", x, NULL, nx, ny, nz);
    } else if (iter == -2) {
        printf("Dumping to fullPSF.tif\n");
        fim_tiff_write_float(c->ftif, "fullPSF.tif", x, NULL, nx, ny, nz);
    }
}

static float *run_hip(int rl, float *restrict im, const int64_t M, const int64_t N, const int64_t P,
                      float *restrict psf, const int64_t pM, const int64_t pN, const int64_t pP,
                      dw_opts *s)
{
    const char *tag = rl ? "rl_hip" : "shb_hip";
    if (s->nIter == 0) {
        fim_free(psf);
        return fim_copy(im, M * N * P);
    }

    dwhip_dwopts_view v;
    v.nIter = s->nIter;
    v.iter_type = (int)s->iter_type;
    v.maxiter = s->maxiter;
    v.err_rel = s->err_rel;
    v.err_abs = s->err_abs;
    v.bg = s->bg;
    v.bg_auto = s->bg_auto;
    v.borderQuality = s->borderQuality;
    v.positivity = s->positivity;
    v.alphamax = s->alphamax;
    v.metric = (int)s->metric;
    v.start_condition = (int)s->start_condition;
    v.verbosity = s->verbosity;
    v.fftw3_planning = s->fftw3_planning;

    dwhip_params p;
    dwhip_adapter_params(&v, &p);

    int64_t work[3], ref[3];
    const int wrc = dwhip_work_size(M, N, P, pM, pN, pP, s->borderQuality, work, ref);
    if (wrc != DWHIP_OK) {
        fprintf(stderr, "%s: job [%" PRId64 "x%" PRId64 "x%" PRId64 "] (border quality %d): %s\n", tag,
                ref[0], ref[1], ref[2], s->borderQuality, dwhip_strerror(wrc));
        if (s->log != NULL && s->log != stdout)
            fprintf(s->log, "%s: job [%" PRId64 "x%" PRId64 "x%" PRId64 "]: %s\n", tag, ref[0], ref[1], ref[2],
                    dwhip_strerror(wrc));
        exit(EXIT_FAILURE);
    }

    float *im_lp = NULL;
    if (s->start_condition == DW_START_LP) {
        /* method_shb.c:278-287: start from the image smoothed with sigma 8 */
        im_lp = fim_copy(im, M * N * P);
        fim_gsmooth(im_lp, M, N, P, 8);
        p.start_image = im_lp;
    }

    dump_ctx dc;
    dc.s = s;
    dc.ftif = NULL;
    if (s->iterdump > 0 || s->tsv != NULL || s->fulldump) {
        dc.ftif = fim_tiff_new(s->log, s->verbosity);
        p.hook = on_iteration;
        p.hook_user = &dc;
        p.hook_every = s->iterdump;
        p.hook_full = s->fulldump ? 1 : 0;
    }

    const int ntrace = dwhip_adapter_trace_len(&v);
    float *trace = calloc((size_t)ntrace, sizeof(float));
    float *out = fim_malloc(M * N * P * sizeof(float));
    dwhip_stats st;
    memset(&st, 0, sizeof(st));
    st.err_trace = trace;
    st.err_trace_len = ntrace;

    fprintf(s->log, "image: [%" PRId64 "x%" PRId64 "x%" PRId64 "]\n"
            "psf: [%" PRId64 "x%" PRId64 "x%" PRId64 "]\n"
            "job: [%" PRId64 "x%" PRId64 "x%" PRId64 "] (transformed as %" PRId64 "x%" PRId64 "x%" PRId64 ")\n",
            M, N, P, pM, pN, pP, ref[0], ref[1], ref[2], work[0], work[1], work[2]);

    const int rc = rl ? dwhip_rl_run(im, M, N, P, psf, pM, pN, pP, &p, out, &st)
                      : dwhip_shb_run(im, M, N, P, psf, pM, pN, pP, &p, out, &st);
    fim_free(psf);
    if (im_lp != NULL)
        fim_free(im_lp);
    if (dc.ftif != NULL)
        fim_tiff_destroy(dc.ftif);
    if (rc != DWHIP_OK) {
        fprintf(stderr, "%s: %s\n", tag, dwhip_strerror(rc));
        if (s->log != NULL)
            fprintf(s->log, "%s: %s\n", tag, dwhip_strerror(rc));
        exit(EXIT_FAILURE);
    }
    if (s->bg_auto) {
        s->bg = st.bg;
        fprintf(s->log, "Setting the background level to %f\n", s->bg);
    }
    dwhip_adapter_log_iterations(s->log, &v, trace, st.iterations);
    if (s->verbosity > 0)
        printf("%s: %d iterations, %.1f ms\n", tag, st.iterations, st.setup_ms + st.iter_ms);
    free(trace);
    return out;
}

float *deconvolve_shb_hip(float *restrict im, const int64_t M, const int64_t N, const int64_t P,
                          float *restrict psf, const int64_t pM, const int64_t pN, const int64_t pP,
                          dw_opts *s)
{
    return run_hip(0, im, M, N, P, psf, pM, pN, pP, s);
}

float *deconvolve_rl_hip(float *restrict im, const int64_t M, const int64_t N, const int64_t P,
                         float *restrict psf, const int64_t pM, const int64_t pN, const int64_t pP,
                         dw_opts *s)
{
    return run_hip(1, im, M, N, P, psf, pM, pN, pP, s);
}
