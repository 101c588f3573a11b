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
#include "fim.h"
#include "dwhip.h"
#include "dwhip_adapter.h"

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
