/*
 * Plain-C part of the deconwolf method module method_shb_hip.c: maps the
 * dw_opts fields that deconvolve_shb reads (src/dw.h:104-211) onto
 * dwhip_params, and writes the per-iteration log lines that deconwolf's
 * iterator writes (dw_iterator_show, src/dw.c:321-333).  Kept free of
 * deconwolf's headers so it builds and is tested on its own.
 */
#ifndef DWHIP_ADAPTER_H
#define DWHIP_ADAPTER_H

#include <stdio.h>

#include "dwhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The dw_opts values on the SHB path, copied field by field by the caller. */
typedef struct {
    int nIter;
    int iter_type;       /* dw_iter_type */
    int maxiter;
    float err_rel;
    float err_abs;
    float bg;
    int bg_auto;
    int borderQuality;
    int positivity;
    float alphamax;
    int metric;          /* dw_metric */
    int start_condition; /* dw_start_condition */
    int verbosity;
    int fftw3_planning;  /* FFTW planner flag: FFTW_ESTIMATE plans by rule, others measure */
} dwhip_dwopts_view;

/* FFTW_ESTIMATE of fftw3.h, the flag dw_opts.fftw3_planning holds for
 * --noplan (src/dw.c:1068) */
#define DWHIP_FFTW_ESTIMATE (1U << 6)

void dwhip_adapter_params(const dwhip_dwopts_view *v, dwhip_params *p);

/* Iterations the error trace must hold for these options. */
int dwhip_adapter_trace_len(const dwhip_dwopts_view *v);

/* "Iteration %3d/%3d, Idiv=%e" (or fMSE) per iteration, as deconwolf
 * writes them to its log file; niter_shown is the iterator's niter
 * (nIter for FIXED, maxiter otherwise, src/dw.c:228-231). */
void dwhip_adapter_log_iterations(FILE *log, const dwhip_dwopts_view *v, const float *err, int n);

#ifdef __cplusplus
}
#endif
#endif
