/* Plain-C part of method_shb_hip.c (see dwhip_adapter.h). */
#include "dwhip_adapter.h"

void dwhip_adapter_params(const dwhip_dwopts_view *v, dwhip_params *p)
{
    dwhip_params_default(p, v->nIter);
    p->niter = v->nIter;
    p->iter_type = v->iter_type;
    p->maxiter = v->maxiter;
    p->err_rel = v->err_rel;
    p->err_abs = v->err_abs;
    p->bg = v->bg;
    p->bg_auto = v->bg_auto;
    p->border_quality = v->borderQuality;
    p->positivity = v->positivity;
    p->alphamax = v->alphamax;
    p->metric = v->metric;
    p->start = v->start_condition;
    p->start_image = NULL;
    p->work[0] = p->work[1] = p->work[2] = 0;
    p->verbosity = v->verbosity > 2 ? 1 : 0;
    p->planning = (unsigned)v->fftw3_planning == DWHIP_FFTW_ESTIMATE ? DWHIP_PLAN_RULE : DWHIP_PLAN_MEASURE;
    p->hook = NULL;
    p->hook_user = NULL;
    p->hook_every = 0;
    p->hook_full = 0;
}

int dwhip_adapter_trace_len(const dwhip_dwopts_view *v)
{
    int n = v->iter_type == DWHIP_ITER_FIXED ? v->nIter : v->maxiter;
    return n > 0 ? n : 1;
}

void dwhip_adapter_log_iterations(FILE *log, const dwhip_dwopts_view *v, const float *err, int n)
{
    if (log == NULL || log == stdout || err == NULL)
        return;
    const int shown = v->iter_type == DWHIP_ITER_FIXED ? v->nIter : v->maxiter;
    for (int k = 0; k < n; k++) {
        if (v->metric == DWHIP_METRIC_MSE)
            fprintf(log, "Iteration %3d/%3d, fMSE=%e\n", k + 1, shown, err[k]);
        else
            fprintf(log, "Iteration %3d/%3d, Idiv=%e\n", k + 1, shown, err[k]);
    }
    fflush(log);
}
