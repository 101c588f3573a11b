/*
 * Optional float FFT backend of the oracle: the FFTW3 single-precision API
 * as exported by Intel MKL (libmkl_rt.so), loaded at run time.  deconwolf
 * supports MKL as its FFTW provider (makefile:46-48,266-276); the oracle
 * uses it only for the timed CPU baseline, where its speed is that of
 * deconwolf's own CPU path.  TEST INFRASTRUCTURE ONLY.
 *
 * The prototypes below are the public FFTW3 interface (fftw3.h is not
 * installed here): plans are opaque pointers, complex values are float[2].
 */
#include <dlfcn.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fft_oracle.h"

typedef void *ext_plan;
typedef ext_plan (*plan_r2c_3d_t)(int, int, int, float *, float *, unsigned);
typedef ext_plan (*plan_c2r_3d_t)(int, int, int, float *, float *, unsigned);
typedef void (*exec_r2c_t)(ext_plan, float *, float *);
typedef void (*exec_c2r_t)(ext_plan, float *, float *);
typedef void (*destroy_t)(ext_plan);
typedef int (*init_threads_t)(void);
typedef void (*nthreads_t)(int);

#define EXT_ESTIMATE (1U << 6)

static struct {
    int tried, ok, enabled, nthreads;
    void *h;
    plan_r2c_3d_t plan_r2c;
    plan_c2r_3d_t plan_c2r;
    exec_r2c_t exec_r2c;
    exec_c2r_t exec_c2r;
    destroy_t destroy;
    init_threads_t init_threads;
    nthreads_t with_nthreads;
    /* one cached plan pair */
    int64_t n[3];
    ext_plan r2c, c2r;
} E;

static int ext_load(void)
{
    if (E.tried)
        return E.ok;
    E.tried = 1;
    const char *names[] = {"libmkl_rt.so", "libmkl_rt.so.2", "libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so"};
    for (size_t i = 0; i < sizeof(names) / sizeof(names[0]) && !E.h; i++)
        E.h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
    if (!E.h)
        return 0;
    E.plan_r2c = (plan_r2c_3d_t)dlsym(E.h, "fftwf_plan_dft_r2c_3d");
    E.plan_c2r = (plan_c2r_3d_t)dlsym(E.h, "fftwf_plan_dft_c2r_3d");
    E.exec_r2c = (exec_r2c_t)dlsym(E.h, "fftwf_execute_dft_r2c");
    E.exec_c2r = (exec_c2r_t)dlsym(E.h, "fftwf_execute_dft_c2r");
    E.destroy = (destroy_t)dlsym(E.h, "fftwf_destroy_plan");
    E.init_threads = (init_threads_t)dlsym(E.h, "fftwf_init_threads");
    E.with_nthreads = (nthreads_t)dlsym(E.h, "fftwf_plan_with_nthreads");
    E.ok = E.plan_r2c && E.plan_c2r && E.exec_r2c && E.exec_c2r && E.destroy;
    if (E.ok && E.init_threads)
        E.init_threads();
    return E.ok;
}

/* 1: use the external float FFT when loadable; returns whether it is in use. */
int oracle_set_fft_backend(int external, int nthreads)
{
    E.enabled = external && ext_load();
    E.nthreads = nthreads > 0 ? nthreads : 1;
    return E.enabled;
}

static int ext_plans(int64_t nx, int64_t ny, int64_t nz)
{
    if (E.r2c && E.n[0] == nx && E.n[1] == ny && E.n[2] == nz)
        return 0;
    if (E.r2c) {
        E.destroy(E.r2c);
        E.destroy(E.c2r);
        E.r2c = E.c2r = NULL;
    }
    if (E.with_nthreads)
        E.with_nthreads(E.nthreads);
    size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
    float *a = malloc(sizeof(float) * 2 * nc);
    float *b = malloc(sizeof(float) * 2 * nc);
    /* FFTW order: slowest dimension first */
    E.r2c = E.plan_r2c((int)nz, (int)ny, (int)nx, a, b, EXT_ESTIMATE);
    E.c2r = E.plan_c2r((int)nz, (int)ny, (int)nx, b, a, EXT_ESTIMATE);
    free(a);
    free(b);
    if (!E.r2c || !E.c2r)
        return -1;
    E.n[0] = nx;
    E.n[1] = ny;
    E.n[2] = nz;
    return 0;
}

int oracle_ext_r2c(const float *in, int64_t nx, int64_t ny, int64_t nz, float *out)
{
    if (!E.enabled || ext_plans(nx, ny, nz) != 0)
        return -1;
    E.exec_r2c(E.r2c, (float *)in, out);
    return 0;
}

int oracle_ext_c2r(float *in, int64_t nx, int64_t ny, int64_t nz, float *out)
{
    if (!E.enabled || ext_plans(nx, ny, nz) != 0)
        return -1;
    E.exec_c2r(E.c2r, in, out);
    return 0;
}
