/*
 * Optional FFT backend of the oracle: the FFTW3 single- and double-precision
 * APIs as exported by Intel MKL (libmkl_rt.so), loaded at run time.
 * deconwolf supports MKL as its FFTW provider (makefile:46-48,266-276); the
 * oracle uses it for the timed CPU baseline, where its speed is that of
 * deconwolf's own CPU path, and for the full-size parity runs, float32 and
 * float64, whose own-FFT transforms would take too long at 200 M voxels.
 * TEST INFRASTRUCTURE ONLY.
 *
 * The prototypes below are the public FFTW3 interface (fftw3.h is not
 * installed here): plans are opaque pointers, complex values are float[2]
 * (fftwf_*) or double[2] (fftw_*).
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
typedef ext_plan (*plan_r2c_3d_d_t)(int, int, int, double *, double *, unsigned);
typedef ext_plan (*plan_c2r_3d_d_t)(int, int, int, double *, double *, unsigned);
typedef void (*exec_r2c_d_t)(ext_plan, double *, double *);
typedef void (*exec_c2r_d_t)(ext_plan, double *, double *);

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
    /* double precision (fftw_*): optional, used when enabled_d */
    int ok_d, enabled_d;
    plan_r2c_3d_d_t plan_r2c_d;
    plan_c2r_3d_d_t plan_c2r_d;
    exec_r2c_d_t exec_r2c_d;
    exec_c2r_d_t exec_c2r_d;
    destroy_t destroy_d;
    init_threads_t init_threads_d;
    nthreads_t with_nthreads_d;
    int64_t nd[3];
    ext_plan r2c_d, c2r_d;
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
    E.plan_r2c_d = (plan_r2c_3d_d_t)dlsym(E.h, "fftw_plan_dft_r2c_3d");
    E.plan_c2r_d = (plan_c2r_3d_d_t)dlsym(E.h, "fftw_plan_dft_c2r_3d");
    E.exec_r2c_d = (exec_r2c_d_t)dlsym(E.h, "fftw_execute_dft_r2c");
    E.exec_c2r_d = (exec_c2r_d_t)dlsym(E.h, "fftw_execute_dft_c2r");
    E.destroy_d = (destroy_t)dlsym(E.h, "fftw_destroy_plan");
    E.init_threads_d = (init_threads_t)dlsym(E.h, "fftw_init_threads");
    E.with_nthreads_d = (nthreads_t)dlsym(E.h, "fftw_plan_with_nthreads");
    E.ok_d = E.plan_r2c_d && E.plan_c2r_d && E.exec_r2c_d && E.exec_c2r_d && E.destroy_d;
    if (E.ok_d && E.init_threads_d)
        E.init_threads_d();
    return E.ok;
}

/* external 1: the float oracle's transforms through fftwf_* when loadable;
 * external 2: also the double oracle's through fftw_*.  Returns the level in
 * use (0, 1 or 2). */
int oracle_set_fft_backend(int external, int nthreads)
{
    E.enabled = external > 0 && ext_load();
    E.enabled_d = external > 1 && E.enabled && E.ok_d;
    E.nthreads = nthreads > 0 ? nthreads : 1;
    return E.enabled ? (E.enabled_d ? 2 : 1) : 0;
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

static int ext_plans_d(int64_t nx, int64_t ny, int64_t nz)
{
    if (E.r2c_d && E.nd[0] == nx && E.nd[1] == ny && E.nd[2] == nz)
        return 0;
    if (E.r2c_d) {
        E.destroy_d(E.r2c_d);
        E.destroy_d(E.c2r_d);
        E.r2c_d = E.c2r_d = NULL;
    }
    if (E.with_nthreads_d)
        E.with_nthreads_d(E.nthreads);
    size_t nc = (size_t)(nx / 2 + 1) * ny * nz;
    double *a = malloc(sizeof(double) * 2 * nc);
    double *b = malloc(sizeof(double) * 2 * nc);
    E.r2c_d = E.plan_r2c_d((int)nz, (int)ny, (int)nx, a, b, EXT_ESTIMATE);
    E.c2r_d = E.plan_c2r_d((int)nz, (int)ny, (int)nx, b, a, EXT_ESTIMATE);
    free(a);
    free(b);
    if (!E.r2c_d || !E.c2r_d)
        return -1;
    E.nd[0] = nx;
    E.nd[1] = ny;
    E.nd[2] = nz;
    return 0;
}

int oracle_ext_r2c_d(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out)
{
    if (!E.enabled_d || ext_plans_d(nx, ny, nz) != 0)
        return -1;
    E.exec_r2c_d(E.r2c_d, (double *)in, out);
    return 0;
}

int oracle_ext_c2r_d(double *in, int64_t nx, int64_t ny, int64_t nz, double *out)
{
    if (!E.enabled_d || ext_plans_d(nx, ny, nz) != 0)
        return -1;
    E.exec_c2r_d(E.c2r_d, in, out);
    return 0;
}
