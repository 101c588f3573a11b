/*
 * SHB oracle: CPU restatement of deconwolf's deconvolve_shb
 * (src/method_shb.c:77-424) in float32 and float64.
 *
 * TEST INFRASTRUCTURE ONLY.  tests/, the smoke check and the benchmark's
 * CPU-baseline leg are the only users; the product never links it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fft_oracle.h"
#include "shb_oracle.h"

#include <stdio.h>

/* With ORACLE_PROGRESS=k in the environment, a line on stderr every k
 * iterations (long full-size runs report as they go). */
static void oracle_progress(int nrun)
{
    static int every = -1;
    if (every < 0) {
        const char *v = getenv("ORACLE_PROGRESS");
        every = v ? atoi(v) : 0;
    }
    if (every > 0 && nrun % every == 0) {
        fprintf(stderr, "oracle: %d iterations\n", nrun);
        fflush(stderr);
    }
}

static double oracle_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int oracle_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

#define REAL float
#define SUFFIX f32
#define ORACLE_IS_F32 1
#include "shb_core.inc"
#undef REAL
#undef SUFFIX
#undef ORACLE_IS_F32

#define REAL double
#define SUFFIX f64
#define ORACLE_IS_F32 0
#include "shb_core.inc"
#undef REAL
#undef SUFFIX
#undef ORACLE_IS_F32
