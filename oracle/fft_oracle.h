/* CPU FFT of the SHB oracle.  Test infrastructure only (see fft_oracle.c). */
#ifndef DWHIP_ORACLE_FFT_H
#define DWHIP_ORACLE_FFT_H
#include <stddef.h>
#include <stdint.h>

typedef struct cfft_plan cfft_plan;

cfft_plan *oracle_cfft_plan(int n);
void oracle_cfft_free(cfft_plan *p);
size_t oracle_cfft_work_size(const cfft_plan *p);
/* In-place complex FFT of interleaved doubles; sign -1 forward, +1 inverse
 * (unnormalised).  `work` holds oracle_cfft_work_size() complex values. */
void oracle_cfft_exec(const cfft_plan *p, double *data, double *work, int sign);
void oracle_cfft(double *data, int n, int sign);

/* x-fastest real volume -> (nx/2+1)*ny*nz complex (interleaved). */
void oracle_rfft3(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out);
/* inverse of the above, unnormalised; `in` is not modified. */
void oracle_irfft3(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out);

/* Optional external FFT (fft_ext.c): each returns 0 when it ran. */
int oracle_set_fft_backend(int external, int nthreads);
int oracle_ext_r2c(const float *in, int64_t nx, int64_t ny, int64_t nz, float *out);
int oracle_ext_c2r(float *in, int64_t nx, int64_t ny, int64_t nz, float *out);
int oracle_ext_r2c_d(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out);
int oracle_ext_c2r_d(double *in, int64_t nx, int64_t ny, int64_t nz, double *out);
#endif
