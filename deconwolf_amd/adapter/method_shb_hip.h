/*
 * deconwolf method module: scaled heavy ball and Richardson-Lucy on an
 * MI355X through libdwhip.  Same dw_function signature as deconvolve_shb
 * and deconvolve_rl (src/method_shb.h, src/method_rl.h, src/dw.h:89-91).  Compiled inside deconwolf's source tree (see the
 * integration notes of this repository).
 */
#ifndef method_shb_hip_h
#define method_shb_hip_h

#include "dw.h"

float *deconvolve_shb_hip(float *restrict im, const int64_t M, const int64_t N, const int64_t P,
                          float *restrict psf, const int64_t pM, const int64_t pN, const int64_t pP,
                          dw_opts *s);

float *deconvolve_rl_hip(float *restrict im, const int64_t M, const int64_t N, const int64_t P,
                         float *restrict psf, const int64_t pM, const int64_t pN, const int64_t pP,
                         dw_opts *s);

#endif
