/*
 * Declaration-only stand-in for the two FFTW3 types that deconwolf's
 * headers name in struct fields and prototypes (src/dw.h:199-200,
 * src/fft.h:42-45): enough to compile the method module
 * deconwolf_amd/adapter/method_shb_hip.c with -fsyntax-only against
 * deconwolf's own dw.h and fim.h.  It declares no functions and is never
 * linked or used to build anything that runs.
 */
#ifndef DWHIP_TEST_FFTW3_SHIM_H
#define DWHIP_TEST_FFTW3_SHIM_H
typedef float fftwf_complex[2];
typedef struct fftwf_plan_s *fftwf_plan;
#endif
