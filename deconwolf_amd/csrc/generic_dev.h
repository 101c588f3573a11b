// Device descriptors of the any-size path: deconwolf's SHB and RL loops for
// work sizes the fused passes cannot transform (border quality 0 and 1 at
// deconwolf's exact size, whose extents may hold any prime factor or be
// odd).  Each axis is a complex FFT of its length n: Stockham passes in LDS
// with radices 2, 3, 4, 5 and 7 when n factors into them, and Bluestein's
// chirp-z form otherwise (a length-m FFT pair with m >= 2n - 1 that does
// factor).  The volumes are real, so x is a real transform of any length:
// two real rows as the real and imaginary parts of one complex row, their
// half spectra (nx/2 + 1 bins) separated after the forward transform and
// recombined before the inverse; y and z run over those bins only.
#pragma once
#include <stdint.h>

#include "shb_dev.h"

#define GEN_MAX_STAGES 16
// radix-8 stages in the any-size FFTs (a build tunable)
#ifndef DWHIP_GEN_R8
#define DWHIP_GEN_R8 0
#endif

struct GenPlanDev {
    int n;           // transform length along the axis
    int m;           // FFT length run in LDS (n, or the Bluestein length)
    int bluestein;   // 1: chirp-z through a length-m FFT pair
    int B;           // pencils per workgroup
    int nst;         // stages of the length-m FFT
    int radix[GEN_MAX_STAGES];
    int ns[GEN_MAX_STAGES];     // product of the radices before the stage
    int twoff[GEN_MAX_STAGES];  // twiddles of the stage (forward sign) in tw
    int twlen;       // stage twiddles in all, contiguous from twoff[0]
    int twlds;       // 1: a workgroup copies them to LDS beside its pencils
    int chirp;       // n values c_j = exp(-i pi j^2 / n) in tw (Bluestein)
    int hf, hi;      // m values each: FFT_m(conj(c) / m) and FFT_m(c / m), circularly extended
};

struct GenDev {
    int nx, ny, nz;
    int hc, hs;       // x bins kept (nx/2 + 1) and the complex row stride of Z and K (>= hc)
    int M, N, P;
    int pM, pN, pP;
    int rl, positivity, metric, has_w;
    float alphamax;
    float2* Z;        // half spectra [nz][ny][hs] (bins < hc), or real rows of 2*hs floats in place
    float2* K;        // OTF / (nx*ny*nz), half spectrum [nz][ny][hs]
    float* U0;        // x_k lives in U[k & 1], nx*ny*nz
    float* U1;
    float* W;         // Bertero weights (has_w)
    const float* im;  // M*N*P image
    const float* psf; // pM*pN*pP PSF
    double* part;     // one partial sum per workgroup of the ratio kernel
    JobState* st;
    const float2* tw; // twiddles, chirps and Bluestein filters
    GenPlanDev plan[3];
};
