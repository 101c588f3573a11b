// HIP kernels of the SHB (scaled heavy ball Richardson-Lucy) loop for
// gfx950.  Restates deconwolf's src/method_shb.c (deconvolve_shb :77-424,
// iter_shb :10-74) as fused passes over a job table (shb_dev.h).
//
// A 3-D R2C/C2R transform of an nx*ny*nz volume is three passes:
//   X: x-rows, real length nx packed as h = nx/2 complex values, complex
//      FFT of length h plus the split step to h+1 bins;
//   Y: columns along y of the (h+1)*ny*nz spectrum, tiles of adjacent kx;
//   Z: columns along z, tiles of adjacent kx.
// One SHB iteration k runs eight passes (A..H):
//   A  X: [C2R of the previous adjoint -> x_k = xt*p_{k-1}*W, clamp]
//         p_k = max(x_k + a_k (x_k - x_{k-1}), bg); R2C(p_k)
//   B  Y forward
//   C  Z forward, * OTF, Z inverse
//   D  Y inverse
//   E  X: C2R -> y = H p_k; error terms; ratio im/y; R2C(ratio)
//   F  Y forward
//   G  Z forward, * conj(OTF), Z inverse
//   H  Y inverse
// and the next iteration's A finishes the update.  The 1/(nx*ny*nz) of both
// inverse transforms is folded into the stored OTF.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "fft_lds.h"
#include "shb_dev.h"
#include "shb_kernels.h"

#define NTHR DWHIP_NTHR

__device__ __forceinline__ int find_job(const JobDev* __restrict__ jobs, int njobs, int kind, int t)
{
    int j = 0;
    while (j + 1 < njobs && t >= jobs[j + 1].tile0[kind]) ++j;
    return j;
}

// alpha of method_shb.c:328-330, in double like the reference.
__device__ __forceinline__ double shb_alpha(int iter, float alphamax)
{
    double a = ((float)iter - 1.0) / ((float)iter + 2.0);
    if (a < 0) a = 0;
    if (a > alphamax) a = alphamax;
    return a;
}

// p = x + alpha*(x - xp), then p = max(p, bg)  (method_shb.c:333-339):
// x - xp in float, the rest in double, rounded to float.
__device__ __forceinline__ float shb_extrap(float x, float xp, double alpha, float bg)
{
    float p = (float)((double)x + alpha * (double)(x - xp));
    return p < bg ? bg : p;
}

__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NTHR / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// X pass
// ---------------------------------------------------------------------------
// C2R split step: builds Z[k] (length h) from the h+1 bins of a row so that
// IFFT_h(Z)[n] = out[2n] + i*out[2n+1] of the unnormalised C2R.
//   Z[k] = (X[k] + conj(X[h-k])) + i*exp(+2 pi i k/nx)*(X[k] - conj(X[h-k]))
__device__ __forceinline__ float2 c2r_split(float2 a, float2 b, float2 w)
{
    // a = X[k], b = X[h-k], w = exp(-2 pi i k/nx)
    float2 s = make_float2(a.x + b.x, a.y - b.y);   // a + conj(b)
    float2 d = make_float2(a.x - b.x, a.y + b.y);   // a - conj(b)
    float2 e = cmulc(d, w);                          // d * exp(+2 pi i k/nx)
    return make_float2(s.x - e.y, s.y + e.x);        // s + i*e
}

// R2C merge step: X[k] from Z[k], Z[h-k] of the packed length-h FFT.
//   X[k] = (Z[k] + conj(Z[h-k]))/2 - i/2 * exp(-2 pi i k/nx) * (Z[k] - conj(Z[h-k]))
__device__ __forceinline__ float2 r2c_merge(float2 a, float2 b, float2 w)
{
    float2 s = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    float2 d = make_float2(0.5f * (a.x - b.x), 0.5f * (a.y + b.y));
    float2 e = cmul(d, w);
    return make_float2(s.x + e.y, s.y - e.x);        // s - i*e
}

template <int MODE>
__global__ __launch_bounds__(NTHR) void k_xpass(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs)
{
    extern __shared__ float2 lds[];
    __shared__ double red[NTHR / 64];
    const int tid = threadIdx.x;
    const int total = bd->ntile[KIND_X];
    const int njobs = bd->njobs;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const JobDev& J = jobs[find_job(jobs, njobs, KIND_X, t)];
        const int lt = t - J.tile0[KIND_X];
        const int h = J.h, nx = J.nx, ny = J.ny, nz = J.nz, hcs = J.hcs;
        const int row0 = lt * J.xrows;
        const int nrows = min(J.xrows, ny * nz - row0);
        const int iter = (MODE == XM_UPDATE || MODE == XM_RATIO) ? J.st->iter : 0;
        const bool inv = MODE == XM_RATIO || MODE == XM_WEIGHTS || MODE == XM_STORE ||
                         (MODE == XM_UPDATE && iter > 0);
        const int hp = h / 2 + 1;
        const uint32_t hpmagic = fdiv_magic((uint32_t)hp);
        const uint32_t hmagic = fdiv_magic((uint32_t)h);

        if (inv) {
            for (int w = tid; w < nrows * hp; w += NTHR) {
                const int b = (int)fdiv((uint32_t)w, hpmagic);
                const int k = w - b * hp;
                const float2* crow = J.C + (size_t)(row0 + b) * hcs;
                const float2 a = crow[k], c = crow[h - k];
                float2* L = lds + b * h;
                L[k] = c2r_split(a, c, J.rtw[k]);
                if (k > 0 && 2 * k != h) L[h - k] = c2r_split(c, a, J.rtw[h - k]);
            }
            __syncthreads();
            fft_lds<DWHIP_XMAXE, +1, false>(lds, J.px, J.tw, nrows, h, tid, NTHR);
        }

        double acc = 0;
        const int nwork = nrows * h;
        for (int w = tid; w < nwork; w += NTHR) {
            const int b = (int)fdiv((uint32_t)w, hmagic);
            const int n = w - b * h;
            const int row = row0 + b;
            const int yy = row % ny, zz = row / ny;
            const int m0 = 2 * n;
            const size_t vi = (size_t)row * nx + m0;
            float2 out;
            if (MODE == XM_PSF) {
                const JobState* st = J.st;
                const int sy = (yy + st->mid[1]) % ny, sz = (zz + st->mid[2]) % nz;
                float v[2] = {0.f, 0.f};
                if (sy < J.pN && sz < J.pP) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int sx = (m0 + e + st->mid[0]) % nx;
                        if (sx < J.pM) v[e] = J.psf[sx + (size_t)J.pM * (sy + (size_t)J.pN * sz)];
                    }
                }
                out = make_float2(v[0], v[1]);
            } else if (MODE == XM_LOAD) {
                out = *reinterpret_cast<const float2*>(J.U0 + vi);
            } else if (MODE == XM_STORE) {
                out = lds[b * h + n];
                *reinterpret_cast<float2*>(J.U0 + vi) = out;
            } else if (MODE == XM_ONES) {
                const bool inr = yy < J.N && zz < J.P;
                out = make_float2(inr && m0 < J.M ? 1.f : 0.f, inr && m0 + 1 < J.M ? 1.f : 0.f);
            } else if (MODE == XM_WEIGHTS) {
                // method_shb.c:225-234: W = P1 > 0.01 ? 1/P1 : 0
                const float2 p1 = lds[b * h + n];
                const float sigma = 0.01f;
                float2 wv;
                wv.x = p1.x > sigma ? (float)(1.0 / p1.x) : 0.f;
                wv.y = p1.y > sigma ? (float)(1.0 / p1.y) : 0.f;
                *reinterpret_cast<float2*>(J.W + vi) = wv;
                out = wv;
            } else if (MODE == XM_UPDATE) {
                const float bg = J.st->bg;
                float* Uk = (iter & 1) ? J.U1 : J.U0;          // holds x_{k-2}, receives x_k
                const float* Ukm1 = (iter & 1) ? J.U0 : J.U1;  // holds x_{k-1}
                float2 xk, xkm1;
                if (iter == 0) {
                    xk = *reinterpret_cast<const float2*>(J.U0 + vi);
                    xkm1 = *reinterpret_cast<const float2*>(J.U1 + vi);
                } else {
                    // finish iteration k-1 (method_shb.c:55-70, 357-376)
                    const float2 xt = lds[b * h + n];
                    xkm1 = *reinterpret_cast<const float2*>(Ukm1 + vi);
                    const float2 xkm2 = *reinterpret_cast<const float2*>(Uk + vi);
                    const double am1 = shb_alpha(iter - 1, J.alphamax);
                    const float p0 = shb_extrap(xkm1.x, xkm2.x, am1, bg);
                    const float p1 = shb_extrap(xkm1.y, xkm2.y, am1, bg);
                    if (J.has_w) {
                        const float2 wv = *reinterpret_cast<const float2*>(J.W + vi);
                        xk = make_float2(xt.x * (p0 * wv.x), xt.y * (p1 * wv.y));
                    } else {
                        xk = make_float2(xt.x * p0, xt.y * p1);
                    }
                    if (J.positivity) {
                        xk.x = xk.x < bg ? bg : xk.x;
                        xk.y = xk.y < bg ? bg : xk.y;
                    }
                    *reinterpret_cast<float2*>(Uk + vi) = xk;
                }
                const double a = shb_alpha(iter, J.alphamax);
                out = make_float2(shb_extrap(xk.x, xkm1.x, a, bg), shb_extrap(xk.y, xkm1.y, a, bg));
            } else {  // XM_RATIO, method_shb.c:27-48 and getError dw.c:1436-1524
                const float2 yv = lds[b * h + n];
                float r[2] = {0.f, 0.f};
                if (yy < J.N && zz < J.P) {
                    const float* imrow = J.im + (size_t)J.M * (yy + (size_t)J.N * zz);
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int m = m0 + e;
                        if (m < J.M) {
                            float y = e ? yv.y : yv.x;
                            const float g = imrow[m];
                            const double yd = y, gd = g;
                            if (J.metric == DWHIP_DEV_METRIC_IDIV) {
                                if (yd > 0 && gd > 0) acc += gd * (double)logf((float)(gd / yd)) - (gd - yd);
                            } else {
                                acc += (yd - gd) * (yd - gd);
                            }
                            if (fabs(yd) < 1e-6) y = (float)copysign(1e-6, yd);
                            r[e] = g / y;
                        }
                    }
                }
                out = make_float2(r[0], r[1]);
            }
            lds[b * h + n] = out;
        }
        if (MODE == XM_RATIO) {
            const double s = block_sum(acc, red);
            if (tid == 0) J.part[lt] = s;
        }
        __syncthreads();

        if (MODE != XM_WEIGHTS && MODE != XM_STORE) {
            fft_lds<DWHIP_XMAXE, -1, false>(lds, J.px, J.tw, nrows, h, tid, NTHR);
            float2* dst = J.C;
            for (int w = tid; w < nrows * hp; w += NTHR) {
                const int b = (int)fdiv((uint32_t)w, hpmagic);
                const int k = w - b * hp;
                const float2* L = lds + b * h;
                float2* crow = dst + (size_t)(row0 + b) * hcs;
                if (k == 0) {
                    const float2 z0 = L[0];
                    crow[0] = make_float2(z0.x + z0.y, 0.f);
                    crow[h] = make_float2(z0.x - z0.y, 0.f);
                } else {
                    const float2 a = L[k], c = L[h - k];
                    crow[k] = r2c_merge(a, c, J.rtw[k]);
                    if (2 * k != h) crow[h - k] = r2c_merge(c, a, J.rtw[h - k]);
                }
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// Column passes (Y and Z)
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(NTHR) void k_colpass(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs)
{
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    constexpr int kind = (MODE == CM_YFWD || MODE == CM_YINV) ? KIND_Y : KIND_Z;
    static_assert(MODE >= CM_YFWD && MODE <= CM_ZINV, "column mode");
    const int total = bd->ntile[kind];
    const int njobs = bd->njobs;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const JobDev& J = jobs[find_job(jobs, njobs, kind, t)];
        const int lt = t - J.tile0[kind];
        const int hcs = J.hcs, ny = J.ny;
        const int B = kind == KIND_Y ? J.ycols : J.zcols;
        const int nct = (J.hc + B - 1) / B;
        const int outer = lt / nct;
        const int kx0 = (lt - outer * nct) * B;
        const int nb = min(B, J.hc - kx0);   // columns past hc are row padding, never touched
        const int n = kind == KIND_Y ? ny : J.nz;
        const size_t stride = kind == KIND_Y ? (size_t)hcs : (size_t)hcs * ny;
        const size_t off = kind == KIND_Y ? (size_t)outer * ny * hcs + kx0 : (size_t)outer * hcs + kx0;
        const float2* src = J.C + off;
        // ZOTF stores the transformed, scaled PSF into K
        float2* dst = (MODE == CM_ZOTF ? J.K : J.C) + off;
        const FftPlanDev& pl = kind == KIND_Y ? J.py : J.pz;
        const uint32_t bmagic = fdiv_magic((uint32_t)B);

        for (int w = tid; w < n * B; w += NTHR) {
            const int i = (int)fdiv((uint32_t)w, bmagic);
            const int b = w - i * B;
            lds[w] = b < nb ? src[i * stride + b] : make_float2(0.f, 0.f);
        }
        __syncthreads();
        fft_lds<DWHIP_CMAXE, (MODE == CM_YINV || MODE == CM_ZINV) ? +1 : -1, true>(lds, pl, J.tw, B, 0, tid, NTHR);
        if (MODE == CM_ZCONV || MODE == CM_ZCONVJ || MODE == CM_ZOTF) {
            const float2* kb = J.K + off;
            // (float)nx*ny*nz, evaluated left to right in float (fft.c:856)
            const float scale = ((float)J.nx * (float)J.ny) * (float)J.nz;
            for (int w = tid; w < n * B; w += NTHR) {
                const int i = (int)fdiv((uint32_t)w, bmagic);
                const int b = w - i * B;
                if (b < nb) {
                    const float2 v = lds[w];
                    if (MODE == CM_ZOTF) {
                        lds[w] = make_float2(v.x / scale, v.y / scale);
                    } else {
                        const float2 kk = kb[i * stride + b];
                        // fft_mul_inplace (fft.c:437-453) / fft_mul_conj_inplace (:490-512)
                        const float a = kk.x, ac = MODE == CM_ZCONVJ ? -kk.y : kk.y;
                        lds[w] = make_float2(a * v.x - ac * v.y, a * v.y + v.x * ac);
                    }
                }
            }
            __syncthreads();
            if (MODE != CM_ZOTF) fft_lds<DWHIP_CMAXE, +1, true>(lds, pl, J.tw, B, 0, tid, NTHR);
        }
        for (int w = tid; w < n * B; w += NTHR) {
            const int i = (int)fdiv((uint32_t)w, bmagic);
            const int b = w - i * B;
            if (b < nb) dst[i * stride + b] = lds[w];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Setup statistics: min(im) and sum(im) (bg_auto, flat start) and argmax(psf)
// ---------------------------------------------------------------------------
// stat layout per job: [nblk][4] doubles: min, sum, max, argmax index
__global__ __launch_bounds__(NTHR) void k_stats_partial(const JobDev* __restrict__ jobs, int j, int nblk)
{
    const JobDev& J = jobs[j];
    __shared__ double smin[NTHR], ssum[NTHR], smax[NTHR], sidx[NTHR];
    const size_t n = (size_t)J.M * J.N * J.P;
    const size_t np = (size_t)J.pM * J.pN * J.pP;
    double vmin = INFINITY, vsum = 0, vmax = -INFINITY, vidx = 1e300;
    for (size_t i = (size_t)blockIdx.x * NTHR + threadIdx.x; i < n; i += (size_t)nblk * NTHR) {
        const double v = J.im[i];
        vmin = v < vmin ? v : vmin;
        vsum += v;
    }
    for (size_t i = (size_t)blockIdx.x * NTHR + threadIdx.x; i < np; i += (size_t)nblk * NTHR) {
        const double v = J.psf[i];
        if (v > vmax) {  // strictly greater keeps the lowest index per thread
            vmax = v;
            vidx = (double)i;
        }
    }
    const int tid = threadIdx.x;
    smin[tid] = vmin; ssum[tid] = vsum; smax[tid] = vmax; sidx[tid] = vidx;
    __syncthreads();
    for (int s = NTHR / 2; s > 0; s >>= 1) {
        if (tid < s) {
            smin[tid] = fmin(smin[tid], smin[tid + s]);
            ssum[tid] += ssum[tid + s];
            const double m2 = smax[tid + s], i2 = sidx[tid + s];
            if (m2 > smax[tid] || (m2 == smax[tid] && i2 < sidx[tid])) {
                smax[tid] = m2;
                sidx[tid] = i2;
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        double* o = J.stat + 4 * (size_t)blockIdx.x;
        o[0] = smin[0]; o[1] = ssum[0]; o[2] = smax[0]; o[3] = sidx[0];
    }
}

__global__ __launch_bounds__(NTHR) void k_stats_final(const JobDev* __restrict__ jobs, int j, int nblk)
{
    const JobDev& J = jobs[j];
    if (threadIdx.x != 0) return;
    double vmin = INFINITY, vsum = 0, vmax = -INFINITY, vidx = 1e300;
    for (int b = 0; b < nblk; ++b) {
        const double* o = J.stat + 4 * (size_t)b;
        vmin = fmin(vmin, o[0]);
        vsum += o[1];
        if (o[2] > vmax || (o[2] == vmax && o[3] < vidx)) {
            vmax = o[2];
            vidx = o[3];
        }
    }
    JobState* st = J.st;
    // bg_auto: bg = max(min(im), 1e-2)  (method_shb.c:108-117)
    float bg = J.bg_user;
    if (J.bg_auto) {
        bg = (float)vmin;
        if (bg < 1e-2) bg = (float)1e-2;
    }
    st->bg = bg;
    // flat start: (float)sum / (float)wMNP_reference  (method_shb.c:255,266)
    const float sumg = (float)vsum;
    st->c0 = sumg / (float)J.refMNP;
    const long long a = (long long)vidx;
    const long long pMN = (long long)J.pM * J.pN;
    st->mid[2] = (int)(a / pMN);
    st->mid[1] = (int)((a - (long long)st->mid[2] * pMN) / J.pM);
    st->mid[0] = (int)(a - (long long)st->mid[2] * pMN - (long long)st->mid[1] * J.pM);
    st->iter = 0;
}

// Start guess into U0 and U1 (method_shb.c:264-287).
__global__ __launch_bounds__(NTHR) void k_start(const JobDev* __restrict__ jobs, int j)
{
    const JobDev& J = jobs[j];
    const size_t n = (size_t)J.nx * J.ny * J.nz;
    const float c0 = J.st->c0;
    const float* src = J.start_img ? J.start_img : J.im;
    for (size_t i = (size_t)blockIdx.x * NTHR + threadIdx.x; i < n; i += (size_t)gridDim.x * NTHR) {
        float v = c0;
        if (J.start_identity) {
            const int m = (int)(i % J.nx);
            const size_t r = i / J.nx;
            const int yy = (int)(r % J.ny), zz = (int)(r / J.ny);
            v = (m < J.M && yy < J.N && zz < J.P) ? src[m + (size_t)J.M * (yy + (size_t)J.N * zz)] : 0.f;
        }
        J.U0[i] = v;
        J.U1[i] = v;
    }
}

// End of a round: error of the iteration from the X-tile partials, in a
// fixed order, then advance the iteration counter.  One block per job.
__global__ __launch_bounds__(NTHR) void k_round_end(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs, int advance)
{
    __shared__ double red[NTHR / 64];
    const JobDev& J = jobs[blockIdx.x];
    const int nt = J.ntile[KIND_X];
    double s = 0;
    for (int i = threadIdx.x; i < nt; i += NTHR) s += J.part[i];
    s = block_sum(s, red);
    if (threadIdx.x == 0) {
        JobState* st = J.st;
        const int it = st->iter;
        const double mean = s / (double)((long long)J.M * J.N * J.P);
        if (it < J.niter) {
            J.err[it] = (float)mean;
            J.errd[it] = mean;
        }
        if (advance) st->iter = it + 1;
    }
}

// Crop x_{k} (k = iteration index given) to the image box.
__global__ __launch_bounds__(NTHR) void k_crop(const JobDev* __restrict__ jobs, int j, int k, float* __restrict__ out)
{
    const JobDev& J = jobs[j];
    const float* U = (k & 1) ? J.U1 : J.U0;
    const size_t n = (size_t)J.M * J.N * J.P;
    for (size_t i = (size_t)blockIdx.x * NTHR + threadIdx.x; i < n; i += (size_t)gridDim.x * NTHR) {
        const int m = (int)(i % J.M);
        const size_t r = i / J.M;
        const int yy = (int)(r % J.N), zz = (int)(r / J.N);
        out[i] = U[m + (size_t)J.nx * (yy + (size_t)J.ny * zz)];
    }
}

// ---------------------------------------------------------------------------
// launch helpers (host)
// ---------------------------------------------------------------------------
// Dynamic LDS above 64 KiB must be allowed per kernel (gfx950 has 160 KiB
// per workgroup).  The limit asked for leaves 1 KiB for static LDS.
static int g_lds_limit = 0;

template <typename K>
static void allow_big_lds(K kernel)
{
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, g_lds_limit);
    if (e != hipSuccess) {
        fprintf(stderr, "dwhip: raising the LDS limit to %d bytes failed: %s\n", g_lds_limit, hipGetErrorString(e));
        (void)hipGetLastError();
    }
}

static void allow_big_lds_all()
{
    static bool done = false;
    if (done) return;
    done = true;
    int v = 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess) v = 65536;
    g_lds_limit = v - 1024;
    allow_big_lds(k_xpass<XM_PSF>);
    allow_big_lds(k_xpass<XM_ONES>);
    allow_big_lds(k_xpass<XM_WEIGHTS>);
    allow_big_lds(k_xpass<XM_UPDATE>);
    allow_big_lds(k_xpass<XM_RATIO>);
    allow_big_lds(k_xpass<XM_LOAD>);
    allow_big_lds(k_xpass<XM_STORE>);
    allow_big_lds(k_colpass<CM_YFWD>);
    allow_big_lds(k_colpass<CM_YINV>);
    allow_big_lds(k_colpass<CM_ZOTF>);
    allow_big_lds(k_colpass<CM_ZCONV>);
    allow_big_lds(k_colpass<CM_ZCONVJ>);
    allow_big_lds(k_colpass<CM_ZFWD>);
    allow_big_lds(k_colpass<CM_ZINV>);
    (void)hipGetLastError();
}

extern "C" {

hipError_t dwhip_launch_xpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s)
{
    allow_big_lds_all();
    switch (mode) {
    case XM_PSF: hipLaunchKernelGGL(k_xpass<XM_PSF>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_ONES: hipLaunchKernelGGL(k_xpass<XM_ONES>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_WEIGHTS: hipLaunchKernelGGL(k_xpass<XM_WEIGHTS>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_UPDATE: hipLaunchKernelGGL(k_xpass<XM_UPDATE>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_RATIO: hipLaunchKernelGGL(k_xpass<XM_RATIO>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_LOAD: hipLaunchKernelGGL(k_xpass<XM_LOAD>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case XM_STORE: hipLaunchKernelGGL(k_xpass<XM_STORE>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t dwhip_launch_colpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s)
{
    allow_big_lds_all();
    switch (mode) {
    case CM_YFWD: hipLaunchKernelGGL(k_colpass<CM_YFWD>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_YINV: hipLaunchKernelGGL(k_colpass<CM_YINV>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_ZOTF: hipLaunchKernelGGL(k_colpass<CM_ZOTF>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_ZCONV: hipLaunchKernelGGL(k_colpass<CM_ZCONV>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_ZCONVJ: hipLaunchKernelGGL(k_colpass<CM_ZCONVJ>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_ZFWD: hipLaunchKernelGGL(k_colpass<CM_ZFWD>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    case CM_ZINV: hipLaunchKernelGGL(k_colpass<CM_ZINV>, dim3(grid), dim3(NTHR), lds, s, bd, jobs); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t dwhip_launch_stats(const JobDev* jobs, int j, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats_partial, dim3(nblk), dim3(NTHR), 0, s, jobs, j, nblk);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(NTHR), 0, s, jobs, j, nblk);
    return hipGetLastError();
}

hipError_t dwhip_launch_start(const JobDev* jobs, int j, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_start, dim3(grid), dim3(NTHR), 0, s, jobs, j);
    return hipGetLastError();
}

hipError_t dwhip_launch_round_end(const BatchDev* bd, const JobDev* jobs, int njobs, int advance, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_end, dim3(njobs), dim3(NTHR), 0, s, bd, jobs, advance);
    return hipGetLastError();
}

hipError_t dwhip_launch_crop(const JobDev* jobs, int j, int k, float* out, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_crop, dim3(grid), dim3(NTHR), 0, s, jobs, j, k, out);
    return hipGetLastError();
}

}  // extern "C"
