// Kernels of the any-size path (generic_dev.h): per-axis complex FFTs of
// any length in LDS (Stockham radix 2/3/4/5/7, Bluestein for other
// factors; x as a real transform, two rows per complex pencil) and the
// elementwise steps of deconwolf's SHB and RL iterations.  The arithmetic
// of the elementwise steps is the same, operation for operation, as the
// fused passes' (shb_passes.h), which restate src/method_shb.c:22-70,
// 333-376 and src/method_rl.c:45-97,357-366.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "generic_dev.h"

#define GTHR 256  // threads per workgroup, elementwise kernels
// The axis kernels are compiled for up to 1024 threads per workgroup (120
// VGPRs, four waves per SIMD) and launched with as many as their tile
// needs (256, 512 or 1024: the host's choice per axis); their loops stride
// by the launch's blockDim.x.
#define GATHR_MAX 1024
#define GATHR ((int)blockDim.x)

namespace {

__device__ __forceinline__ float2 gmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 gmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 gconj(float2 a) { return make_float2(a.x, -a.y); }

// alpha of method_shb.c:328-330 and the extrapolation of :333-339, as in
// the fused update pass
__device__ __forceinline__ double g_alpha(int iter, float alphamax)
{
    double a = ((float)iter - 1.0) / ((float)iter + 2.0);
    if (a < 0) a = 0;
    if (a > alphamax) a = alphamax;
    return a;
}
__device__ __forceinline__ float g_extrap(float x, float xp, double alpha, float bg)
{
    float p = (float)((double)x + alpha * (double)(x - xp));
    return p < bg ? bg : p;
}

template <int R>
__device__ __forceinline__ void gstage(const float2* __restrict__ src, float2* __restrict__ dst, int m, int nb,
                                       int ns, const float2* __restrict__ tw, int sign)
{
    const int L = m / R;
    // indices by magic multiply (fft_engine.h's fdiv; j < 2^24, so exact)
    const uint32_t Lm = fdiv_magic((uint32_t)L), nsm = fdiv_magic((uint32_t)ns);
    for (int j = threadIdx.x; j < nb * L; j += GATHR) {
        const int b = (int)fdiv((uint32_t)j, Lm), jj = j - b * L;
        const int k = jj - (int)fdiv((uint32_t)jj, nsm) * ns;
        float2 v[R];
#pragma unroll
        for (int t = 0; t < R; ++t) v[t] = src[b * m + jj + t * L];
        if (ns > 1) {
#pragma unroll
            for (int t = 1; t < R; ++t) {
                const float2 w = tw[k * (R - 1) + t - 1];
                v[t] = sign < 0 ? gmul(v[t], w) : gmulc(v[t], w);
            }
        }
        // the fused passes' small DFTs (fft_engine.h)
        if (sign < 0)
            DftC<R, -1>::run(v);
        else
            DftC<R, +1>::run(v);
        const int o = b * m + (jj - k) * R + k;
#pragma unroll
        for (int t = 0; t < R; ++t) dst[o + t * ns] = v[t];
    }
}

// Unnormalised length-m FFT of nb pencils (pencil b at a + b*m), ping-pong
// between a and t; returns the buffer holding the result.  stw: the stage
// twiddles (stage s at stw + twoff[s] - twoff[0]), in LDS or memory.
__device__ __forceinline__ float2* gfft(float2* a, float2* t, const GenPlanDev& pl, const float2* __restrict__ stw,
                                        int nb, int sign)
{
    float2 *src = a, *dst = t;
    for (int s = 0; s < pl.nst; ++s) {
        const float2* tws = stw + (pl.twoff[s] - pl.twoff[0]);
        switch (pl.radix[s]) {
        case 2: gstage<2>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
        case 3: gstage<3>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
        case 4: gstage<4>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
        case 5: gstage<5>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
#if DWHIP_GEN_R8
        case 8: gstage<8>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
#endif
        default: gstage<7>(src, dst, pl.m, nb, pl.ns[s], tws, sign); break;
        }
        __syncthreads();
        float2* x = src;
        src = dst;
        dst = x;
    }
    return src;
}

// Column geometry of tile t along y (AXIS 1) or z (AXIS 2): B adjacent
// bins kx < hc of one z plane (y) or one y row (z).  Base offset of pencil
// 0 in Z and K, element stride, pencils in the tile.
template <int AXIS>
__device__ __forceinline__ void gpencils(const GenDev& G, int B, int t, size_t* base, size_t* estep, int* nb)
{
    const int hc = G.hc, hs = G.hs, ny = G.ny;
    const int nxb = (hc + B - 1) / B;
    const int xb = t % nxb, o = t / nxb;  // o: z for AXIS 1, y for AXIS 2
    const int x0 = xb * B;
    *base = AXIS == 1 ? (size_t)x0 + (size_t)hs * ny * o : (size_t)x0 + (size_t)hs * o;
    *estep = AXIS == 1 ? (size_t)hs : (size_t)hs * ny;
    *nb = min(B, hc - x0);
}

// Tiles of an axis pass: B row pairs along x (the pair of rows 2p, 2p+1 is
// one complex pencil), B bins of a plane or row along y and z.
template <int AXIS>
__device__ __forceinline__ int gntiles(const GenDev& G, int B)
{
    if (AXIS == 0) return ((G.ny * G.nz + 1) / 2 + B - 1) / B;
    const int nxb = (G.hc + B - 1) / B;
    return nxb * (AXIS == 1 ? G.nz : G.ny);
}

// Length-n transform of the nb pencils in `a` (the pre-chirp applied when
// Bluestein), `t` scratch; returns the buffer holding the result (before
// the Bluestein post-chirp).
__device__ __forceinline__ float2* gtransform(float2* a, float2* t, const GenPlanDev& pl,
                                              const float2* __restrict__ tw, const float2* __restrict__ stw, int nb,
                                              int sign)
{
    float2* r = gfft(a, t, pl, stw, nb, pl.bluestein ? -1 : sign);
    if (pl.bluestein) {
        const float2* H = tw + (sign < 0 ? pl.hf : pl.hi);
        const uint32_t mm = fdiv_magic((uint32_t)pl.m);
        for (int idx = threadIdx.x; idx < nb * pl.m; idx += GATHR) {
            const int k = idx - (int)fdiv((uint32_t)idx, mm) * pl.m;
            r[idx] = gmul(r[idx], H[k]);
        }
        __syncthreads();
        r = gfft(r, r == a ? t : a, pl, stw, nb, +1);
    }
    return r;
}

// In-place FFT of every pencil of Z along AXIS; sign -1 forward, +1
// inverse (unnormalised).  x (AXIS 0) is the real transform: forward, the
// real rows 2p and 2p+1 (row stride 2*hs floats) are the real and imaginary
// parts of pencil p, whose spectrum Y gives the half spectra
// (Y[k] + conj(Y[n-k])) / 2 and (Y[k] - conj(Y[n-k])) / 2i of the two rows,
// written over them; inverse, the two half spectra are extended by
// Hermitian symmetry (bins 0 and n/2 by their real parts, as a C2R
// transform takes them) into one pencil, whose inverse holds the two real
// rows.  y and z are complex transforms of the bins kx < hc.  JUNC > 0 (the
// z axis): forward transform, times K (JUNC 1, fft_mul_inplace,
// fft.c:437-453) or conj(K) (JUNC 2, fft_mul_conj_inplace, fft.c:490-512),
// inverse transform, in LDS: the junction of a convolution in one pass
// instead of three.
template <int AXIS, int JUNC, bool TWL>
__device__ __forceinline__ void gaxis_body(const GenDev& G, int sign)
{
    extern __shared__ float2 sh[];
    const GenPlanDev& pl = G.plan[AXIS];
    const int n = pl.n, m = pl.m, B = pl.B;
    float2* a = sh;
    float2* t = sh + (size_t)B * m;
    const float2* __restrict__ tw = G.tw;
    // the stage twiddles: copied to LDS after the two pencil buffers (TWL),
    // or read from memory
    const float2* stw = tw + pl.twoff[0];
    if (TWL) {
        float2* twl = sh + 2 * (size_t)B * m;
        for (int i = threadIdx.x; i < pl.twlen; i += GATHR) twl[i] = tw[pl.twoff[0] + i];
        __syncthreads();
        stw = twl;
    }
    const float2* chirp = tw + pl.chirp;
    const int s_in = JUNC ? -1 : sign, s_out = JUNC ? +1 : sign;
    float2* __restrict__ Z = G.Z;
    float* __restrict__ Zr = reinterpret_cast<float*>(G.Z);
    const float2* __restrict__ K = G.K;
    const int ntile = gntiles<AXIS>(G, B);
    const int nrows = G.ny * G.nz, npairs = (nrows + 1) / 2, hc = G.hc;
    const size_t hs = (size_t)G.hs, rs = 2 * hs;  // complex and real row strides
    // index splits by magic multiply (fft_engine.h's fdiv; indices < 2^24)
    const uint32_t mmag = fdiv_magic((uint32_t)m), nmag = fdiv_magic((uint32_t)n), bmag = fdiv_magic((uint32_t)B);
    const uint32_t hmag = fdiv_magic((uint32_t)hc);
    for (int t0 = blockIdx.x; t0 < ntile; t0 += gridDim.x) {
        // y and z: the tiles that share a row's lines (adjacent bins) on one
        // XCD, consecutive there (workgroup b runs on XCD b mod 8), when the
        // grid covers the tiles once
        int tile = t0;
        if (AXIS != 0 && (int)gridDim.x == ntile) {
            const int q = ntile >> 3, rr = ntile & 7, x = t0 & 7, kk = t0 >> 3;
            tile = x * q + min(x, rr) + kk;
        }
        size_t base = 0, estep = 0;
        int nb;
        if (AXIS == 0)
            nb = min(B, npairs - tile * B);
        else
            gpencils<AXIS>(G, B, tile, &base, &estep, &nb);
        const int p0 = tile * B;  // x: the tile's first row pair
        for (int idx = threadIdx.x; idx < B * m; idx += GATHR) {
            // x rows: element-fastest (consecutive addresses); y/z columns:
            // pencil-fastest (adjacent bins, consecutive addresses)
            const int q = (int)fdiv((uint32_t)idx, AXIS == 0 ? mmag : bmag);
            const int b = AXIS == 0 ? q : idx - q * B;
            const int i = AXIS == 0 ? idx - q * m : q;
            float2 v = make_float2(0.f, 0.f);
            if (b < nb && i < n) {
                if (AXIS == 0) {
                    const size_t r0 = 2 * (size_t)(p0 + b);
                    const bool two = r0 + 1 < (size_t)nrows;
                    if (sign < 0) {
                        v = make_float2(Zr[r0 * rs + i], two ? Zr[(r0 + 1) * rs + i] : 0.f);
                    } else {
                        const int k = i < hc ? i : n - i;
                        const float2 A = Z[r0 * hs + k];
                        const float2 C = two ? Z[(r0 + 1) * hs + k] : make_float2(0.f, 0.f);
                        if (i == 0 || 2 * i == n)
                            v = make_float2(A.x, C.x);
                        else if (i < hc)
                            v = make_float2(A.x - C.y, A.y + C.x);  // A + iC
                        else
                            v = make_float2(A.x + C.y, C.x - A.y);  // conj(A) + i conj(C)
                    }
                } else {
                    v = Z[base + b + i * estep];
                }
                if (pl.bluestein) v = s_in < 0 ? gmul(v, chirp[i]) : gmulc(v, chirp[i]);
            }
            a[b * m + i] = v;
        }
        __syncthreads();
        float2* r = gtransform(a, t, pl, tw, stw, nb, s_in);
        if (JUNC) {
            // spectrum (post-chirp) times K, then the inverse's pre-chirp,
            // into the other buffer (zero beyond n, as the load step pads)
            float2* o = r == a ? t : a;
            for (int idx = threadIdx.x; idx < B * m; idx += GATHR) {
                const int q = (int)fdiv((uint32_t)idx, bmag);
                const int b = idx - q * B;
                const int i = q;
                float2 v = make_float2(0.f, 0.f);
                if (b < nb && i < n) {
                    v = r[b * m + i];
                    if (pl.bluestein) v = gmul(v, chirp[i]);
                    const float2 k = K[base + b + i * estep];
                    v = JUNC == 2 ? gmulc(v, k) : gmul(v, k);
                    if (pl.bluestein) v = gmulc(v, chirp[i]);
                }
                o[b * m + i] = v;
            }
            __syncthreads();
            r = gtransform(o, o == a ? t : a, pl, tw, stw, nb, s_out);
        }
        if (AXIS == 0 && sign < 0) {
            // the two rows' half spectra from the pencil's spectrum Y
            for (int idx = threadIdx.x; idx < B * hc; idx += GATHR) {
                const int b = (int)fdiv((uint32_t)idx, hmag);
                const int k = idx - b * hc;
                if (b >= nb) continue;
                const int kn = k ? n - k : 0;
                float2 y = r[b * m + k], yn = r[b * m + kn];
                if (pl.bluestein) {
                    y = gmul(y, chirp[k]);
                    yn = gmul(yn, chirp[kn]);
                }
                const float2 A = make_float2((y.x + yn.x) * 0.5f, (y.y - yn.y) * 0.5f);  // (Y[k] + conj(Y[n-k])) / 2
                const float2 C = make_float2((y.y + yn.y) * 0.5f, (yn.x - y.x) * 0.5f);  // (Y[k] - conj(Y[n-k])) / 2i
                const size_t r0 = 2 * (size_t)(p0 + b);
                Z[r0 * hs + k] = A;
                if (r0 + 1 < (size_t)nrows) Z[(r0 + 1) * hs + k] = C;
            }
        } else if (AXIS == 0) {
            // the two real rows from the pencil's inverse
            for (int idx = threadIdx.x; idx < B * n; idx += GATHR) {
                const int b = (int)fdiv((uint32_t)idx, nmag);
                const int i = idx - b * n;
                if (b >= nb) continue;
                float2 v = r[b * m + i];
                if (pl.bluestein) v = gmulc(v, chirp[i]);
                const size_t r0 = 2 * (size_t)(p0 + b);
                Zr[r0 * rs + i] = v.x;
                if (r0 + 1 < (size_t)nrows) Zr[(r0 + 1) * rs + i] = v.y;
            }
        } else {
            for (int idx = threadIdx.x; idx < B * n; idx += GATHR) {
                const int q = (int)fdiv((uint32_t)idx, bmag);
                const int b = idx - q * B;
                const int i = q;
                if (b >= nb) continue;
                float2 v = r[b * m + i];
                if (pl.bluestein) v = s_out < 0 ? gmul(v, chirp[i]) : gmulc(v, chirp[i]);
                Z[base + b + i * estep] = v;
            }
        }
        __syncthreads();
    }
}

template <int AXIS, int JUNC>
__global__ __launch_bounds__(GATHR_MAX) void gk_axis(const GenDev* __restrict__ gd, int sign)
{
    if (gd->plan[AXIS].twlds)
        gaxis_body<AXIS, JUNC, true>(*gd, sign);
    else
        gaxis_body<AXIS, JUNC, false>(*gd, sign);
}

// The elementwise kernels walk the volume by x-rows (row r = z*ny + y):
// workgroups over rows, threads along the row.  Real volumes U0, U1 and W
// have rows of nx floats; Z's real rows are 2*hs floats apart.
#define GROWS(G)                                                                       \
    for (size_t r = blockIdx.x, nrows_ = (size_t)(G).ny * (G).nz; r < nrows_; r += gridDim.x) \
        for (int x = threadIdx.x; x < (G).nx; x += GTHR)

__device__ __forceinline__ float* gzr(const GenDev& G, size_t r) { return reinterpret_cast<float*>(G.Z) + r * 2 * (size_t)G.hs; }

// OTF set-up (method_shb.c:184-212): the PSF inserted at the corner and
// shifted so its maximum (st->mid) is at the origin.
__global__ __launch_bounds__(GTHR) void gk_psf(const GenDev* __restrict__ gd)
{
    const GenDev& G = *gd;
    const int* mid = G.st->mid;
    GROWS(G)
    {
        const int y = (int)(r % G.ny), z = (int)(r / G.ny);
        const int sx = (x + mid[0]) % G.nx, sy = (y + mid[1]) % G.ny, sz = (z + mid[2]) % G.nz;
        float v = 0.f;
        if (sx < G.pM && sy < G.pN && sz < G.pP) v = G.psf[sx + (size_t)G.pM * (sy + (size_t)G.pN * sz)];
        gzr(G, r)[x] = v;
    }
}

// K = FFT(psf) / (float)nx*ny*nz (the divide of fft.c:856 folded into the
// OTF), over the half spectrum
__global__ __launch_bounds__(GTHR) void gk_otf(const GenDev* __restrict__ gd)
{
    const GenDev& G = *gd;
    const float scale = ((float)G.nx * (float)G.ny) * (float)G.nz;
    const size_t V = (size_t)G.hs * G.ny * G.nz;
    for (size_t i = (size_t)blockIdx.x * GTHR + threadIdx.x; i < V; i += (size_t)gridDim.x * GTHR) {
        const float2 v = G.Z[i];
        G.K[i] = make_float2(v.x / scale, v.y / scale);
    }
}

// the image box as ones (Bertero weights, method_shb.c:216-253)
__global__ __launch_bounds__(GTHR) void gk_ones(const GenDev* __restrict__ gd)
{
    const GenDev& G = *gd;
    GROWS(G)
    {
        const int y = (int)(r % G.ny), z = (int)(r / G.ny);
        gzr(G, r)[x] = x < G.M && y < G.N && z < G.P ? 1.f : 0.f;
    }
}

// W = P1 > 0.01 ? 1/P1 : 0 (method_shb.c:225-234)
__global__ __launch_bounds__(GTHR) void gk_weights(const GenDev* __restrict__ gd)
{
    const GenDev& G = *gd;
    GROWS(G)
    {
        const float p1 = gzr(G, r)[x];
        G.W[r * G.nx + x] = p1 > 0.01f ? (float)(1.0 / p1) : 0.f;
    }
}

// Z *= K (fft_mul_inplace, fft.c:437-453) or Z *= conj(K) (fft_mul_conj_inplace, fft.c:490-512)
__global__ __launch_bounds__(GTHR) void gk_mulk(const GenDev* __restrict__ gd, int conj)
{
    const GenDev& G = *gd;
    const size_t V = (size_t)G.hs * G.ny * G.nz;
    for (size_t i = (size_t)blockIdx.x * GTHR + threadIdx.x; i < V; i += (size_t)gridDim.x * GTHR)
        G.Z[i] = conj ? gmulc(G.Z[i], G.K[i]) : gmul(G.Z[i], G.K[i]);
}

// Update of iteration k (as the fused pass A): SHB finishes iteration k-1,
// x_k = x~ (p_{k-1} W) clamped to bg with positivity, then extrapolates
// p_k = max(x_k + a_k (x_k - x_{k-1}), bg) into Z; RL sets x_k = x~ (x_{k-1}
// W), clamped to bg when bg > 0, into Z.  Iteration 0 takes x_0 (U0 = U1 =
// start guess).
__global__ __launch_bounds__(GTHR) void gk_update(const GenDev* __restrict__ gd)
{
    const GenDev& G = *gd;
    const int iter = G.st->iter;
    const float bg = G.st->bg;
    float* __restrict__ Uk = (iter & 1) ? G.U1 : G.U0;
    const float* __restrict__ Ukm1 = (iter & 1) ? G.U0 : G.U1;
    const double am1 = g_alpha(iter - 1, G.alphamax), al = g_alpha(iter, G.alphamax);
    GROWS(G)
    {
        const size_t i = r * G.nx + x;
        float* zp = gzr(G, r) + x;
        const float xkm1 = Ukm1[i];
        float xk, zv;
        if (G.rl) {
            xk = xkm1;
            if (iter > 0) {
                const float xt = *zp;
                xk = G.has_w ? xt * (xkm1 * G.W[i]) : xt * xkm1;
                if (bg > 0) xk = xk < bg ? bg : xk;
                Uk[i] = xk;
            }
            zv = xk;
        } else {
            if (iter == 0) {
                xk = Uk[i];
            } else {
                const float xt = *zp;
                const float p = g_extrap(xkm1, Uk[i], am1, bg);
                xk = G.has_w ? xt * (p * G.W[i]) : xt * p;
                if (G.positivity) xk = xk < bg ? bg : xk;
                Uk[i] = xk;
            }
            zv = g_extrap(xk, xkm1, al, bg);
        }
        *zp = zv;
    }
}

// Error terms of y (Z's real rows) against the image (getError,
// dw.c:1436-1524) and the ratio into Z: SHB im/y with the copysign rule
// inside the image, 0 outside (method_shb.c:27-48); RL im/y, bg where
// y <= 0, 1e-6 outside (method_rl.c:45-65).  One partial sum per workgroup
// (fixed grid).
__global__ __launch_bounds__(GTHR) void gk_ratio(const GenDev* __restrict__ gd, int metric_idiv)
{
    __shared__ double red[GTHR / 64];
    const GenDev& G = *gd;
    const float bg = G.st->bg;
    const float outside = G.rl ? 1e-6f : 0.f;
    double acc = 0;
    GROWS(G)
    {
        const int yy = (int)(r % G.ny), zz = (int)(r / G.ny);
        float* zp = gzr(G, r) + x;
        float rv = outside;
        if (x < G.M && yy < G.N && zz < G.P) {
            const float y = *zp;
            const float gv = G.im[x + (size_t)G.M * (yy + (size_t)G.N * zz)];
            const double yd = y, gdd = gv;
            const float q = gv / y;
            if (metric_idiv) {
                if (yd > 0 && gdd > 0) acc += gdd * (double)logf(q) - (gdd - yd);
            } else {
                acc += (yd - gdd) * (yd - gdd);
            }
            if (G.rl)
                rv = y > 0 ? q : bg;
            else if (fabs(yd) < 1e-6)
                rv = gv / (float)copysign(1e-6, yd);
            else
                rv = q;
        }
        *zp = rv;
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int w = 0; w < GTHR / 64; ++w) s += red[w];
        G.part[blockIdx.x] = s;
    }
}

}  // namespace

extern "C" {

hipError_t dwhip_gen_launch_axis(const GenDev* gd, int axis, int sign, int grid, size_t lds, int threads,
                                 hipStream_t s)
{
    if (threads != 256 && threads != 512 && threads != GATHR_MAX) return hipErrorInvalidValue;
    // axis 3 / 4: the z axis with the K / conj(K) junction
    switch (axis) {
    case 0: hipLaunchKernelGGL((gk_axis<0, 0>), dim3(grid), dim3(threads), lds, s, gd, sign); break;
    case 1: hipLaunchKernelGGL((gk_axis<1, 0>), dim3(grid), dim3(threads), lds, s, gd, sign); break;
    case 2: hipLaunchKernelGGL((gk_axis<2, 0>), dim3(grid), dim3(threads), lds, s, gd, sign); break;
    case 3: hipLaunchKernelGGL((gk_axis<2, 1>), dim3(grid), dim3(threads), lds, s, gd, sign); break;
    case 4: hipLaunchKernelGGL((gk_axis<2, 2>), dim3(grid), dim3(threads), lds, s, gd, sign); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Raises the dynamic-LDS limit of the axis kernels (once per process).
hipError_t dwhip_gen_lds_limit(size_t bytes)
{
    const void* fs[5] = {(const void*)gk_axis<0, 0>, (const void*)gk_axis<1, 0>, (const void*)gk_axis<2, 0>,
                         (const void*)gk_axis<2, 1>, (const void*)gk_axis<2, 2>};
    hipError_t e = hipSuccess;
    for (int i = 0; i < 5 && e == hipSuccess; ++i)
        e = hipFuncSetAttribute(fs[i], hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    return e;
}

// elementwise steps: 0 psf, 1 otf, 2 ones, 3 weights, 4 mul K, 5 mul conj K,
// 6 update, 7 ratio (IDIV), 8 ratio (MSE)
hipError_t dwhip_gen_launch_elem(const GenDev* gd, int op, int grid, hipStream_t s)
{
    switch (op) {
    case 0: hipLaunchKernelGGL(gk_psf, dim3(grid), dim3(GTHR), 0, s, gd); break;
    case 1: hipLaunchKernelGGL(gk_otf, dim3(grid), dim3(GTHR), 0, s, gd); break;
    case 2: hipLaunchKernelGGL(gk_ones, dim3(grid), dim3(GTHR), 0, s, gd); break;
    case 3: hipLaunchKernelGGL(gk_weights, dim3(grid), dim3(GTHR), 0, s, gd); break;
    case 4: hipLaunchKernelGGL(gk_mulk, dim3(grid), dim3(GTHR), 0, s, gd, 0); break;
    case 5: hipLaunchKernelGGL(gk_mulk, dim3(grid), dim3(GTHR), 0, s, gd, 1); break;
    case 6: hipLaunchKernelGGL(gk_update, dim3(grid), dim3(GTHR), 0, s, gd); break;
    case 7: hipLaunchKernelGGL(gk_ratio, dim3(grid), dim3(GTHR), 0, s, gd, 1); break;
    case 8: hipLaunchKernelGGL(gk_ratio, dim3(grid), dim3(GTHR), 0, s, gd, 0); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // extern "C"
