// Set-up, reduction and crop kernels of the SHB loop, and the dispatch of
// the per-mode pass launchers.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include <algorithm>

#include "shb_dev.h"
#include "shb_kernels.h"

template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < NT / 64; ++i) s += red[i];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Setup statistics: min(im) and sum(im) (bg_auto, flat start) and argmax(psf)
// ---------------------------------------------------------------------------
#define STHR 256
// stat layout per job: [nblk][4] doubles: min, sum, max, argmax index
__global__ __launch_bounds__(STHR) void k_stats_partial(const JobDev* __restrict__ jobs, int j, int nblk)
{
    const JobDev& J = jobs[j];
    __shared__ double smin[STHR], ssum[STHR], smax[STHR], sidx[STHR];
    const size_t n = (size_t)J.M * J.N * J.P;
    const size_t np = (size_t)J.pM * J.pN * J.pP;
    double vmin = INFINITY, vsum = 0, vmax = -INFINITY, vidx = 1e300;
    // each thread visits i, i + S, i + 2S, ... in that order; eight loads
    // are issued before they are added (the same sum, more loads in flight)
    const size_t S = (size_t)nblk * STHR;
    size_t i = (size_t)blockIdx.x * STHR + threadIdx.x;
    for (; i + 7 * S < n; i += 8 * S) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = J.im[i + q * S];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const double d = v[q];
            vmin = d < vmin ? d : vmin;
            vsum += d;
        }
    }
    for (; i < n; i += S) {
        const double v = J.im[i];
        vmin = v < vmin ? v : vmin;
        vsum += v;
    }
    for (size_t i = (size_t)blockIdx.x * STHR + threadIdx.x; i < np; i += (size_t)nblk * STHR) {
        const double v = J.psf[i];
        if (v > vmax) {  // strictly greater keeps the lowest index per thread
            vmax = v;
            vidx = (double)i;
        }
    }
    const int tid = threadIdx.x;
    smin[tid] = vmin; ssum[tid] = vsum; smax[tid] = vmax; sidx[tid] = vidx;
    __syncthreads();
    for (int s = STHR / 2; s > 0; s >>= 1) {
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

__global__ __launch_bounds__(STHR) void k_stats_final(const JobDev* __restrict__ jobs, int j, int nblk)
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
    st->stop = 0;
}

// Start guess into U0 and U1 (method_shb.c:264-287).
__global__ __launch_bounds__(STHR) void k_start(const JobDev* __restrict__ jobs, int j)
{
    const JobDev& J = jobs[j];
    const size_t n = (size_t)J.nx * J.ny * J.nz;
    const float c0 = J.st->c0;
    const float* src = J.start_img ? J.start_img : J.im;
    for (size_t i = (size_t)blockIdx.x * STHR + threadIdx.x; i < n; i += (size_t)gridDim.x * STHR) {
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
// dw_iterator_next (src/dw.c:237-269) before iteration n, the errors of
// iterations n-1 and n-2 in J.err: whether iteration n runs.
__device__ bool iterator_continue_dev(const JobDev& J, int n)
{
    const float e = J.err[n - 1];
    if (J.iter_type == DWHIP_DEV_ITER_REL) {
        const float last = n < 2 ? 2 * e * J.err_rel : J.err[n - 2];
        // C's fabs: the float difference, then double arithmetic
        if (fabs((double)(e - last)) / (double)e < (double)J.err_rel) return false;
    } else if (J.iter_type == DWHIP_DEV_ITER_ABS) {
        if (n > 0 && e < J.err_abs) return false;
    }
    return n < J.niter;
}

__global__ __launch_bounds__(STHR) void k_round_end(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs,
                                                    int advance)
{
    __shared__ double red[STHR / 64];
    const JobDev& J = jobs[blockIdx.x];
    if (J.dev_stop && J.st->stop) return;  // stopped in an earlier round of this step
    const int nt = J.ntile[KIND_XR];  // one partial per tile of pass E
    // each thread sums its strided partials in index order; the loads of
    // eight of them are issued together
    const double* __restrict__ part = J.part;
    double s = 0;
    int i = threadIdx.x;
    for (; i + 7 * STHR < nt; i += 8 * STHR) {
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) a[q] = part[i + q * STHR];
#pragma unroll
        for (int q = 0; q < 8; ++q) s += a[q];
    }
    for (; i < nt; i += STHR) s += part[i];
    s = block_sum<STHR>(s, red);
    if (threadIdx.x == 0) {
        JobState* st = J.st;
        const int it = st->iter;
        const double mean = s / (double)((long long)J.M * J.N * J.P);
        if (it < J.niter) {
            J.err[it] = (float)mean;
            J.errd[it] = mean;
        }
        if (advance) {
            st->iter = it + 1;
            if (J.dev_stop && it < J.niter && !iterator_continue_dev(J, it + 1)) st->stop = 1;
        }
    }
}

// Crop x_k (k = iteration index given) to the image box.
__global__ __launch_bounds__(STHR) void k_crop(const JobDev* __restrict__ jobs, int j, int k, float* __restrict__ out)
{
    const JobDev& J = jobs[j];
    const float* U = (k & 1) ? J.U1 : J.U0;
    // by image rows: workgroups over the rows, threads along a row
    const int nrows = J.N * J.P;
    for (int r = blockIdx.x; r < nrows; r += gridDim.x) {
        const int zz = r / J.N, yy = r - zz * J.N;
        const float* src = U + (size_t)J.nx * (yy + (size_t)J.ny * zz);
        float* dst = out + (size_t)J.M * r;
        for (int m = threadIdx.x; m < J.M; m += STHR) dst[m] = src[m];
    }
}

// ---------------------------------------------------------------------------
// launch helpers (host)
// ---------------------------------------------------------------------------
#define DECL_LAUNCH(name) \
    extern "C" hipError_t name(const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s);
DECL_LAUNCH(dwhip_launch_x0)
DECL_LAUNCH(dwhip_launch_x1)
DECL_LAUNCH(dwhip_launch_x2)
DECL_LAUNCH(dwhip_launch_x3)
DECL_LAUNCH(dwhip_launch_x4)
DECL_LAUNCH(dwhip_launch_x5)
DECL_LAUNCH(dwhip_launch_x6)
DECL_LAUNCH(dwhip_launch_c0)
DECL_LAUNCH(dwhip_launch_c1)
DECL_LAUNCH(dwhip_launch_c2)
DECL_LAUNCH(dwhip_launch_c3)
DECL_LAUNCH(dwhip_launch_c4)
DECL_LAUNCH(dwhip_launch_c5)
DECL_LAUNCH(dwhip_launch_c6)
DECL_LAUNCH(dwhip_launch_c7)
DECL_LAUNCH(dwhip_launch_c8)

extern "C" {

hipError_t dwhip_launch_xpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s)
{
    static_assert(XM_PSF == 0 && XM_STORE == 6, "X modes 0..6");
    switch (mode) {
    case 0: return dwhip_launch_x0(bd, jobs, grid, lds, s);
    case 1: return dwhip_launch_x1(bd, jobs, grid, lds, s);
    case 2: return dwhip_launch_x2(bd, jobs, grid, lds, s);
    case 3: return dwhip_launch_x3(bd, jobs, grid, lds, s);
    case 4: return dwhip_launch_x4(bd, jobs, grid, lds, s);
    case 5: return dwhip_launch_x5(bd, jobs, grid, lds, s);
    case 6: return dwhip_launch_x6(bd, jobs, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t dwhip_launch_colpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s)
{
    static_assert(CM_YFWD == 0 && CM_YFWD_IMG == 8 && DWHIP_N_CMODES == 9, "column modes 0..8");
    switch (mode) {
    case 0: return dwhip_launch_c0(bd, jobs, grid, lds, s);
    case 1: return dwhip_launch_c1(bd, jobs, grid, lds, s);
    case 2: return dwhip_launch_c2(bd, jobs, grid, lds, s);
    case 3: return dwhip_launch_c3(bd, jobs, grid, lds, s);
    case 4: return dwhip_launch_c4(bd, jobs, grid, lds, s);
    case 5: return dwhip_launch_c5(bd, jobs, grid, lds, s);
    case 6: return dwhip_launch_c6(bd, jobs, grid, lds, s);
    case 7: return dwhip_launch_c7(bd, jobs, grid, lds, s);
    case 8: return dwhip_launch_c8(bd, jobs, grid, lds, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t dwhip_launch_stats(const JobDev* jobs, int j, int nblk, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats_partial, dim3(nblk), dim3(STHR), 0, s, jobs, j, nblk);
    hipLaunchKernelGGL(k_stats_final, dim3(1), dim3(STHR), 0, s, jobs, j, nblk);
    return hipGetLastError();
}

hipError_t dwhip_launch_start(const JobDev* jobs, int j, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_start, dim3(grid), dim3(STHR), 0, s, jobs, j);
    return hipGetLastError();
}

hipError_t dwhip_launch_round_end(const BatchDev* bd, const JobDev* jobs, int njobs, int advance, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_end, dim3(njobs), dim3(STHR), 0, s, bd, jobs, advance);
    return hipGetLastError();
}

hipError_t dwhip_launch_crop(const JobDev* jobs, int j, int k, float* out, int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_crop, dim3(grid), dim3(STHR), 0, s, jobs, j, k, out);
    return hipGetLastError();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// XY tiling (deconwolf's tiling.c / deconvolve_tiles, src/dw.c:2000-2229)
// ---------------------------------------------------------------------------
// Tile i: out[(m - x0) + xm*((n - y0) + xn*(p - z0))] = im[m + M*(n + N*p)] + offset
// over the tile's grown box (tiling_get_tile, tiling.c:300-330, then
// fim_add_scalar as deconvolve_tiles does, dw.c:2131-2134).
__global__ __launch_bounds__(STHR) void k_tile_extract(const float* __restrict__ im, int M, int N, int x0, int y0,
                                                      int z0, int xm, int xn, int xp, float offset,
                                                      float* __restrict__ out)
{
    const size_t n = (size_t)xm * xn * xp;
    for (size_t i = (size_t)blockIdx.x * STHR + threadIdx.x; i < n; i += (size_t)gridDim.x * STHR) {
        const int a = (int)(i % xm);
        const size_t r = i / xm;
        const int b = (int)(r % xn), c = (int)(r / xn);
        float v = im[(size_t)(x0 + a) + (size_t)M * ((y0 + b) + (size_t)N * (z0 + c))];
        if (offset > 0) v += offset;
        out[i] = v;
    }
}

// getWeight1d (tiling.c:169-195), float arithmetic as in the reference
__device__ __forceinline__ float tile_w1d(float a, float b, float c, float d, int xi)
{
#pragma clang fp contract(off)
    const float x = (float)xi;
    if (x < a || x > d) return 0.f;
    if (x >= b && x <= c) return 1.f;
    if (x >= a && x <= b) return (x - a + 1.f) / (b - a + 1.f);
    return 1.f - (x - c) / (d - c + 1.f);
}

__device__ __forceinline__ float tile_w(const TileDev& t, int m, int n)
{
    const float wm = tile_w1d((float)t.xpos[0], (float)t.pos[0], (float)t.pos[1], (float)t.xpos[1], m);
    const float wn = tile_w1d((float)t.xpos[2], (float)t.pos[2], (float)t.pos[3], (float)t.xpos[3], n);
    return fminf(wm, wn);
}

// Blend (tiling_put_tile_raw, tiling.c:353-410, over all tiles in order):
// out = sum_t (w_t / sum_s w_s) * tile_t, accumulated in float in tile
// order from zero, each tile first shifted by -offset and clipped at 0 when
// offset > 0 (dw.c:2141-2145).  One thread per (m, n) column of the image.
__global__ __launch_bounds__(STHR) void k_tile_blend(const TileDev* __restrict__ tiles, int ntiles, int M, int N,
                                                    int P, float offset, float* __restrict__ out)
{
    // products and sums rounded separately, as the reference's
    // buf[pos] += w*S compiled for x86-64 without FMA (no contraction)
#pragma clang fp contract(off)
    const size_t ncol = (size_t)M * N;
    for (size_t col = (size_t)blockIdx.x * STHR + threadIdx.x; col < ncol; col += (size_t)gridDim.x * STHR) {
        const int m = (int)(col % M), n = (int)(col / M);
        float wsum = 0.f;
        for (int t = 0; t < ntiles; ++t) wsum = wsum + tile_w(tiles[t], m, n);
        for (int p = 0; p < P; ++p) out[col + ncol * p] = 0.f;
        for (int t = 0; t < ntiles; ++t) {
            const TileDev& T = tiles[t];
            if (m < T.xpos[0] || m > T.xpos[1] || n < T.xpos[2] || n > T.xpos[3]) continue;
            const float w = tile_w(T, m, n) / wsum;
            const size_t base = (size_t)(m - T.xpos[0]) + (size_t)T.xsize[0] * (n - T.xpos[2]);
            const size_t plane = (size_t)T.xsize[0] * T.xsize[1];
            for (int p = T.xpos[4]; p <= T.xpos[5]; ++p) {
                float s = T.out[base + plane * (p - T.xpos[4])];
                if (offset > 0) {
                    s += -offset;
                    s = s < 0 ? 0.f : s;
                }
                out[col + ncol * p] = out[col + ncol * p] + w * s;
            }
        }
    }
}

extern "C" hipError_t dwhip_launch_tile_extract(const float* im, int M, int N, int x0, int y0, int z0, int xm, int xn,
                                               int xp, float offset, float* out, hipStream_t s)
{
    const size_t n = (size_t)xm * xn * xp;
    const int grid = (int)std::min<size_t>((n + STHR - 1) / STHR, 4096);
    hipLaunchKernelGGL(k_tile_extract, dim3(grid), dim3(STHR), 0, s, im, M, N, x0, y0, z0, xm, xn, xp, offset, out);
    return hipGetLastError();
}

extern "C" hipError_t dwhip_launch_tile_blend(const TileDev* tiles, int ntiles, int M, int N, int P, float offset,
                                             float* out, hipStream_t s)
{
    const size_t n = (size_t)M * N;
    const int grid = (int)std::min<size_t>((n + STHR - 1) / STHR, 4096);
    hipLaunchKernelGGL(k_tile_blend, dim3(grid), dim3(STHR), 0, s, tiles, ntiles, M, N, P, offset, out);
    return hipGetLastError();
}
