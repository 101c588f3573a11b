// Pass kernels of the SHB (scaled heavy ball Richardson-Lucy) loop for
// gfx950, as templates; k_xpass.hip and k_colpass.hip instantiate one mode
// per translation unit.  Restates deconwolf's src/method_shb.c (deconvolve_shb :77-424,
// iter_shb :10-74) as fused passes over a job table (shb_dev.h).
//
// A 3-D R2C/C2R transform of an nx*ny*nz volume is three passes:
//   X: x-rows, real length nx packed as h = nx/2 complex values, complex
//      FFT of length h plus the split/merge step to h+1 bins;
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
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#include "fft_engine.h"
#include "shb_dev.h"
#include "shb_kernels.h"

#define XTHR DWHIP_XTHR
#define CTHR DWHIP_CTHR
#define ZTHR DWHIP_ZTHR

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
// x - xp in float, the rest in double, rounded to float.  With
// DWHIP_EXTRAP_F32 the product and sum are one float FMA with alpha rounded
// to float (one rounding of the result instead of the double's; a build
// tunable to measure the double arithmetic's cost).
#ifndef DWHIP_EXTRAP_F32
#define DWHIP_EXTRAP_F32 0
#endif
__device__ __forceinline__ float shb_extrap(float x, float xp, double alpha, float bg)
{
#if DWHIP_EXTRAP_F32
    float p = fmaf((float)alpha, x - xp, x);
#else
    float p = (float)((double)x + alpha * (double)(x - xp));
#endif
    return p < bg ? bg : p;
}

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

#ifndef XCHUNK
#define XCHUNK 8  // elements per thread whose loads are issued together
#endif

// C2R split source read straight from a tile's half-spectrum rows (row
// stride hcs): element i of row pen is formed from X[i], X[h - i] and the
// split twiddle.  With DWHIP_SPLIT_GROUP > 0 the loads of that many
// elements of a radix-25 butterfly are issued before their splits are
// formed; formed per element (0, the default), the compiler waits for each
// element's loads before issuing the next element's.  Measured on the
// update pass of a 1250-wide transform: 1084 us per element, 1208 us
// grouped (182 VGPRs, two waves per SIMD), 1104 us grouped and held to
// three waves; the waves of the other workgroups hide the per-element
// latency at three waves per SIMD.
// R2C merge of a tile: 2 by output bin (a thread writes bin m of every
// row, one twiddle load), 1 by bin pair (bins k and h - k of every row, two
// twiddle loads; a wave's mirrored stores leave partly written lines), 0
// by element (a twiddle load per element).
#ifndef DWHIP_MERGE_BY_BIN
#define DWHIP_MERGE_BY_BIN 2
#endif

#ifndef DWHIP_SPLIT_GROUP
#define DWHIP_SPLIT_GROUP 0
#endif
// Split twiddles of a first-stage butterfly: element i = j + r*h/R needs
// exp(-2 pi i i/nx) = exp(-2 pi i j/nx) * exp(-i pi r/R), one table load
// and R - 1 products by compile-time constants instead of R table loads
// (build tunables).  Measured on a 1250-wide transform: the ratio pass's
// LDS-staged split 316 -> 313 us; the update pass's split from global
// memory 1032 -> 1149 us (182 VGPRs), so that one stays off.
#ifndef DWHIP_SPLIT_ROT_STAGED
#define DWHIP_SPLIT_ROT_STAGED 1
#endif
#ifndef DWHIP_SPLIT_ROT_ROWS
#define DWHIP_SPLIT_ROT_ROWS 0
#endif
template <int R>
__device__ __forceinline__ float2 split_tw(float2 w0, int r)
{
    return r == 0 ? w0 : cmul(w0, make_float2(Roots<2 * R>::t.c[r], -Roots<2 * R>::t.s[r]));
}

// C2R split source from half-spectrum rows staged in LDS (row stride hc)
struct SplitStaged {
    static constexpr bool kStrided = true;
    const float2* L;
    TwPtr rtw;
    int hc, h;
    __device__ __forceinline__ float2 operator()(int pen, int i) const
    {
        const float2* Lr = L + pen * hc;
        return c2r_split(Lr[i], Lr[h - i], ldg(rtw + i));
    }
    template <int R>
    __device__ __forceinline__ void load(int pen, int j, int nb, float2* v) const
    {
        const float2* Lr = L + pen * hc;
        const float2 w0 = ldg(rtw + j);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = j + r * nb;
            v[r] = c2r_split(Lr[i], Lr[h - i], split_tw<R>(w0, r));
        }
    }
};

template <bool NT>
struct SplitRows {
    static constexpr bool kStrided = true;
    const DWHIP_GL float2* C;  // the tile's first row
    TwPtr rtw;
    int hcs, h;
    __device__ __forceinline__ float2 operator()(int pen, int i) const  // one element (single-stage plans)
    {
        const DWHIP_GL float2* crow = C + (size_t)pen * hcs;
        return c2r_split(ldgs<NT>(crow + i), ldgs<NT>(crow + h - i), ldg(rtw + i));
    }
    template <int R>
    __device__ __forceinline__ void load(int pen, int j, int nb, float2* v) const
    {
        if constexpr (DWHIP_SPLIT_ROT_ROWS) {
            const DWHIP_GL float2* crow = C + (size_t)pen * hcs;
            const float2 w0 = ldg(rtw + j);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = j + r * nb;
                v[r] = c2r_split(ldgs<NT>(crow + i), ldgs<NT>(crow + h - i), split_tw<R>(w0, r));
            }
            return;
        }
        if constexpr (R != 25 || DWHIP_SPLIT_GROUP == 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = (*this)(pen, j + r * nb);
            return;
        }
        const DWHIP_GL float2* crow = C + (size_t)pen * hcs;
        constexpr int G = DWHIP_SPLIT_GROUP > 0 && DWHIP_SPLIT_GROUP < R ? DWHIP_SPLIT_GROUP : R;
#pragma unroll
        for (int r0 = 0; r0 < R; r0 += G) {
            float2 a[G], b[G], w[G];
#pragma unroll
            for (int r = 0; r < G; ++r) {
                if (r0 + r >= R) break;
                const int i = j + (r0 + r) * nb;
                a[r] = ldgs<NT>(crow + i);
                b[r] = ldgs<NT>(crow + h - i);
                w[r] = ldg(rtw + i);
            }
#pragma unroll
            for (int r = 0; r < G; ++r) {
                if (r0 + r >= R) break;
                v[r0 + r] = c2r_split(a[r], b[r], w[r]);
            }
            // the next group's loads stay behind this group's splits (its
            // registers would otherwise be live across them)
            __builtin_amdgcn_sched_barrier(0);
        }
    }
};

template <bool B>
struct BoolC {
    static constexpr bool value = B;
};

// Passes that stage their half-spectrum rows in LDS before the C2R (faster
// for the ratio pass, whose tiles are otherwise bound by the latency of the
// split's scattered loads; slower for the update pass: DWHIP_X_STAGE_UPDATE
// measured 1130 us against 1084 on a 1250-wide transform, 1157 us with the
// nontemporal hint on its streams).
#ifndef DWHIP_X_STAGE_UPDATE
#define DWHIP_X_STAGE_UPDATE 0
#endif
__host__ __device__ constexpr bool x_stage_rows(int mode)
{
    return DWHIP_X_STAGE_LDS && (mode == XM_RATIO || (DWHIP_X_STAGE_UPDATE && mode == XM_UPDATE));
}

// Nontemporal hints on the passes' read-once/write-once streams.  Measured
// on c2: the column passes and the ratio pass gain 2-10 %.  The update pass
// keeps plain accesses: its C2R split reads each spectrum line twice
// (elements k and h - k, in different loads), and with the hint the line is
// gone before the second read (x_update 8 % slower); the hint on its stores
// alone changes nothing.
#ifndef DWHIP_NT_X
#define DWHIP_NT_X 1
#endif
#ifndef DWHIP_NT_UPDATE
#define DWHIP_NT_UPDATE 0
#endif
#ifndef DWHIP_NT_COLS
#define DWHIP_NT_COLS 1
#endif
__host__ __device__ constexpr bool x_nt(int mode) { return mode == XM_UPDATE ? DWHIP_NT_UPDATE : DWHIP_NT_X; }
// the update's state streams (x_{k-1}, x_{k-2} in and x_k out, W), read or
// written once, hinted apart from its spectrum reads: they then leave more
// of L2 to the spectrum lines its split reads twice and to the spectrum it
// writes, which the next pass (y) reads
#ifndef DWHIP_NT_STATE
#define DWHIP_NT_STATE 1
#endif

// Element loops of the two iteration passes specialised for their common
// case (SHB iteration k > 0 of the update pass; the SHB ratio pass, whose
// tiles hold image rows only): contiguous tile addressing, the job's flags
// read once, selects instead of branches.  0 runs the generic loop for
// every case (a build tunable for measurement).
#ifndef DWHIP_X_FAST
#define DWHIP_X_FAST 1
#endif

#ifdef DWHIP_X_WAVES
#define DWHIP_X_OCC __attribute__((amdgpu_waves_per_eu(DWHIP_X_WAVES)))
#else
#define DWHIP_X_OCC
#endif

// Plans of the x-row transform compiled with their lengths known (h, the
// stage geometry and the butterfly indexing fold to constants): two-stage
// plans (R1, R2) of the forward transform, the inverse running R2, R1.
// 25 x 25 is the x-row plan of 1250-wide transforms (1024-px images with
// 181-px PSFs).  Off by default (a build tunable): with the geometry known
// the compiler issues a whole butterfly's loads at once, and the passes
// need 218-238 VGPRs (two waves per SIMD; measured 6 % slower on the update
// pass, 2 % slower still when held to three waves with spills).
#ifndef DWHIP_X_FIXED
#define DWHIP_X_FIXED 0
#endif
__host__ __device__ constexpr bool x_fixed_mode(int mode)
{
    return DWHIP_X_FIXED && (mode == XM_UPDATE || mode == XM_RATIO);
}

// x-row transform of a tile with plan p, or with the compile-time plan
// (FR1, FR2) when FR1 > 0 (p then gives only the stage twiddle offsets)
template <int FR1, int FR2, int DIR, bool SRC_LDS, bool DST_LDS, class Src, class Dst>
__device__ __forceinline__ void x_fft(const FftPlanDev& p, const StageGeo& g, int tid, float2* lds, TwPtr __restrict__ tw,
                                      Src src, Dst dst)
{
    if constexpr (FR1 > 0) {
        constexpr int RA = DIR < 0 ? FR1 : FR2, RB = DIR < 0 ? FR2 : FR1;
        stage_first<RA, mb_for(DWHIP_XMAXE, RA), DIR, false, SRC_LDS>(g, tid, src, lds);
        __syncthreads();
        stage_last<RB, mb_for(DWHIP_XMAXE, RB), DIR, false, DST_LDS>(g, tid, lds, tw + p.twoff[1], dst);
    } else {
        fft_tile<DWHIP_XMAXE, DIR, false, DWHIP_RMAX_X, SRC_LDS, DST_LDS>(p, g, tid, lds, tw, src, dst);
    }
}

// One X tile t of job J; FR1 > 0: the job's x plan is (FR1, FR2).
template <int MODE, int FR1, int FR2>
__device__ __forceinline__ void xpass_tile(const JobDev& J, int t, int tid, float2* lds, double* red)
{
    constexpr bool XNT = x_nt(MODE);
    constexpr int kind = xmode_kind(MODE);
    {
        const int lt = t - J.tile0[kind];
        const int h = FR1 > 0 ? FR1 * FR2 : J.h;
        const int nx = FR1 > 0 ? 2 * h : J.nx;
        const int ny = J.ny, hcs = J.hcs;
        int row0 = lt * J.xrows;
        int nrows = min(J.xrows, ny * J.nz - row0);
        if (MODE == XM_RATIO && !J.rl) {
            // the SHB ratio is 0 outside the image box and only y < N, z < P
            // rows of y are used: pass E's tiles cover exactly those rows,
            // ceil(N / xrows) tiles within each plane z < P (the RL ratio is
            // 1e-6 outside the image, so RL runs the plain x tiles)
            const int tpp = (J.N + J.xrows - 1) / J.xrows;
            const int z = lt / tpp, k = lt - z * tpp;
            row0 = z * ny + k * J.xrows;
            nrows = min(J.xrows, J.N - k * J.xrows);
        }
        const DWHIP_GL JobState* __restrict__ st = gp(J.st);
        const int iter = (MODE == XM_UPDATE || MODE == XM_RATIO) ? st->iter : 0;
        const bool inv = MODE == XM_RATIO || MODE == XM_WEIGHTS || MODE == XM_STORE ||
                         (MODE == XM_UPDATE && iter > 0);
        const bool fwd = MODE != XM_WEIGHTS && MODE != XM_STORE;
        StageGeo g;
        g.n = h;
        g.npen = nrows;
        g.pstride = h;
        g.nthr = XTHR;
        g.npshift = 0;
        const uint32_t hmagic = fdiv_magic((uint32_t)h);
        const DWHIP_GL float2* __restrict__ C = gp(J.C);
        TwPtr __restrict__ rtw = gp(J.rtw);

        auto lds_src = [&](int pen, int i) { return lds[pen * h + i]; };
        auto lds_dst = [&](int pen, int i, float2 v) { lds[pen * h + i] = v; };

        if (inv && x_stage_rows(MODE)) {
            // The tile's half-spectrum rows (hc values each) go to LDS first,
            // with coalesced loads of XCHUNK values per thread in flight; the
            // split then reads both of its operands (k and h - k) from there.
            // Rows the ratio pass skips are zeros.
            const int hc = h + 1;
            const uint32_t hcmagic = fdiv_magic((uint32_t)hc);
            uint64_t live = ~0ull;  // bit b: row row0 + b is transformed
            if (MODE == XM_RATIO) {
                live = 0;
                int z = row0 / ny, y = row0 - (row0 / ny) * ny;
                for (int b = 0; b < nrows; ++b) {
                    if (z < J.P && y < J.N) live |= 1ull << b;
                    if (++y == ny) { y = 0; ++z; }
                }
            }
            const int nraw = nrows * hc;
            for (int w0 = tid; w0 < nraw; w0 += XCHUNK * XTHR) {
                float2 v[XCHUNK];
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q) {
                    const int w = w0 + q * XTHR;
                    v[q] = make_float2(0.f, 0.f);
                    if (w < nraw) {
                        const int b = (int)fdiv((uint32_t)w, hcmagic);
                        if ((live >> b) & 1ull) v[q] = ldgs<XNT>(C + (size_t)(row0 + b) * hcs + (w - b * hc));
                    }
                }
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q)
                    if (w0 + q * XTHR < nraw) lds[w0 + q * XTHR] = v[q];
            }
            __syncthreads();
            if (DWHIP_SPLIT_ROT_STAGED) {
                SplitStaged split_src{lds, rtw, hc, h};
                x_fft<FR1, FR2, +1, true, true>(J.pxi, g, tid, lds, gp(J.tw), split_src, lds_dst);
            } else {
                auto split_src = [&](int pen, int i) {
                    const float2* L = lds + pen * hc;
                    return c2r_split(L[i], L[h - i], ldg(rtw + i));
                };
                x_fft<FR1, FR2, +1, true, true>(J.pxi, g, tid, lds, gp(J.tw), split_src, lds_dst);
            }
            __syncthreads();
        } else if (inv && (DWHIP_SPLIT_GROUP > 0 || DWHIP_SPLIT_ROT_ROWS) && MODE != XM_RATIO) {
            SplitRows<false> split_src{C + (size_t)row0 * hcs, rtw, hcs, h};
            x_fft<FR1, FR2, +1, false, true>(J.pxi, g, tid, lds, gp(J.tw), split_src, lds_dst);
            __syncthreads();
        } else if (inv) {
            auto split_src = [&](int pen, int i) {
                const int row = row0 + pen;
                if (MODE == XM_RATIO) {
                    const int zz = row / ny;
                    if (zz >= J.P || row - zz * ny >= J.N) return make_float2(0.f, 0.f);
                }
                // plain loads: each spectrum line is read twice (elements k
                // and h - k), and a nontemporal hint on the first read drops
                // the line before the second (the weights pass fetched its
                // spectrum twice)
                const DWHIP_GL float2* crow = C + (size_t)row * hcs;
                return c2r_split(ldg(crow + i), ldg(crow + h - i), ldg(rtw + i));
            };
            x_fft<FR1, FR2, +1, false, true>(J.pxi, g, tid, lds, gp(J.tw), split_src, lds_dst);
            __syncthreads();
        }

        // elementwise step on packed real pairs (elements 2n, 2n+1 of a row),
        // in natural order in LDS; loads of XCHUNK elements are issued together
        double acc = 0;
        const int nwork = nrows * h;
        DWHIP_GL float* Uk = gp((iter & 1) ? J.U1 : J.U0);          // x_{k-2} in, x_k out (k > 0)
        const DWHIP_GL float* Ukm1 = gp((iter & 1) ? J.U0 : J.U1);  // x_{k-1}
        DWHIP_GL float* U0 = gp(J.U0);
        DWHIP_GL float* W = gp(J.W);
        const DWHIP_GL float* im = gp(J.im);
        const float bg = st->bg;
        // alpha of iterations k-1 and k (uniform)
        const double am1 = MODE == XM_UPDATE ? shb_alpha(iter - 1, J.alphamax) : 0.0;
        const double al = MODE == XM_UPDATE ? shb_alpha(iter, J.alphamax) : 0.0;
        // the SHB ratio pass's tiles lie within one plane: its rows' y and z
        // come from the tile (no division per element)
        const bool plane_tile = MODE == XM_RATIO && !J.rl;
        const int tz = plane_tile ? row0 / ny : 0;
        const int ty0 = plane_tile ? row0 - tz * ny : 0;
        // chunks of XCHUNK elements per thread; full chunks carry no bounds
        // tests, so their loads and arithmetic schedule as straight-line code
        auto chunk = [&](int w0, auto full) {
            using FULL = decltype(full);
            float2 a[XCHUNK], b[XCHUNK], c[XCHUNK], wv[XCHUNK];
            size_t vi[XCHUNK];
            int yy[XCHUNK], zz[XCHUNK], m0[XCHUNK];
#pragma unroll
            for (int q = 0; q < XCHUNK; ++q) {
                const int w = min(w0 + q * XTHR, nwork - 1);
                const int bb = (int)fdiv((uint32_t)w, hmagic);
                const int n = w - bb * h;
                const int row = row0 + bb;
                if (plane_tile) {
                    yy[q] = ty0 + bb;
                    zz[q] = tz;
                } else {
                    yy[q] = row % ny;
                    zz[q] = row / ny;
                }
                m0[q] = 2 * n;
                vi[q] = (size_t)row * nx + 2 * n;
                a[q] = b[q] = c[q] = wv[q] = make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int q = 0; q < XCHUNK; ++q) {
                if (!FULL::value && w0 + q * XTHR >= nwork) continue;
                if (inv) a[q] = lds[w0 + q * XTHR];
                if (MODE == XM_UPDATE) {
                    b[q] = ldgs<XNT>(reinterpret_cast<const DWHIP_GL float2*>(Ukm1 + vi[q]));
                    if (!J.rl) c[q] = ldgs<XNT>(reinterpret_cast<const DWHIP_GL float2*>(Uk + vi[q]));
                    if (iter > 0 && J.has_w) wv[q] = ldgs<XNT>(reinterpret_cast<const DWHIP_GL float2*>(W + vi[q]));
                } else if (MODE == XM_LOAD) {
                    a[q] = ldg(reinterpret_cast<const DWHIP_GL float2*>(U0 + vi[q]));
                } else if (MODE == XM_RATIO) {
                    if (yy[q] < J.N && zz[q] < J.P) {
                        const DWHIP_GL float* imrow = im + (size_t)J.M * (yy[q] + (size_t)J.N * zz[q]);
                        if (m0[q] < J.M) b[q].x = imrow[m0[q]];
                        if (m0[q] + 1 < J.M) b[q].y = imrow[m0[q] + 1];
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < XCHUNK; ++q) {
                const int w = w0 + q * XTHR;
                if (!FULL::value && w >= nwork) continue;
                float2 out;
                if (MODE == XM_PSF) {
                    const int sy = (yy[q] + st->mid[1]) % ny, sz = (zz[q] + st->mid[2]) % J.nz;
                    float v[2] = {0.f, 0.f};
                    if (sy < J.pN && sz < J.pP) {
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            const int sx = (m0[q] + e + st->mid[0]) % nx;
                            if (sx < J.pM) v[e] = gp(J.psf)[sx + (size_t)J.pM * (sy + (size_t)J.pN * sz)];
                        }
                    }
                    out = make_float2(v[0], v[1]);
                } else if (MODE == XM_LOAD) {
                    out = a[q];
                } else if (MODE == XM_STORE) {
                    out = a[q];
                    stg(reinterpret_cast<DWHIP_GL float2*>(U0 + vi[q]), out);
                } else if (MODE == XM_ONES) {
                    const bool inr = yy[q] < J.N && zz[q] < J.P;
                    out = make_float2(inr && m0[q] < J.M ? 1.f : 0.f, inr && m0[q] + 1 < J.M ? 1.f : 0.f);
                } else if (MODE == XM_WEIGHTS) {
                    // method_shb.c:225-234: W = P1 > 0.01 ? 1/P1 : 0
                    const float2 p1 = a[q];
                    const float sigma = 0.01f;
                    out.x = p1.x > sigma ? (float)(1.0 / p1.x) : 0.f;
                    out.y = p1.y > sigma ? (float)(1.0 / p1.y) : 0.f;
                    stg(reinterpret_cast<DWHIP_GL float2*>(W + vi[q]), out);
                } else if (MODE == XM_UPDATE && J.rl) {
                    // Richardson-Lucy (method_rl.c:85-97, 357-366): f = x_{k-1},
                    // x_k = x~ * (f * W), clamped to bg when bg > 0
                    const float2 f = b[q];
                    float2 xk = f;  // iteration 0: the start guess
                    if (iter > 0) {
                        const float2 xt = a[q];
                        if (J.has_w)
                            xk = make_float2(xt.x * (f.x * wv[q].x), xt.y * (f.y * wv[q].y));
                        else
                            xk = make_float2(xt.x * f.x, xt.y * f.y);
                        if (bg > 0) {
                            xk.x = xk.x < bg ? bg : xk.x;
                            xk.y = xk.y < bg ? bg : xk.y;
                        }
                        stgs<XNT>(reinterpret_cast<DWHIP_GL float2*>(Uk + vi[q]), xk);
                    }
                    out = xk;
                } else if (MODE == XM_UPDATE) {
                    const float2 xkm1 = b[q];
                    float2 xk;
                    if (iter == 0) {
                        xk = c[q];  // U0 = x_0, U1 = x_{-1} = x_0 (start guess)
                    } else {
                        // finish iteration k-1 (method_shb.c:55-70, 357-376)
                        const float2 xt = a[q];
                        const float2 xkm2 = c[q];
                        const float p0 = shb_extrap(xkm1.x, xkm2.x, am1, bg);
                        const float p1 = shb_extrap(xkm1.y, xkm2.y, am1, bg);
                        if (J.has_w)
                            xk = make_float2(xt.x * (p0 * wv[q].x), xt.y * (p1 * wv[q].y));
                        else
                            xk = make_float2(xt.x * p0, xt.y * p1);
                        if (J.positivity) {
                            xk.x = xk.x < bg ? bg : xk.x;
                            xk.y = xk.y < bg ? bg : xk.y;
                        }
                        stgs<XNT>(reinterpret_cast<DWHIP_GL float2*>(Uk + vi[q]), xk);
                    }
                    out = make_float2(shb_extrap(xk.x, xkm1.x, al, bg), shb_extrap(xk.y, xkm1.y, al, bg));
                } else {  // XM_RATIO, method_shb.c:27-48 and getError dw.c:1436-1524
                    const float2 yv = a[q];
                    // outside the image: 0 for SHB (method_shb.c:44), 1e-6 for RL
                    // (method_rl.c:62)
                    const float outside = J.rl ? 1e-6f : 0.f;
                    float r[2] = {outside, outside};
                    if (yy[q] < J.N && zz[q] < J.P) {
#pragma unroll
                        for (int e = 0; e < 2; ++e) {
                            const int m = m0[q] + e;
                            if (m < J.M) {
                                float y = e ? yv.y : yv.x;
                                const float gv = e ? b[q].y : b[q].x;
                                const double yd = y, gd = gv;
                                // g / y of two floats: the correctly rounded float
                                // quotient equals the double quotient rounded to
                                // float (53 >= 2 * 24 + 2), so get_fIdiv's
                                // logf(gval / yval) over doubles needs no double
                                // division, and the ratio reuses it
                                const float q = gv / y;
                                if (J.metric == DWHIP_DEV_METRIC_IDIV) {
                                    if (yd > 0 && gd > 0) acc += gd * (double)logf(q) - (gd - yd);
                                } else {
                                    acc += (yd - gd) * (yd - gd);
                                }
                                if (J.rl) {
                                    r[e] = y > 0 ? q : bg;  // method_rl.c:52-60
                                } else if (fabs(yd) < 1e-6) {
                                    r[e] = gv / (float)copysign(1e-6, yd);  // method_shb.c:41-42
                                } else {
                                    r[e] = q;
                                }
                            }
                        }
                    }
                    out = make_float2(r[0], r[1]);
                }
                lds[w] = out;
            }
                };
        // job flags read once, before the element loops' stores (the
        // compiler otherwise reloads them from the job table per element)
        const bool rl = J.rl;
        const bool has_w = J.has_w;
        const float lo = J.positivity ? bg : -INFINITY;  // x_k < lo -> lo (method_shb.c:366-374)
        if (DWHIP_X_FAST && MODE == XM_UPDATE && !rl && iter > 0) {
            // SHB iteration k > 0.  The tile's x-rows row0 .. row0+nrows-1
            // are consecutive in every real volume (row stride nx = 2h), so
            // packed pair w of the tile is pair w from the tile's first voxel.
            const size_t tb = (size_t)row0 * nx;
            const DWHIP_GL float2* __restrict__ xm1 = reinterpret_cast<const DWHIP_GL float2*>(Ukm1 + tb);
            DWHIP_GL float2* __restrict__ xio = reinterpret_cast<DWHIP_GL float2*>(Uk + tb);  // x_{k-2} in, x_k out
            const DWHIP_GL float2* __restrict__ wt = reinterpret_cast<const DWHIP_GL float2*>(W + tb);
            auto upd = [&](int w0, auto full, auto hw) {
                constexpr bool FULL = decltype(full)::value, HW = decltype(hw)::value;
                float2 a[XCHUNK], b[XCHUNK], c[XCHUNK], wv[XCHUNK];
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q) {
                    const int w = w0 + q * XTHR;
                    if (!FULL && w >= nwork) continue;
                    a[q] = lds[w];
                    b[q] = ldgs<DWHIP_NT_STATE>(xm1 + w);
                    c[q] = ldgs<DWHIP_NT_STATE>(xio + w);
                    wv[q] = HW ? ldgs<DWHIP_NT_STATE>(wt + w) : make_float2(1.f, 1.f);
                }
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q) {
                    const int w = w0 + q * XTHR;
                    if (!FULL && w >= nwork) continue;
                    // finish iteration k-1 (method_shb.c:55-70, 357-376)
                    const float p0 = shb_extrap(b[q].x, c[q].x, am1, bg);
                    const float p1 = shb_extrap(b[q].y, c[q].y, am1, bg);
                    float2 xk = HW ? make_float2(a[q].x * (p0 * wv[q].x), a[q].y * (p1 * wv[q].y))
                                   : make_float2(a[q].x * p0, a[q].y * p1);
                    xk.x = xk.x < lo ? lo : xk.x;
                    xk.y = xk.y < lo ? lo : xk.y;
                    stgs<DWHIP_NT_STATE>(xio + w, xk);
                    lds[w] = make_float2(shb_extrap(xk.x, b[q].x, al, bg), shb_extrap(xk.y, b[q].y, al, bg));
                }
            };
            auto rows = [&](auto hw) {
                int w0 = tid;
                for (; w0 + (XCHUNK - 1) * XTHR < nwork; w0 += XCHUNK * XTHR) upd(w0, BoolC<true>(), hw);
                for (; w0 < nwork; w0 += XCHUNK * XTHR) upd(w0, BoolC<false>(), hw);
            };
            if (has_w)
                rows(BoolC<true>());
            else
                rows(BoolC<false>());
        } else if (DWHIP_X_FAST && MODE == XM_RATIO && plane_tile && h > 1) {
            // SHB ratio over a plane tile: every row is an image row (y < N,
            // z < P); pair w holds columns 2n, 2n+1 of tile row w / h.  Image
            // loads are clamped into the tile's rows and dropped beyond
            // column M (method_shb.c:27-48, getError dw.c:1436-1524).
            const int M = J.M;
            const DWHIP_GL float* __restrict__ imt = im + (size_t)M * (ty0 + (size_t)J.N * tz);
            const int ilast = nrows * M - 1;
            auto rat = [&](int w0, auto full, auto idiv) {
                constexpr bool FULL = decltype(full)::value, IDIV = decltype(idiv)::value;
                float2 a[XCHUNK], gv[XCHUNK];
                int m0[XCHUNK];
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q) {
                    const int w = FULL ? w0 + q * XTHR : min(w0 + q * XTHR, nwork - 1);
                    const int bb = (int)__umulhi((uint32_t)w, hmagic);  // h > 1: hmagic is the magic of h
                    const int n = w - bb * h;
                    m0[q] = 2 * n;
                    const int i0 = bb * M + 2 * n;
                    a[q] = lds[w];
                    gv[q] = make_float2(imt[min(i0, ilast)], imt[min(i0 + 1, ilast)]);
                }
#pragma unroll
                for (int q = 0; q < XCHUNK; ++q) {
                    const int w = w0 + q * XTHR;
                    if (!FULL && w >= nwork) continue;
                    float r[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const bool valid = m0[q] + e < M;
                        const float y = e ? a[q].y : a[q].x;
                        const float g = e ? gv[q].y : gv[q].x;
                        const double yd = y, gd = g;
                        // the correctly rounded float quotient equals the
                        // double quotient rounded to float, so get_fIdiv's
                        // logf(gval / yval) needs no double division
                        const float qt = g / y;
                        double term;
                        if (IDIV)
                            term = (y > 0 && g > 0) ? gd * (double)logf(qt) - (gd - yd) : 0.0;
                        else
                            term = (yd - gd) * (yd - gd);
                        acc += valid ? term : 0.0;
                        float rv = qt;
                        if (fabs(yd) < 1e-6) rv = g / (float)copysign(1e-6, yd);  // method_shb.c:41-42
                        r[e] = valid ? rv : 0.f;
                    }
                    lds[w] = make_float2(r[0], r[1]);
                }
            };
            auto rows = [&](auto idiv) {
                int w0 = tid;
                for (; w0 + (XCHUNK - 1) * XTHR < nwork; w0 += XCHUNK * XTHR) rat(w0, BoolC<true>(), idiv);
                for (; w0 < nwork; w0 += XCHUNK * XTHR) rat(w0, BoolC<false>(), idiv);
            };
            if (J.metric == DWHIP_DEV_METRIC_IDIV)
                rows(BoolC<true>());
            else
                rows(BoolC<false>());
        } else {
            int w0 = tid;
            for (; w0 + (XCHUNK - 1) * XTHR < nwork; w0 += XCHUNK * XTHR) chunk(w0, BoolC<true>());
            for (; w0 < nwork; w0 += XCHUNK * XTHR) chunk(w0, BoolC<false>());
        }
        if (MODE == XM_RATIO) {
            const double s = block_sum<XTHR>(acc, red);
            if (tid == 0) gp(J.part)[lt] = s;
        }
        __syncthreads();

        if (fwd) {
            x_fft<FR1, FR2, -1, true, true>(J.pxf, g, tid, lds, gp(J.tw), lds_src, lds_dst);
            __syncthreads();
            [[maybe_unused]] const int hp = h / 2 + 1;
            DWHIP_GL float2* __restrict__ dst = gp(J.C) + (size_t)row0 * hcs;
            // (the SHB ratio pass's plane tiles hold image rows only, so
            // every row of a tile is written)
#if DWHIP_MERGE_BY_BIN == 2
            // a thread writes bin m (0 <= m <= h) of every row of the tile,
            // X[m] = merge(Z[m], Z[h - m], exp(-2 pi i m/nx)), with one twiddle;
            // a wave's stores cover whole lines of each row
            for (int m = tid; m <= h; m += XTHR) {
                const float2 wm = ldg(rtw + (m < h ? m : 0));
                for (int b = 0; b < nrows; ++b) {
                    const float2* L = lds + b * h;
                    float2 v;
                    if (m == 0 || m == h) {
                        const float2 z0 = L[0];
                        v = make_float2(m == 0 ? z0.x + z0.y : z0.x - z0.y, 0.f);
                    } else {
                        v = r2c_merge(L[m], L[h - m], wm);
                    }
                    stgs<XNT>(dst + (size_t)b * hcs + m, v);
                }
            }
#elif DWHIP_MERGE_BY_BIN
            // a thread merges bin k and its mirror h - k of every row of the
            // tile, with one pair of twiddles
            for (int k = tid; k < hp; k += XTHR) {
                const float2 wk = ldg(rtw + k);
                const float2 wm = ldg(rtw + (k == 0 ? 0 : h - k));
                for (int b = 0; b < nrows; ++b) {
                    const float2* L = lds + b * h;
                    DWHIP_GL float2* crow = dst + (size_t)b * hcs;
                    if (k == 0) {
                        const float2 z0 = L[0];
                        stgs<XNT>(crow, make_float2(z0.x + z0.y, 0.f));
                        stgs<XNT>(crow + h, make_float2(z0.x - z0.y, 0.f));
                    } else {
                        const float2 a = L[k], c = L[h - k];
                        stgs<XNT>(crow + k, r2c_merge(a, c, wk));
                        if (2 * k != h) stgs<XNT>(crow + h - k, r2c_merge(c, a, wm));
                    }
                }
            }
#else
            const uint32_t hpmagic = fdiv_magic((uint32_t)hp);
            for (int w = tid; w < nrows * hp; w += XTHR) {
                const int b = (int)fdiv((uint32_t)w, hpmagic);
                const int k = w - b * hp;
                const float2* L = lds + b * h;
                DWHIP_GL float2* crow = dst + (size_t)b * hcs;
                if (k == 0) {
                    const float2 z0 = L[0];
                    stgs<XNT>(crow, make_float2(z0.x + z0.y, 0.f));
                    stgs<XNT>(crow + h, make_float2(z0.x - z0.y, 0.f));
                } else {
                    const float2 a = L[k], c = L[h - k];
                    stgs<XNT>(crow + k, r2c_merge(a, c, ldg(rtw + k)));
                    if (2 * k != h) stgs<XNT>(crow + h - k, r2c_merge(c, a, ldg(rtw + h - k)));
                }
            }
#endif
        }
        __syncthreads();
    }
}

// Tile order of a pass: the update and ratio passes can walk their tiles
// from the last plane to the first (DWHIP_REV_UPDATE, DWHIP_REV_RATIO), so
// that they start on the planes the pass before them wrote last and end on
// the planes the pass after them reads first, which the device's last-level
// cache may still hold.  The update pass does (its y neighbours run from
// the first plane to the last).
#ifndef DWHIP_REV_UPDATE
#define DWHIP_REV_UPDATE 1
#endif
#ifndef DWHIP_REV_RATIO
#define DWHIP_REV_RATIO 0
#endif
template <int MODE>
constexpr bool x_reversed()
{
    return (MODE == XM_UPDATE && DWHIP_REV_UPDATE) || (MODE == XM_RATIO && DWHIP_REV_RATIO);
}

template <int MODE>
__global__ __launch_bounds__(XTHR) DWHIP_X_OCC void k_xpass(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs)
{
    extern __shared__ float2 lds[];
    __shared__ double red[XTHR / 64];
    constexpr int kind = xmode_kind(MODE);
    const int total = bd->ntile[kind];
    const int njobs = bd->njobs;
    for (int t0 = blockIdx.x; t0 < total; t0 += gridDim.x) {
        const int t = x_reversed<MODE>() ? total - 1 - t0 : t0;
        const JobDev& J = jobs[find_job(jobs, njobs, kind, t)];
        if (J.dev_stop && gp(J.st)->stop) continue;  // the stopping rule ended this job
        if constexpr (x_fixed_mode(MODE)) {
            if (J.pxf.nst == 2 && J.pxf.radix[0] == 25 && J.pxf.radix[1] == 25) {
                xpass_tile<MODE, 25, 25>(J, t, threadIdx.x, lds, red);
                continue;
            }
        }
        xpass_tile<MODE, 0, 0>(J, t, threadIdx.x, lds, red);
    }
}

// ---------------------------------------------------------------------------
// Column passes (Y and Z)
// ---------------------------------------------------------------------------
// OTF multiply at the Z junction: fft_mul_inplace (fft.c:437-453) and
// fft_mul_conj_inplace (fft.c:490-512) with their operand order.
template <bool CONJ, bool NT>
struct OtfOp {
    const DWHIP_GL float2* __restrict__ K;
    size_t stride;
    int npv;    // pencils holding data
    bool full;  // npv is the whole tile
    static constexpr bool kStrided = true;
    __device__ __forceinline__ float2 fetch(int pen, int i) const
    {
        return pen < npv ? ldgs<NT>(K + i * stride + pen) : make_float2(0.f, 0.f);
    }
    // elements j + r*nb: one address per butterfly, then a uniform step
    template <int R>
    __device__ __forceinline__ void fetch_r(int pen, int j, int nb, float2* v) const
    {
        const DWHIP_GL float2* p = K + pen + (size_t)j * stride;
        const size_t step = (size_t)nb * stride;
        if (full) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = ldgs<NT>(p + r * step);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = pen < npv ? ldgs<NT>(p + r * step) : make_float2(0.f, 0.f);
        }
    }
    __device__ __forceinline__ float2 apply(float2 v, float2 kk) const
    {
        const float a = kk.x, ac = CONJ ? -kk.y : kk.y;
        return make_float2(a * v.x - ac * v.y, a * v.y + v.x * ac);
    }
};

// Pencils of a column tile in global memory: element i of pencil pen at
// p[i * stride + pen].  Pencils from npv on (row padding past hc) and
// elements from nv on (rows a pruned pass skips) read as zero and are not
// written; `full` (every pencil and element in range, the usual tile) skips
// those tests.
template <bool NT>
struct ColPencils {
    DWHIP_GL float2* p;
    size_t stride;
    int npv, nv;
    bool full;
    static constexpr bool kStrided = true;
    __device__ __forceinline__ float2 operator()(int pen, int i) const
    {
        return (pen < npv && i < nv) ? ldgs<NT>(p + i * stride + pen) : make_float2(0.f, 0.f);
    }
    __device__ __forceinline__ void operator()(int pen, int i, float2 v) const
    {
        if (pen < npv && i < nv) stgs<NT>(p + i * stride + pen, v);
    }
    template <int R>
    __device__ __forceinline__ void load(int pen, int j, int nb, float2* v) const
    {
        const DWHIP_GL float2* q = p + pen + (size_t)j * stride;
        const size_t step = (size_t)nb * stride;
        if (full) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[r] = ldgs<NT>(q + r * step);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                v[r] = (pen < npv && j + r * nb < nv) ? ldgs<NT>(q + r * step) : make_float2(0.f, 0.f);
        }
    }
    template <int R>
    __device__ __forceinline__ void store(int pen, int j, int nb, const float2* v) const
    {
        DWHIP_GL float2* q = p + pen + (size_t)j * stride;
        const size_t step = (size_t)nb * stride;
        if (full) {
#pragma unroll
            for (int r = 0; r < R; ++r) stgs<NT>(q + r * step, v[r]);
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (pen < npv && j + r * nb < nv) stgs<NT>(q + r * step, v[r]);
        }
    }
};

// Workgroup b runs on XCD b mod 8 (dispatch is round robin over the 8 XCDs,
// each with its own L2).  Column tile t covers B adjacent kx columns, so
// tiles t and t+1 share the 128-byte lines of a row when B*8 < 128: map
// consecutive workgroups of one XCD to consecutive tiles, so that both
// halves of a line are read through the same L2.  A bijection of [0, G).
__device__ __forceinline__ int xcd_tile(int b, int G)
{
    const int q = G >> 3, r = G & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

template <int MODE>
constexpr int col_threads()
{
    return cmode_is_y(MODE) ? CTHR : ZTHR;
}

// waves per SIMD the column kernels are compiled for (the VGPR budget):
// Y tiles (1024 threads, up to 159 KiB of LDS, one per CU) need 4 and stay
// within 128 VGPRs; Z tiles (512 threads, 10 values each, 35 KiB for
// nz = 140) fit 8 in 64 VGPRs, four workgroups per CU
#ifndef DWHIP_Y_WAVES
#define DWHIP_Y_WAVES 4
#endif
#ifndef DWHIP_Z_WAVES
#define DWHIP_Z_WAVES 8
#endif
template <int MODE>
constexpr int col_waves()
{
    return cmode_is_y(MODE) ? DWHIP_Y_WAVES : DWHIP_Z_WAVES;
}

template <int MODE>
__global__ __launch_bounds__(col_threads<MODE>()) __attribute__((amdgpu_waves_per_eu(col_waves<MODE>())))
void k_colpass(const BatchDev* __restrict__ bd, const JobDev* __restrict__ jobs)
{
    extern __shared__ float2 lds[];
    const int tid = threadIdx.x;
    constexpr int kind = cmode_kind(MODE);
    constexpr bool ykind = cmode_is_y(MODE);
    constexpr int NT = col_threads<MODE>();
    constexpr bool cnt = DWHIP_NT_COLS;
    static_assert(MODE >= 0 && MODE < DWHIP_N_CMODES, "column mode");
    const int total = bd->ntile[kind];
    const int njobs = bd->njobs;
    for (int t0 = blockIdx.x; t0 < total; t0 += gridDim.x) {
        // XCD-aware order when the grid covers the tiles exactly once
        const int t = (ykind && (int)gridDim.x == total) ? xcd_tile(t0, total) : t0;
        const JobDev& J = jobs[find_job(jobs, njobs, kind, t)];
        if (J.dev_stop && gp(J.st)->stop) continue;  // the stopping rule ended this job
        const int lt = t - J.tile0[kind];
        const int hcs = J.hcs, ny = J.ny;
        const int B = ykind ? J.ycols : J.zcols;
        const int nct = (J.hc + B - 1) / B;
        const int outer = lt / nct;
        const int kx0 = (lt - outer * nct) * B;
        const int nb = min(B, J.hc - kx0);   // columns past hc are row padding, never touched
        const size_t stride = ykind ? (size_t)hcs : (size_t)hcs * ny;
        const size_t off = ykind ? (size_t)outer * ny * hcs + kx0 : (size_t)outer * hcs + kx0;
        DWHIP_GL float2* __restrict__ dst = gp(MODE == CM_ZOTF ? J.K : J.C) + off;
        StageGeo g;
        g.n = ykind ? ny : J.nz;
        g.npen = B;
        g.pstride = 0;
        g.nthr = NT;
        g.npshift = __ffs(B) - 1;  // B is a power of two
        // elements read (i < nin) and written (i < nout) along the pencil
        const int nin = (MODE == CM_YFWD_IMG && !J.rl) ? J.N : (MODE == CM_ZCONVJ && !J.rl) ? J.P : g.n;
        const int nout = (MODE == CM_YINV_IMG) ? J.N : (MODE == CM_ZCONV) ? J.P : g.n;

        // (Z tiles only: the unmasked forms push 1024-thread Y tiles past 128 VGPRs)
        const ColPencils<cnt> gsrc{gp(J.C) + off, stride, nb, nin, !ykind && nb == B && nin == g.n};
        const ColPencils<cnt> gdst{dst, stride, nb, nout, !ykind && nb == B && nout == g.n};
        if (MODE == CM_YFWD || MODE == CM_YFWD_IMG) {
            fft_tile<DWHIP_YMAXE, -1, true, DWHIP_RMAX_Y, false, false>(J.pyf, g, tid, lds, gp(J.tw), gsrc, gdst);
        } else if (MODE == CM_YINV || MODE == CM_YINV_IMG) {
            fft_tile<DWHIP_YMAXE, +1, true, DWHIP_RMAX_Y, false, false>(J.pyi, g, tid, lds, gp(J.tw), gsrc, gdst);
        } else if (MODE == CM_ZFWD) {
            fft_tile<DWHIP_ZMAXE, -1, true, DWHIP_RMAX_Z, false, false>(J.pzf, g, tid, lds, gp(J.tw), gsrc, gdst);
        } else if (MODE == CM_ZINV) {
            fft_tile<DWHIP_ZMAXE, +1, true, DWHIP_RMAX_Z, false, false>(J.pzi, g, tid, lds, gp(J.tw), gsrc, gdst);
        } else if (MODE == CM_ZOTF) {
            // (float)nx*ny*nz, evaluated left to right in float (fft.c:856)
            const float scale = ((float)J.nx * (float)J.ny) * (float)J.nz;
            auto sdst = [&](int pen, int i, float2 v) {
                if (pen < nb) stg(dst + i * stride + pen, make_float2(v.x / scale, v.y / scale));
            };
            fft_tile<DWHIP_ZMAXE, -1, true, DWHIP_RMAX_Z, false, false>(J.pzf, g, tid, lds, gp(J.tw), gsrc, sdst);
        } else {
            OtfOp<MODE == CM_ZCONVJ, cnt> op{gp(J.K) + off, stride, nb, nb == B};
            fft_pair_tile<DWHIP_ZMAXE, -1, +1, true, DWHIP_RMAX_Z>(J.pzf, J.pzi, g, tid, lds, gp(J.tw), gsrc, op, gdst);
        }
        __syncthreads();
    }
}

// Dynamic LDS above 64 KiB must be allowed per kernel (gfx950 has 160 KiB
// per workgroup).  The limit asked for leaves 1 KiB for static LDS.
static inline void dwhip_allow_big_lds(const void* kernel)
{
    int v = 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, 0) != hipSuccess) v = 65536;
    hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, v - 1024);
    if (e != hipSuccess) {
        fprintf(stderr, "dwhip: raising the LDS limit to %d bytes failed: %s\n", v - 1024, hipGetErrorString(e));
        (void)hipGetLastError();
    }
}
