// Mixed-radix Stockham FFT staged in LDS, for CDNA4 (gfx950).
//
// A workgroup transforms `npen` pencils of length n that sit in LDS.  Each
// stage moves every element LDS -> registers -> LDS once: a thread first
// loads all of its butterflies' inputs, the group synchronises, then the
// thread applies twiddles, the radix-R butterfly, and stores the outputs in
// Stockham order.  Holding a full stage in registers makes the stage
// in-place in LDS, so one LDS buffer per pencil suffices.
//
// Stage (radix R, sub-transform length ns, nb = n/R butterflies per pencil):
//   butterfly j (0 <= j < nb), k = j mod ns:
//     v[r]  = in[j + r*nb] * exp(sign*2*pi*i*r*k/(ns*R))      r = 0..R-1
//     V     = DFT_R(v)
//     out[(j/ns)*ns*R + k + r*ns] = V[r]
// This is the autosort (Stockham) formulation; after the last stage the
// pencil is in natural order.
//
// Two LDS layouts:
//   COLS: element i of pencil b at lds[i*npen + b] (column tiles, where
//         adjacent pencils are adjacent in global memory),
//   rows: element i of pencil b at lds[b*pstride + i] (x-rows).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DWHIP_MAX_STAGES 12

struct FftPlanDev {
    int n;
    int nst;
    int radix[DWHIP_MAX_STAGES];
    int ns[DWHIP_MAX_STAGES];
    int twoff[DWHIP_MAX_STAGES];   // offset of the stage's twiddles in the table
    uint32_t nsmagic[DWHIP_MAX_STAGES];  // ceil(2^32/ns) for k = j mod ns
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)  // a * conj(b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (DIR<0) or +i (DIR>0)
template <int DIR>
__device__ __forceinline__ float2 mul_mi(float2 a)
{
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// q = floor(j / d) with magic = ceil(2^32 / d), exact whenever j*d < 2^32
// (the rounding excess of magic is below d, so j*excess/2^32 < 1/d).
// d = 1 is encoded as magic = 0.
__device__ __forceinline__ uint32_t fdiv(uint32_t j, uint32_t magic)
{
    return magic ? __umulhi(j, magic) : j;
}
__host__ __device__ __forceinline__ uint32_t fdiv_magic(uint32_t d)
{
    return d <= 1 ? 0u : (uint32_t)((0x100000000ull + d - 1) / (uint64_t)d);
}

// ---- small DFTs, in place on v[0..R-1]; DIR = -1 forward, +1 inverse ----
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Dft<3, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = -0.5f;
        const float s1 = DIR * 0.86602540378443864676f;  // sin(2pi/3)*DIR
        float2 a = v[0], b = v[1], c = v[2];
        float2 t = cadd(b, c);
        float2 d = csub(b, c);
        v[0] = cadd(a, t);
        float2 m = make_float2(a.x + c1 * t.x, a.y + c1 * t.y);
        // +/- i*s1*d
        float2 r = make_float2(-s1 * d.y, s1 * d.x);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    }
};

template <int DIR> struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 t0 = cadd(v[0], v[2]);
        float2 t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]);
        float2 t3 = mul_mi<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR> struct Dft<5, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = 0.30901699437494742410f;   // cos(2pi/5)
        const float c2 = -0.80901699437494742410f;  // cos(4pi/5)
        const float s1 = DIR * 0.95105651629515357212f;  // sin(2pi/5)
        const float s2 = DIR * 0.58778525229247312917f;  // sin(4pi/5)
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        float2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        v[0] = make_float2(a.x + b1.x + b2.x, a.y + b1.y + b2.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
        // i*(s1*d1 + s2*d2) and i*(s2*d1 - s1*d2)
        float2 n1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        float2 n2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        float2 in1 = make_float2(-n1.y, n1.x);
        float2 in2 = make_float2(-n2.y, n2.x);
        v[1] = cadd(m1, in1);
        v[4] = csub(m1, in1);
        v[2] = cadd(m2, in2);
        v[3] = csub(m2, in2);
    }
};

template <int DIR> struct Dft<7, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = 0.62348980185873353053f;   // cos(2pi/7)
        const float c2 = -0.22252093395631440429f;  // cos(4pi/7)
        const float c3 = -0.90096886790241912624f;  // cos(6pi/7)
        const float s1 = DIR * 0.78183148246802980871f;
        const float s2 = DIR * 0.97492791218182360702f;
        const float s3 = DIR * 0.43388373911755812048f;
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[6]), d1 = csub(v[1], v[6]);
        float2 b2 = cadd(v[2], v[5]), d2 = csub(v[2], v[5]);
        float2 b3 = cadd(v[3], v[4]), d3 = csub(v[3], v[4]);
        v[0] = make_float2(a.x + b1.x + b2.x + b3.x, a.y + b1.y + b2.y + b3.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x + c3 * b3.x,
                                a.y + c1 * b1.y + c2 * b2.y + c3 * b3.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c3 * b2.x + c1 * b3.x,
                                a.y + c2 * b1.y + c3 * b2.y + c1 * b3.y);
        float2 m3 = make_float2(a.x + c3 * b1.x + c1 * b2.x + c2 * b3.x,
                                a.y + c3 * b1.y + c1 * b2.y + c2 * b3.y);
        // sin(2pi*q*m/7) for m=1..3 against d1..d3:
        // m=1: s1 d1 + s2 d2 + s3 d3
        // m=2: s2 d1 - s3 d2 - s1 d3
        // m=3: s3 d1 - s1 d2 + s2 d3
        float2 n1 = make_float2(s1 * d1.x + s2 * d2.x + s3 * d3.x, s1 * d1.y + s2 * d2.y + s3 * d3.y);
        float2 n2 = make_float2(s2 * d1.x - s3 * d2.x - s1 * d3.x, s2 * d1.y - s3 * d2.y - s1 * d3.y);
        float2 n3 = make_float2(s3 * d1.x - s1 * d2.x + s2 * d3.x, s3 * d1.y - s1 * d2.y + s2 * d3.y);
        float2 i1 = make_float2(-n1.y, n1.x);
        float2 i2 = make_float2(-n2.y, n2.x);
        float2 i3 = make_float2(-n3.y, n3.x);
        v[1] = cadd(m1, i1);
        v[6] = csub(m1, i1);
        v[2] = cadd(m2, i2);
        v[5] = csub(m2, i2);
        v[3] = cadd(m3, i3);
        v[4] = csub(m3, i3);
    }
};

template <int DIR> struct Dft<8, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float r2 = 0.70710678118654752440f;
        // radix-2 over stride 4
        float2 a0 = cadd(v[0], v[4]), a4 = csub(v[0], v[4]);
        float2 a1 = cadd(v[1], v[5]), a5 = csub(v[1], v[5]);
        float2 a2 = cadd(v[2], v[6]), a6 = csub(v[2], v[6]);
        float2 a3 = cadd(v[3], v[7]), a7 = csub(v[3], v[7]);
        // twiddles w8^k on the odd half: w8 = exp(DIR*i*pi/4)
        a5 = make_float2(r2 * (a5.x - DIR * a5.y), r2 * (a5.y + DIR * a5.x));
        a6 = mul_mi<DIR>(a6);  // times exp(DIR*i*pi/2) = DIR*i
        a7 = make_float2(r2 * (-a7.x - DIR * a7.y), r2 * (-a7.y + DIR * a7.x));
        // two radix-4 on (a0,a1,a2,a3) and (a4,a5,a6,a7)
        float2 e[4] = {a0, a1, a2, a3};
        float2 o[4] = {a4, a5, a6, a7};
        Dft<4, DIR>::run(e);
        Dft<4, DIR>::run(o);
        v[0] = e[0]; v[2] = e[1]; v[4] = e[2]; v[6] = e[3];
        v[1] = o[0]; v[3] = o[1]; v[5] = o[2]; v[7] = o[3];
    }
};

// LDS index of element i of pencil b.
template <bool COLS>
__device__ __forceinline__ int lds_idx(int b, int i, int npen, int pstride)
{
    return COLS ? i * npen + b : b * pstride + i;
}

// One Stockham stage over all pencils of the group.
template <int R, int MB, int DIR, bool COLS>
__device__ __forceinline__ void fft_stage(float2* lds, int n, int npen, int pstride, int ns,
                                          uint32_t nsmagic, const float2* __restrict__ tw,
                                          uint32_t npmagic, uint32_t nbmagic, int tid, int nthr)
{
    const int nb = n / R;
    const int total = nb * npen;
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * nthr;
        if (w < total) {
            int b, j;
            if (COLS) {
                j = (int)fdiv((uint32_t)w, npmagic);
                b = w - j * npen;
            } else {
                b = (int)fdiv((uint32_t)w, nbmagic);
                j = w - b * nb;
            }
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = lds[lds_idx<COLS>(b, j + r * nb, npen, pstride)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * nthr;
        if (w < total) {
            int b, j;
            if (COLS) {
                j = (int)fdiv((uint32_t)w, npmagic);
                b = w - j * npen;
            } else {
                b = (int)fdiv((uint32_t)w, nbmagic);
                j = w - b * nb;
            }
            const int jq = (int)fdiv((uint32_t)j, nsmagic);
            const int k = j - jq * ns;
            if (ns > 1) {
                const float2* t = tw + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    float2 wv = t[r - 1];
                    v[q][r] = DIR < 0 ? cmul(v[q][r], wv) : cmulc(v[q][r], wv);
                }
            }
            Dft<R, DIR>::run(v[q]);
            const int base = jq * ns * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lds_idx<COLS>(b, base + r * ns, npen, pstride)] = v[q][r];
        }
    }
    __syncthreads();
}

// Full transform of the npen pencils in LDS.  MAXE bounds the elements a
// thread holds per stage: the host guarantees n*npen <= MAXE*nthr.
// Twiddles (forward sign) come from `tw`; DIR>0 conjugates them.
template <int MAXE, int DIR, bool COLS>
__device__ void fft_lds(float2* lds, const FftPlanDev& p, const float2* __restrict__ tw, int npen,
                        int pstride, int tid, int nthr)
{
    const uint32_t npmagic = fdiv_magic((uint32_t)npen);
    for (int s = 0; s < p.nst; ++s) {
        const int R = p.radix[s];
        const int nb = p.n / R;
        const uint32_t nbmagic = fdiv_magic((uint32_t)nb);
        const float2* t = tw + p.twoff[s];
        switch (R) {
        case 2: fft_stage<2, (MAXE + 1) / 2, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        case 3: fft_stage<3, (MAXE + 2) / 3, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        case 4: fft_stage<4, (MAXE + 3) / 4, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        case 5: fft_stage<5, (MAXE + 4) / 5, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        case 7: fft_stage<7, (MAXE + 6) / 7, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        case 8: fft_stage<8, (MAXE + 7) / 8, DIR, COLS>(lds, p.n, npen, pstride, p.ns[s], p.nsmagic[s], t, npmagic, nbmagic, tid, nthr); break;
        default: break;
        }
    }
}
