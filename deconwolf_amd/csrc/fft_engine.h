// Register-first mixed-radix Stockham FFT for CDNA4 (gfx950).
//
// A workgroup transforms `npen` pencils of length n.  A stage with radix R
// has nb = n/R butterflies per pencil; butterfly j of a pencil reads
// elements j + r*nb (r < R), multiplies them by exp(s 2 pi i r k/(ns R)),
// k = j mod ns, applies the length-R DFT and writes element
// (j/ns)*ns*R + k + r*ns (the autosort Stockham step; ns is the product of
// the earlier radices).  After the last stage the pencil is in natural
// order.
//
// Data stays in registers inside a stage; consecutive stages exchange
// through LDS (one LDS buffer, written in place: all reads of a stage finish
// before its writes).  The first stage reads straight from global memory
// (or from a generator), the last one writes straight to global memory, and
// two stages that meet at a junction (last of one transform, first of the
// next) are done in one register pass:
//   forward-last | elementwise op | inverse-first  (OTF multiply), or
//   inverse-last | elementwise op | forward-first  (SHB update, ratio).
// That needs the radix of the junction's two sides to agree: the forward
// plan uses radices R1..Rk and the inverse plan Rk..R1.
//
// Thread mapping: work item w = tid + q*nthr (q < MB) is butterfly (pen, j)
// with, for column tiles (COLS), pen = w mod npen, j = w / npen, so that a
// wavefront reads adjacent pencils (adjacent addresses in global memory);
// for rows, j = w mod nb, pen = w / nb.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DWHIP_MAX_STAGES 12

struct FftPlanDev {
    int n;
    int nst;
    int radix[DWHIP_MAX_STAGES];
    int ns[DWHIP_MAX_STAGES];
    int twoff[DWHIP_MAX_STAGES];         // stage twiddles in the table (forward sign)
    uint32_t nsmagic[DWHIP_MAX_STAGES];  // fdiv magic of ns
};

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
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

// q = floor(j / d) with magic = ceil(2^32 / d), exact whenever j*d < 2^32
// (the rounding excess of magic is below d, so j*excess/2^32 < 1/d).
// d = 1 is encoded as magic = 0.
__device__ __forceinline__ uint32_t fdiv(uint32_t j, uint32_t magic)
{
    return magic ? __umulhi(j, magic) : j;
}
// ceil(2^32 / d) = floor((2^32 - 1) / d) + 1 for every d >= 2 (d divides
// 2^32 only as a power of two, where both sides are 2^32 / d), so one 32-bit
// division gives it; for d = 1 the sum wraps to the encoding 0.
__host__ __device__ __forceinline__ uint32_t fdiv_magic(uint32_t d)
{
    return d == 0 ? 0u : 0xffffffffu / d + 1u;
}

// ---------------------------------------------------------------------------
// compile-time trigonometry (Taylor series in double, |x| <= pi)
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_red(double x)
{
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k < 24; ++k) {
        term *= -x2 / (double)((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_red(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 24; ++k) {
        term *= -x2 / (double)((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
// exp(2 pi i m / R) components, m taken mod R into (-R/2, R/2]
constexpr double cx_cos2pi(int m, int R)
{
    m %= R;
    if (m < 0) m += R;
    if (2 * m > R) m -= R;
    return cx_cos_red(2.0 * kPi * (double)m / (double)R);
}
constexpr double cx_sin2pi(int m, int R)
{
    m %= R;
    if (m < 0) m += R;
    if (2 * m > R) m -= R;
    return cx_sin_red(2.0 * kPi * (double)m / (double)R);
}

// cos/sin(2 pi m / R) for m < R, evaluated by the compiler
template <int R> struct RootTable {
    float c[R];
    float s[R];
    constexpr RootTable() : c(), s()
    {
        for (int m = 0; m < R; ++m) {
            c[m] = (float)cx_cos2pi(m, R);
            s[m] = (float)cx_sin2pi(m, R);
        }
    }
};
template <int R> struct Roots {
    static constexpr RootTable<R> t{};
};

// ---------------------------------------------------------------------------
// small DFTs: v[m] <- sum_q v[q] exp(DIR 2 pi i q m / R)
// ---------------------------------------------------------------------------
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
    __device__ __forceinline__ static void run(float2*) {}
};

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
        const float s1 = DIR * 0.86602540378443864676f;
        float2 a = v[0];
        float2 t = cadd(v[1], v[2]);
        float2 d = csub(v[1], v[2]);
        v[0] = cadd(a, t);
        float2 m = make_float2(a.x + c1 * t.x, a.y + c1 * t.y);
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
        float2 d = csub(v[1], v[3]);
        // times DIR*i
        float2 t3 = DIR < 0 ? make_float2(d.y, -d.x) : make_float2(-d.y, d.x);
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR> struct Dft<5, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = 0.30901699437494742410f;
        const float c2 = -0.80901699437494742410f;
        const float s1 = DIR * 0.95105651629515357212f;
        const float s2 = DIR * 0.58778525229247312917f;
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        float2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        v[0] = make_float2(a.x + b1.x + b2.x, a.y + b1.y + b2.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
        float2 n1 = make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        float2 n2 = make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        float2 i1 = make_float2(-n1.y, n1.x);
        float2 i2 = make_float2(-n2.y, n2.x);
        v[1] = cadd(m1, i1);
        v[4] = csub(m1, i1);
        v[2] = cadd(m2, i2);
        v[3] = csub(m2, i2);
    }
};

template <int DIR> struct Dft<7, DIR> {
    __device__ __forceinline__ static void run(float2* v)
    {
        const float c1 = 0.62348980185873353053f;
        const float c2 = -0.22252093395631440429f;
        const float c3 = -0.90096886790241912624f;
        const float s1 = DIR * 0.78183148246802980871f;
        const float s2 = DIR * 0.97492791218182360702f;
        const float s3 = DIR * 0.43388373911755812048f;
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[6]), d1 = csub(v[1], v[6]);
        float2 b2 = cadd(v[2], v[5]), d2 = csub(v[2], v[5]);
        float2 b3 = cadd(v[3], v[4]), d3 = csub(v[3], v[4]);
        v[0] = make_float2(a.x + b1.x + b2.x + b3.x, a.y + b1.y + b2.y + b3.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x + c3 * b3.x, a.y + c1 * b1.y + c2 * b2.y + c3 * b3.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c3 * b2.x + c1 * b3.x, a.y + c2 * b1.y + c3 * b2.y + c1 * b3.y);
        float2 m3 = make_float2(a.x + c3 * b1.x + c1 * b2.x + c2 * b3.x, a.y + c3 * b1.y + c1 * b2.y + c2 * b3.y);
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

// Composite R = F*Q (F = 4 when 4 | R and R != 4, else the smallest prime):
// decimation in time, x[F*n2 + n1]:
//   Y_n1 = DFT_Q(x[F*n2 + n1]) ,  X[k1 + Q*k2] = sum_n1 w_F^(n1 k2) w_R^(n1 k1) Y_n1[k1]
constexpr int first_factor(int R)
{
    if (R % 4 == 0 && R != 4) return 4;
    for (int f = 2; f <= R; ++f)
        if (R % f == 0) return f;
    return R;
}
constexpr bool is_base_radix(int R) { return R == 1 || R == 2 || R == 3 || R == 4 || R == 5 || R == 7; }

template <int R, int DIR, bool BASE = is_base_radix(R)> struct DftC;

template <int R, int DIR> struct DftC<R, DIR, true> {
    __device__ __forceinline__ static void run(float2* v) { Dft<R, DIR>::run(v); }
};

template <int R, int DIR> struct DftC<R, DIR, false> {
    static constexpr int F = first_factor(R);
    static constexpr int Q = R / F;
    __device__ __forceinline__ static void run(float2* v)
    {
        float2 y[F][Q];
#pragma unroll
        for (int n1 = 0; n1 < F; ++n1) {
#pragma unroll
            for (int n2 = 0; n2 < Q; ++n2) y[n1][n2] = v[F * n2 + n1];
            DftC<Q, DIR>::run(y[n1]);
        }
#pragma unroll
        for (int n1 = 1; n1 < F; ++n1)
#pragma unroll
            for (int k1 = 1; k1 < Q; ++k1) {
                const int m = (n1 * k1) % R;
                const float c = Roots<R>::t.c[m];
                const float s = DIR < 0 ? -Roots<R>::t.s[m] : Roots<R>::t.s[m];
                y[n1][k1] = cmul(y[n1][k1], make_float2(c, s));
            }
#pragma unroll
        for (int k1 = 0; k1 < Q; ++k1) {
            float2 z[F];
#pragma unroll
            for (int n1 = 0; n1 < F; ++n1) z[n1] = y[n1][k1];
            DftC<F, DIR>::run(z);
#pragma unroll
            for (int k2 = 0; k2 < F; ++k2) v[k1 + Q * k2] = z[k2];
        }
    }
};

// ---------------------------------------------------------------------------
// global-memory pointers
// ---------------------------------------------------------------------------
// Pointers read from the job table are generic (flat) to the compiler, and
// flat loads also count against the LDS wait counter, so every LDS wait
// would wait for outstanding memory loads too.  Casting them to the global
// address space gives global loads with a scalar base.
#define DWHIP_GL __attribute__((address_space(1)))
typedef const DWHIP_GL float2* TwPtr;
template <class T>
__device__ __forceinline__ DWHIP_GL T* gp(T* p)
{
    return (DWHIP_GL T*)p;
}
// float2's constructors take generic references, so complex values move
// through global pointers as native two-float vectors
typedef float dwhip_f2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ldg(const DWHIP_GL float2* p)
{
    const dwhip_f2v t = *(const DWHIP_GL dwhip_f2v*)p;
    return make_float2(t.x, t.y);
}
__device__ __forceinline__ void stg(DWHIP_GL float2* p, float2 v)
{
    dwhip_f2v t;
    t.x = v.x;
    t.y = v.y;
    *(DWHIP_GL dwhip_f2v*)p = t;
}
// Streams: values a pass reads once or writes once (spectra, state, OTF).
// NT adds the nontemporal hint; whether it pays is a property of the pass
// (the callers choose).  Tables every tile reads again (twiddles) use ldg.
template <bool NT>
__device__ __forceinline__ float2 ldgs(const DWHIP_GL float2* p)
{
    if constexpr (NT) {
        const dwhip_f2v t = __builtin_nontemporal_load((const DWHIP_GL dwhip_f2v*)p);
        return make_float2(t.x, t.y);
    } else {
        return ldg(p);
    }
}
template <bool NT>
__device__ __forceinline__ void stgs(DWHIP_GL float2* p, float2 v)
{
    if constexpr (NT) {
        dwhip_f2v t;
        t.x = v.x;
        t.y = v.y;
        __builtin_nontemporal_store(t, (DWHIP_GL dwhip_f2v*)p);
    } else {
        stg(p, v);
    }
}

// ---------------------------------------------------------------------------
// stage geometry
// ---------------------------------------------------------------------------
struct StageGeo {
    int n;        // pencil length
    int npen;     // pencils in the tile (a power of two for column tiles)
    int pstride;  // LDS pencil stride (rows layout)
    int nthr;
    int npshift;  // log2(npen) for column tiles
};

// butterfly (pen, j) of work item w for a stage with nb butterflies per pencil
template <bool COLS>
__device__ __forceinline__ void bfly_of(const StageGeo& g, int w, int nb, uint32_t nbmagic, int& pen, int& j)
{
    if (COLS) {
        j = w >> g.npshift;
        pen = w & (g.npen - 1);
    } else {
        pen = (int)fdiv((uint32_t)w, nbmagic);
        j = w - pen * nb;
    }
}

// LDS indices, strength-reduced: element r of a butterfly sits at
// base + r * step.  All LDS indices are below 2^24 (at most 159 KiB of
// float2), so products use the full-rate 24-bit multiply.
//   in:   elements j + r*nb of butterfly (pen, j), work item w
//   outR: elements j*R + r (first stage, junction)
//   outM: elements (j/ns)*ns*R + j%ns + r*ns (middle stages), jq = j/ns
template <bool COLS>
__device__ __forceinline__ int lds_in(const StageGeo& g, int w, int pen, int j)
{
    return COLS ? w : (int)__umul24((uint32_t)pen, (uint32_t)g.pstride) + j;
}
template <bool COLS>
__device__ __forceinline__ int lds_in_step(const StageGeo& g, int nb)
{
    return COLS ? nb * g.npen : nb;
}
template <int R, bool COLS>
__device__ __forceinline__ int lds_outR(const StageGeo& g, int w, int pen, int j)
{
    return COLS ? w + (int)__umul24((uint32_t)j, (uint32_t)((R - 1) * g.npen))
                : (int)__umul24((uint32_t)pen, (uint32_t)g.pstride) + j * R;
}
template <bool COLS>
__device__ __forceinline__ int lds_outR_step(const StageGeo& g)
{
    return COLS ? g.npen : 1;
}
template <int R, bool COLS>
__device__ __forceinline__ int lds_outM(const StageGeo& g, int w, int pen, int j, int jq, int ns)
{
    return COLS ? w + (int)__umul24((uint32_t)jq, (uint32_t)((R - 1) * ns * g.npen))
                : (int)__umul24((uint32_t)pen, (uint32_t)g.pstride) + j + (int)__umul24((uint32_t)jq, (uint32_t)((R - 1) * ns));
}
template <bool COLS>
__device__ __forceinline__ int lds_outM_step(const StageGeo& g, int ns)
{
    return COLS ? ns * g.npen : ns;
}

// Global operands.  A functor with `static constexpr bool kStrided = true`
// moves the R elements j + r*nb of a butterfly in one call (load<R>,
// store<R>, fetch<R>), so that it can form one address per butterfly and
// step from there; any other functor is called per element.
template <class F, class = void>
struct io_strided {
    static constexpr bool value = false;
};
template <class F>
struct io_strided<F, decltype((void)F::kStrided, void())> {
    static constexpr bool value = F::kStrided;
};
template <int R, class Src>
__device__ __forceinline__ void io_load(const Src& src, int pen, int j, int nb, float2* v)
{
    if constexpr (io_strided<Src>::value) {
        src.template load<R>(pen, j, nb, v);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = src(pen, j + r * nb);
    }
}
template <int R, class Dst>
__device__ __forceinline__ void io_store(const Dst& dst, int pen, int j, int nb, const float2* v)
{
    if constexpr (io_strided<Dst>::value) {
        dst.template store<R>(pen, j, nb, v);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) dst(pen, j + r * nb, v[r]);
    }
}
template <int R, class Op>
__device__ __forceinline__ void io_fetch(const Op& op, int pen, int j, int nb, float2* v)
{
    if constexpr (io_strided<Op>::value) {
        op.template fetch_r<R>(pen, j, nb, v);
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = op.fetch(pen, j + r * nb);
    }
}

template <int R, int DIR>
__device__ __forceinline__ void twiddle(float2* v, TwPtr __restrict__ tw, int k)
{
    if (k == 0) return;
    TwPtr t = tw + k * (R - 1);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = DIR < 0 ? cmul(v[r], ldg(t + r - 1)) : cmulc(v[r], ldg(t + r - 1));
}

// ---------------------------------------------------------------------------
// stage kinds.  Functor interfaces:
//   Src:  float2 operator()(int pen, int i) const         element i of pencil pen
//   Dst:  void operator()(int pen, int i, float2 v) const
//   Op:   void operator()(int pen, int j, int nb, float2* v) const
//         v[r] is element j + r*nb of pencil pen, in natural order
// ---------------------------------------------------------------------------

// first stage (ns = 1): global/generated input -> DFT -> LDS
template <int R, int MB, int DIR, bool COLS, bool SYNC, class Src>
__device__ __forceinline__ void stage_first(const StageGeo& g, int tid, Src src, float2* lds)
{
    const int nb = g.n / R;
    const uint32_t nbm = fdiv_magic((uint32_t)nb);
    const int total = nb * g.npen;
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            io_load<R>(src, pen, j, nb, v[q]);
        }
    }
    if (SYNC) __syncthreads();
    const int ostep = lds_outR_step<COLS>(g);
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            DftC<R, DIR>::run(v[q]);
            float2* o = lds + lds_outR<R, COLS>(g, w, pen, j);
#pragma unroll
            for (int r = 0; r < R; ++r) o[r * ostep] = v[q][r];
        }
    }
}

// middle stage: LDS -> twiddle -> DFT -> LDS (in place, two barriers)
template <int R, int MB, int DIR, bool COLS>
__device__ __forceinline__ void stage_mid(const StageGeo& g, int tid, float2* lds, int ns, uint32_t nsmagic,
                                          TwPtr __restrict__ tw)
{
    const int nb = g.n / R;
    const uint32_t nbm = fdiv_magic((uint32_t)nb);
    const int total = nb * g.npen;
    const int istep = lds_in_step<COLS>(g, nb);
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            const float2* in = lds + lds_in<COLS>(g, w, pen, j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = in[r * istep];
        }
    }
    __syncthreads();
    const int ostep = lds_outM_step<COLS>(g, ns);
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            const int jq = (int)fdiv((uint32_t)j, nsmagic);
            const int k = j - jq * ns;
            twiddle<R, DIR>(v[q], tw, k);
            DftC<R, DIR>::run(v[q]);
            float2* o = lds + lds_outM<R, COLS>(g, w, pen, j, jq, ns);
#pragma unroll
            for (int r = 0; r < R; ++r) o[r * ostep] = v[q][r];
        }
    }
    __syncthreads();
}

// last stage (ns = n/R, so the outputs of butterfly j are j + r*nb):
// LDS -> twiddle -> DFT -> Dst
template <int R, int MB, int DIR, bool COLS, bool SYNC, class Dst>
__device__ __forceinline__ void stage_last(const StageGeo& g, int tid, const float2* lds, TwPtr __restrict__ tw,
                                           Dst dst)
{
    const int nb = g.n / R;
    const uint32_t nbm = fdiv_magic((uint32_t)nb);
    const int total = nb * g.npen;
    const int istep = lds_in_step<COLS>(g, nb);
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            const float2* in = lds + lds_in<COLS>(g, w, pen, j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = in[r * istep];
        }
    }
    if (SYNC) __syncthreads();
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            twiddle<R, DIR>(v[q], tw, j);
            DftC<R, DIR>::run(v[q]);
            io_store<R>(dst, pen, j, nb, v[q]);
        }
    }
}

// junction: LDS -> twiddle -> DFT(DIR1) -> op -> DFT(DIR2) -> LDS at j*R + r
// (last stage of a DIR1 transform met by the first stage of a DIR2 one).
// Op: float2 fetch(int pen, int i) const        operand of element i
//     float2 apply(float2 v, float2 f) const     (element j + r*nb)
template <int R, int MB, int DIR1, int DIR2, bool COLS, bool PF, class Op>
__device__ __forceinline__ void stage_junction(const StageGeo& g, int tid, float2* lds, TwPtr __restrict__ tw,
                                               Op op, const float2* pf)
{
    const int nb = g.n / R;
    const uint32_t nbm = fdiv_magic((uint32_t)nb);
    const int total = nb * g.npen;
    const int istep = lds_in_step<COLS>(g, nb);
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            const float2* in = lds + lds_in<COLS>(g, w, pen, j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = in[r * istep];
        }
    }
    __syncthreads();
    const int ostep = lds_outR_step<COLS>(g);
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            twiddle<R, DIR1>(v[q], tw, j);
            DftC<R, DIR1>::run(v[q]);
            float2 f[R];
            if (PF) {
#pragma unroll
                for (int r = 0; r < R; ++r) f[r] = pf[q * R + r];
            } else {
                io_fetch<R>(op, pen, j, nb, f);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = op.apply(v[q][r], f[r]);
            DftC<R, DIR2>::run(v[q]);
            float2* o = lds + lds_outR<R, COLS>(g, w, pen, j);
#pragma unroll
            for (int r = 0; r < R; ++r) o[r * ostep] = v[q][r];
        }
    }
    __syncthreads();
}

// Operand prefetch for a junction: the values op.fetch returns there, for
// the same butterflies and in the same order, loaded before the first stage
// of the pair so that their latency overlaps the forward stages.
template <int R, int MB, bool COLS, class Op>
__device__ __forceinline__ void stage_prefetch(const StageGeo& g, int tid, const Op& op, float2* pf)
{
    const int nb = g.n / R;
    const uint32_t nbm = fdiv_magic((uint32_t)nb);
    const int total = nb * g.npen;
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            int pen, j;
            bfly_of<COLS>(g, w, nb, nbm, pen, j);
            io_fetch<R>(op, pen, j, nb, pf + q * R);
        }
    }
}

// single-stage transform pair (n == R): Src -> DFT(DIR1) -> op -> DFT(DIR2)
// -> Dst.  With PAIR false only DFT(DIR1) runs.
template <int R, int MB, int DIR1, int DIR2, bool COLS, bool PAIR, bool SYNC, class Src, class Op, class Dst>
__device__ __forceinline__ void stage_single(const StageGeo& g, int tid, Src src, Op op, Dst dst)
{
    const int total = g.npen;  // one butterfly per pencil
    float2 v[MB][R];
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = src(w, r);
        }
    }
    if (SYNC) __syncthreads();
#pragma unroll
    for (int q = 0; q < MB; ++q) {
        const int w = tid + q * g.nthr;
        if (w < total) {
            DftC<R, DIR1>::run(v[q]);
            if (PAIR) {
#pragma unroll
                for (int r = 0; r < R; ++r) v[q][r] = op.apply(v[q][r], op.fetch(w, r));
                DftC<R, DIR2>::run(v[q]);
            }
#pragma unroll
            for (int r = 0; r < R; ++r) dst(w, r, v[q][r]);
        }
    }
}

// no-op for single transforms
struct NoOp {
    __device__ __forceinline__ float2 fetch(int, int) const { return make_float2(0.f, 0.f); }
    __device__ __forceinline__ float2 apply(float2 v, float2) const { return v; }
};

// ---------------------------------------------------------------------------
// radix dispatch
// ---------------------------------------------------------------------------
// Radices the planner may use; each maps to a template instantiation.
// Each axis instantiates the radices up to its own limit RMAX: x-rows up
// to 25, y columns up to 16, z columns up to 7 (the z kernels run at
// eight waves per SIMD, within 64 VGPRs).
#define DWHIP_FOR_RADICES(X) \
    X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(15) X(16) X(20) X(25)
#ifndef DWHIP_RMAX_X
#define DWHIP_RMAX_X 25
#endif
#ifndef DWHIP_RMAX_Y
#define DWHIP_RMAX_Y 16
#endif
#ifndef DWHIP_RMAX_Z
#define DWHIP_RMAX_Z 7
#endif
__host__ __device__ constexpr bool radix_ok(int rmax, int R) { return R <= rmax; }
#define DWHIP_MAX_FFT_STAGES 4

__host__ __device__ constexpr int mb_for(int maxe, int R) { return (maxe + R - 1) / R; }

template <int MAXE, int DIR, bool COLS, int RMAX, bool SYNC, class Src>
__device__ __forceinline__ void run_first(int R, const StageGeo& g, int tid, Src src, float2* lds)
{
    switch (R) {
#define CASE(RR) \
    case RR: if constexpr (radix_ok(RMAX, RR)) { stage_first<RR, mb_for(MAXE, RR), DIR, COLS, SYNC>(g, tid, src, lds); } break;
        DWHIP_FOR_RADICES(CASE)
#undef CASE
    default: break;
    }
}

template <int MAXE, int DIR, bool COLS, int RMAX>
__device__ __forceinline__ void run_mid(int R, const StageGeo& g, int tid, float2* lds, int ns, uint32_t nsmagic,
                                        TwPtr __restrict__ tw)
{
    switch (R) {
#define CASE(RR) \
    case RR: if constexpr (radix_ok(RMAX, RR)) { stage_mid<RR, mb_for(MAXE, RR), DIR, COLS>(g, tid, lds, ns, nsmagic, tw); } break;
        DWHIP_FOR_RADICES(CASE)
#undef CASE
    default: break;
    }
}

template <int MAXE, int DIR, bool COLS, int RMAX, bool SYNC, class Dst>
__device__ __forceinline__ void run_last(int R, const StageGeo& g, int tid, const float2* lds,
                                         TwPtr __restrict__ tw, Dst dst)
{
    switch (R) {
#define CASE(RR) \
    case RR: if constexpr (radix_ok(RMAX, RR)) { stage_last<RR, mb_for(MAXE, RR), DIR, COLS, SYNC>(g, tid, lds, tw, dst); } break;
        DWHIP_FOR_RADICES(CASE)
#undef CASE
    default: break;
    }
}

// Junction operands held in registers between the prefetch and the
// junction: at most DWHIP_PF_MAX per thread; junction radices needing more
// fetch their operands at the junction.  Off by default: the registers
// cost more occupancy than the early loads gain.
#ifndef DWHIP_PF_MAX
#define DWHIP_PF_MAX 0
#endif
__host__ __device__ constexpr bool pf_ok(int maxe, int R) { return mb_for(maxe, R) * R <= DWHIP_PF_MAX; }

template <int MAXE, int DIR1, int DIR2, bool COLS, int RMAX, bool PAIR, bool SYNC, class Src, class Op, class Dst>
__device__ __forceinline__ void run_single(int R, const StageGeo& g, int tid, Src src, Op op, Dst dst)
{
    switch (R) {
#define CASE(RR) \
    case RR: if constexpr (radix_ok(RMAX, RR)) { stage_single<RR, mb_for(MAXE, RR), DIR1, DIR2, COLS, PAIR, SYNC>(g, tid, src, op, dst); } break;
        DWHIP_FOR_RADICES(CASE)
#undef CASE
    default: break;
    }
}

// ---------------------------------------------------------------------------
// whole transforms on a tile
// ---------------------------------------------------------------------------
// Stages 1..nst-2 of plan p (all but the first and the last); nst <= 4.
template <int MAXE, int DIR, bool COLS, int RMAX>
__device__ __forceinline__ void run_mids(const FftPlanDev& p, const StageGeo& g, int tid, float2* lds,
                                         TwPtr __restrict__ tw)
{
    // At most four stages: straight-line code, so that the register
    // allocator sees disjoint stage bodies (a loop around the radix switch
    // makes their live ranges overlap).
    if (p.nst > 2) run_mid<MAXE, DIR, COLS, RMAX>(p.radix[1], g, tid, lds, p.ns[1], p.nsmagic[1], tw + p.twoff[1]);
    if (p.nst > 3) run_mid<MAXE, DIR, COLS, RMAX>(p.radix[2], g, tid, lds, p.ns[2], p.nsmagic[2], tw + p.twoff[2]);
}

// Src -> transform -> Dst.  The LDS must be free on entry.  SRC_LDS: Src
// reads the LDS (a barrier then separates its reads from the first
// stage's writes); DST_LDS: Dst writes the LDS (likewise for the last
// stage).  On return the caller synchronises before reusing the LDS.
template <int MAXE, int DIR, bool COLS, int RMAX, bool SRC_LDS, bool DST_LDS, class Src, class Dst>
__device__ __forceinline__ void fft_tile(const FftPlanDev& p, const StageGeo& g, int tid, float2* lds,
                                         TwPtr __restrict__ tw, Src src, Dst dst)
{
    if (p.nst == 0) {  // n == 1
        float2 v[MAXE];
#pragma unroll
        for (int q = 0; q < MAXE; ++q)
            if (tid + q * g.nthr < g.npen) v[q] = src(tid + q * g.nthr, 0);
        if (SRC_LDS || DST_LDS) __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXE; ++q)
            if (tid + q * g.nthr < g.npen) dst(tid + q * g.nthr, 0, v[q]);
        return;
    }
    if (p.nst == 1) {
        run_single<MAXE, DIR, 0, COLS, RMAX, false, SRC_LDS || DST_LDS>(p.radix[0], g, tid, src, NoOp(), dst);
        return;
    }
    run_first<MAXE, DIR, COLS, RMAX, SRC_LDS>(p.radix[0], g, tid, src, lds);
    __syncthreads();
    run_mids<MAXE, DIR, COLS, RMAX>(p, g, tid, lds, tw);
    run_last<MAXE, DIR, COLS, RMAX, DST_LDS>(p.radix[p.nst - 1], g, tid, lds, tw + p.twoff[p.nst - 1], dst);
}

// Src -> transform (plan a, DIR1) -> op -> transform (plan b, DIR2) -> Dst,
// b's first radix equal to a's last.  Src and Dst address global memory.
template <int MAXE, int DIR1, int DIR2, bool COLS, int RMAX, class Src, class Op, class Dst>
__device__ __forceinline__ void fft_pair_tile(const FftPlanDev& a, const FftPlanDev& b, const StageGeo& g, int tid,
                                              float2* lds, TwPtr __restrict__ tw, Src src, Op op, Dst dst)
{
    if (a.nst == 0) {  // n == 1: the op alone
        for (int w = tid; w < g.npen; w += g.nthr) dst(w, 0, op.apply(src(w, 0), op.fetch(w, 0)));
        return;
    }
    if (a.nst == 1) {
        run_single<MAXE, DIR1, DIR2, COLS, RMAX, true, false>(a.radix[0], g, tid, src, op, dst);
        return;
    }
    // One code path per junction radix, so that the prefetched operands
    // live in registers of a statically known count across the forward
    // stages.
    switch (a.radix[a.nst - 1]) {
#define CASE(RR)                                                                                           \
    case RR:                                                                                               \
        if constexpr (radix_ok(RMAX, RR)) {                                                                \
            constexpr int MB = mb_for(MAXE, RR);                                                           \
            constexpr bool PF = pf_ok(MAXE, RR);                                                           \
            float2 pf[PF ? MB * RR : 1];                                                                   \
            if (PF) stage_prefetch<RR, MB, COLS>(g, tid, op, pf);                                          \
            run_first<MAXE, DIR1, COLS, RMAX, false>(a.radix[0], g, tid, src, lds);                        \
            __syncthreads();                                                                               \
            run_mids<MAXE, DIR1, COLS, RMAX>(a, g, tid, lds, tw);                                          \
            stage_junction<RR, MB, DIR1, DIR2, COLS, PF>(g, tid, lds, tw + a.twoff[a.nst - 1], op, pf);    \
            run_mids<MAXE, DIR2, COLS, RMAX>(b, g, tid, lds, tw);                                          \
            run_last<MAXE, DIR2, COLS, RMAX, false>(b.radix[b.nst - 1], g, tid, lds, tw + b.twoff[b.nst - 1], dst); \
        }                                                                                                  \
        break;
        DWHIP_FOR_RADICES(CASE)
#undef CASE
    default: break;
    }
}
