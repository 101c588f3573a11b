// Born-Wolf widefield PSF on the device: the algorithm of deconwolf's dw_bw
// (src/dw_bwpsf.c:538-775 with src/bw_gsl.c:79-157 and src/lanczos.c:102-160).
//
//   k_bw_profile  |A(r, z)|^2, A = int_0^1 J0(k0 NA r rho) e^{-i k0 NA^2 z rho^2/(2 ni)} rho drho,
//                 at 17 radii per pixel for each plane z <= (P-1)/2, by
//                 Gauss-Legendre quadrature (nodes from the host, per plane)
//   k_bw_pixels   the mean over each pixel of the Lanczos-5 interpolation of
//                 that profile (8 x 8 Gauss-Legendre nodes), zero beyond
//                 radmax; one octant per plane, mirrored eight ways and in z
//
// Double precision throughout; the pixel values are stored as float, as
// dw_bw stores them.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"

#define BTHR 256
#define BW_OVERSAMPLING 17  // radial samples per pixel (dw_bwpsf.c:181)
#define BW_PIXEL_NODES 8

__global__ __launch_bounds__(BTHR) void k_bw_profile(const double* __restrict__ nodes, const int* __restrict__ qoff,
                                                     const int* __restrict__ qn, const double* __restrict__ defocus,
                                                     int nplanes, int nr, double k0NA, double kphase, double step_nm,
                                                     double* __restrict__ prof)
{
    const int i = blockIdx.x * BTHR + threadIdx.x;
    if (i >= nplanes * nr) return;
    const int z = i / nr, n = i - z * nr;
    const double a = k0NA * (double)n * step_nm;   // J0 argument per unit rho
    const double b = -kphase * defocus[z];         // phase per unit rho^2
    const double* x = nodes + 2 * qoff[z];         // (rho, weight) pairs on [0, 1]
    double re = 0, im = 0;
    for (int q = 0; q < qn[z]; ++q) {
        const double rho = x[2 * q], w = x[2 * q + 1];
        const double f = j0(a * rho) * rho * w;
        double s, c;
        sincos(b * rho * rho, &s, &c);
        re += f * c;
        im += f * s;
    }
    prof[i] = re * re + im * im;
}

__device__ __forceinline__ double bw_sinc(double t)
{
    const double x = M_PI * t;
    if (fabs(x) > 1e-6) return sin(x) / x;
    return 1 - x * x / 6 + x * x * x * x / 120;
}

__device__ __forceinline__ double lanczos5_w(double t) { return bw_sinc(t) * bw_sinc(t / 5.0); }

// lanczos5 (lanczos.c:102-160): v symmetric about 0, x >= 0, n + 5 < nv
__device__ __forceinline__ double lanczos5(const double* __restrict__ v, double x)
{
    const int n = (int)floor(x);
    const double d = x - (double)n;
    if (d < 1e-9) return v[n];
    if (d > 1 - 1e-9) return v[n + 1];
    double y = 0;
    for (int k = 0; k < 5; ++k) y += v[abs(n - k)] * lanczos5_w(d + (double)k);
    for (int k = 1; k < 6; ++k) y += v[n + k] * lanczos5_w((double)k - d);
    return y;
}

struct BwPixelNodes {
    double x[BW_PIXEL_NODES];
    double w[BW_PIXEL_NODES];
};

// one thread per (plane, octant pixel x <= y <= N/2)
__global__ __launch_bounds__(BTHR) void k_bw_pixels(const double* __restrict__ prof, int nr, int N, int P, int nplanes,
                                                    int radmax, BwPixelNodes g, float* __restrict__ V)
{
    const int h = N / 2 + 1;  // quadrant side: 2x <= N
    const long long i = (long long)blockIdx.x * BTHR + threadIdx.x;
    if (i >= (long long)nplanes * h * h) return;
    const int z = (int)(i / ((long long)h * h));
    const int rem = (int)(i - (long long)z * h * h);
    const int x = rem / h, y = rem - (rem / h) * h;
    if (y < x) return;
    const double c = ((double)N - 1.0) / 2.0;
    const double* __restrict__ pr = prof + (size_t)z * nr;
    double val = 0;
    const double rpix = sqrt((x - c) * (x - c) + (y - c) * (y - c));
    if (rpix < radmax) {
        // mean over [x - 1/2, x + 1/2] x [y - 1/2, y + 1/2] (dw_bwpsf.c:599-634)
        for (int a = 0; a < BW_PIXEL_NODES; ++a) {
            const double X = (double)x - c + (g.x[a] - 0.5);
            double row = 0;
            for (int b = 0; b < BW_PIXEL_NODES; ++b) {
                const double Y = (double)y - c + (g.x[b] - 0.5);
                const double r2 = X * X + Y * Y;
                const double r = fabs(r2) > 1e-10 ? sqrt(r2) : 0.0;
                row += g.w[b] * lanczos5(pr, r * (double)BW_OVERSAMPLING);
            }
            val += g.w[a] * row;
        }
    }
    const float v = (float)val;
    const int xf = N - x - 1, yf = N - y - 1;
    const size_t plane = (size_t)N * N;
    float* __restrict__ planes[2] = {V + (size_t)z * plane, V + (size_t)(P - 1 - z) * plane};
    for (int k = 0; k < (z == P - 1 - z ? 1 : 2); ++k) {
        float* p = planes[k];
        p[x + (size_t)N * y] = v;
        p[x + (size_t)N * yf] = v;
        p[y + (size_t)N * x] = v;
        p[y + (size_t)N * xf] = v;
        p[xf + (size_t)N * y] = v;
        p[xf + (size_t)N * yf] = v;
        p[yf + (size_t)N * x] = v;
        p[yf + (size_t)N * xf] = v;
    }
}

// n-point Gauss-Legendre rule on [0, 1]: Newton iteration on P_n from the
// asymptotic guesses cos(pi (i - 1/4) / (n + 1/2)), three-term recurrence.
static void gauss_legendre01(int n, std::vector<double>& xw)
{
    const int m = (n + 1) / 2;
    std::vector<double> x(n), w(n);
    for (int i = 0; i < m; ++i) {
        double t = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1;
            dp = n * (t * p1 - p0) / (t * t - 1.0);
            const double dt = p1 / dp;
            t -= dt;
            if (fabs(dt) < 1e-16) break;
        }
        {
            // derivative at the converged node for the weight
            double p0 = 1, p1 = t;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (t * p1 - p0) / (t * t - 1.0);
        }
        const double wt = 2.0 / ((1.0 - t * t) * dp * dp);
        x[i] = -t;
        x[n - 1 - i] = t;
        w[i] = w[n - 1 - i] = wt;
    }
    for (int i = 0; i < n; ++i) {
        xw.push_back(0.5 * (x[i] + 1.0));
        xw.push_back(0.5 * w[i]);
    }
}

extern "C" int dwhip_bw_psf_device(float* out, int64_t N, int64_t P, double lambda_nm, double NA, double ni,
                                   double dxy_nm, double dz_nm, void* stream)
{
    if (!out || N < 1 || P < 1 || N > 4095 || P > 4095 || !(lambda_nm > 0) || !(NA > 0) || !(ni > 0) ||
        !(dxy_nm > 0) || !(dz_nm > 0))
        return DWHIP_ERR_ARG;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const double c = ((double)N - 1.0) / 2.0;
    const int radmax = (int)lround(sqrt(2.0 * ((double)N - c) * ((double)N - c))) + 1;
    const int nr = (radmax + 1) * BW_OVERSAMPLING + 2;
    const int nplanes = (int)((P + 1) / 2);
    const double k0 = 2.0 * M_PI / lambda_nm;
    const double step_nm = dxy_nm / BW_OVERSAMPLING;
    const double rmax_nm = (double)(nr - 1) * step_nm;
    // per-plane quadrature: enough nodes for J0 up to k0 NA rmax and for the
    // defocus phase
    std::vector<double> xw, defocus(nplanes);
    std::vector<int> qoff(nplanes), qn(nplanes);
    for (int z = 0; z < nplanes; ++z) {
        defocus[z] = (double)(float)(dz_nm * ((double)z - ((double)P - 1.0) / 2.0));  // dw_bwpsf.c:549
        const double vmax = k0 * NA * rmax_nm;
        const double wmax = fabs(k0 * NA * NA * defocus[z] / (2.0 * ni));
        qn[z] = (int)(64 + 2.0 * (vmax + wmax) / M_PI);
        qoff[z] = (int)(xw.size() / 2);
        gauss_legendre01(qn[z], xw);
    }
    BwPixelNodes g;
    {
        std::vector<double> p;
        gauss_legendre01(BW_PIXEL_NODES, p);
        for (int a = 0; a < BW_PIXEL_NODES; ++a) {
            g.x[a] = p[2 * a];
            g.w[a] = p[2 * a + 1];
        }
    }
    const size_t bytes = xw.size() * sizeof(double) + 2 * nplanes * sizeof(int) + nplanes * sizeof(double) +
                         (size_t)nplanes * nr * sizeof(double);
    char* mem = nullptr;
    hipError_t e = hipMalloc((void**)&mem, bytes);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? DWHIP_ERR_NOMEM : DWHIP_ERR_HIP;
    double* d_xw = (double*)mem;
    double* d_def = d_xw + xw.size();
    double* d_prof = d_def + nplanes;
    int* d_qoff = (int*)(d_prof + (size_t)nplanes * nr);
    int* d_qn = d_qoff + nplanes;
    e = hipMemcpyAsync(d_xw, xw.data(), xw.size() * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_def, defocus.data(), nplanes * sizeof(double), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_qoff, qoff.data(), nplanes * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_qn, qn.data(), nplanes * sizeof(int), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const int n1 = nplanes * nr;
        hipLaunchKernelGGL(k_bw_profile, dim3((n1 + BTHR - 1) / BTHR), dim3(BTHR), 0, s, d_xw, d_qoff, d_qn, d_def,
                           nplanes, nr, k0 * NA, k0 * NA * NA / (2.0 * ni), step_nm, d_prof);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        const int h = (int)(N / 2 + 1);
        const long long n2 = (long long)nplanes * h * h;
        hipLaunchKernelGGL(k_bw_pixels, dim3((unsigned)((n2 + BTHR - 1) / BTHR)), dim3(BTHR), 0, s, d_prof, nr,
                           (int)N, (int)P, nplanes, radmax, g, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(mem);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "Born-Wolf PSF", __LINE__);
}
