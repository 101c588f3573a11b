// Host side of the any-size path (generic_dev.h, k_generic.hip): plans,
// twiddle and Bluestein tables, the job layout, and the set-up, round and
// final-update sequences that the per-image driver (dwhip.cpp) runs for a
// job whose work size the fused passes cannot transform.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <complex>
#include <mutex>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "generic_dev.h"

extern "C" {
hipError_t dwhip_gen_launch_axis(const GenDev* gd, int axis, int sign, int grid, size_t lds, int threads,
                                 hipStream_t s);
hipError_t dwhip_gen_launch_elem(const GenDev* gd, int op, int grid, hipStream_t s);
hipError_t dwhip_gen_lds_limit(size_t bytes);
hipError_t dwhip_launch_stats(const JobDev* jobs, int j, int nblk, hipStream_t s);
hipError_t dwhip_launch_start(const JobDev* jobs, int j, int grid, hipStream_t s);
hipError_t dwhip_launch_round_end(const BatchDev* bd, const JobDev* jobs, int njobs, int advance, hipStream_t s);
}

namespace {

enum { OP_PSF = 0, OP_OTF, OP_ONES, OP_WEIGHTS, OP_MULK, OP_MULKC, OP_UPDATE, OP_RATIO_IDIV, OP_RATIO_MSE };

// LDS for the two ping-pong buffers of an axis workgroup: at most kGenLds,
// and pencils per workgroup chosen within kGenWgLds (a build tunable, KiB).
// The axis kernels are bound by latency more than by bytes: a long pencil
// (a Bluestein length, 2240 for n = 1115) runs alone in a workgroup of 512
// threads, two workgroups per CU (54 KiB each with the stage twiddles),
// short ones 16 to a workgroup of 256 threads, several per CU.
#ifndef DWHIP_GEN_WG_LDS_KB
#define DWHIP_GEN_WG_LDS_KB 75
#endif
#ifndef DWHIP_GEN_VALS
#define DWHIP_GEN_VALS 4
#endif
constexpr size_t kGenLds = 150 * 1024;
constexpr size_t kGenWgLds = DWHIP_GEN_WG_LDS_KB * 1024;
constexpr int kElemGrid = 4096;    // grid of the elementwise kernels
constexpr int kRatioGrid = 2048;   // fixed grid of the ratio kernel: its partial sums

bool smooth2357(int64_t n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5, 7})
        while (n % f == 0) n /= f;
    return n == 1;
}

// radices of a 2,3,5,7-smooth m: eights first (DWHIP_GEN_R8), then fours,
// then 2, 3, 5, 7
std::vector<int> gen_radices(int m)
{
    std::vector<int> r;
    while (DWHIP_GEN_R8 && m % 8 == 0) {
        r.push_back(8);
        m /= 8;
    }
    while (m % 4 == 0) {
        r.push_back(4);
        m /= 4;
    }
    for (int f : {2, 3, 5, 7})
        while (m % f == 0) {
            r.push_back(f);
            m /= f;
        }
    return r;
}

// Length run in LDS for an axis of length n: n itself when it factors
// into 2, 3, 5, 7, else the smallest such m >= 2n - 1 (Bluestein).
int gen_fft_len(int n, bool* bluestein)
{
    *bluestein = !smooth2357(n);
    if (!*bluestein) return n;
    int m = 2 * n - 1;
    while (!smooth2357(m)) ++m;
    return m;
}

// In-place complex FFT in double (host), sign -1 forward; n 2,3,5,7-smooth.
// Recursive decimation in time over the smallest factor.
void host_fft(std::vector<std::complex<double>>& a, int sign)
{
    const int n = (int)a.size();
    if (n <= 1) return;
    int r = 2;
    while (n % r) ++r;
    const int q = n / r;
    std::vector<std::vector<std::complex<double>>> sub(r, std::vector<std::complex<double>>(q));
    for (int j = 0; j < q; ++j)
        for (int t = 0; t < r; ++t) sub[t][j] = a[j * r + t];
    for (int t = 0; t < r; ++t) host_fft(sub[t], sign);
    for (int k = 0; k < n; ++k) {
        std::complex<double> acc = 0;
        for (int t = 0; t < r; ++t) {
            const double ang = sign * 2.0 * M_PI * (double)((int64_t)t * k % n) / (double)n;
            acc += sub[t][k % q] * std::complex<double>(cos(ang), sin(ang));
        }
        a[k] = acc;
    }
}

std::once_flag g_lds_once;
hipError_t g_lds_err = hipSuccess;

}  // namespace

bool dwhip_generic_size_ok(const int64_t* w)
{
    for (int d = 0; d < 3; ++d) {
        if (w[d] < 1 || w[d] > (1 << 20)) return false;
        bool b;
        const int m = gen_fft_len((int)w[d], &b);
        if (2 * (size_t)m * sizeof(float2) > kGenLds) return false;
        if ((int)gen_radices(m).size() > GEN_MAX_STAGES) return false;
    }
    return true;
}

// Plans the any-size job: per-axis FFT plans and tables (appended to
// j->tw), pencils per workgroup, the layout of the job block.
int dwhip_generic_plan(DwhipJob* j)
{
    GenDev& g = j->g;
    memset(&g, 0, sizeof(g));
    const JobDev& d = j->d;
    g.nx = d.nx; g.ny = d.ny; g.nz = d.nz;
    g.hc = d.hc; g.hs = d.hcs;
    g.M = d.M; g.N = d.N; g.P = d.P;
    g.pM = d.pM; g.pN = d.pN; g.pP = d.pP;
    g.rl = d.rl; g.positivity = d.positivity; g.metric = d.metric; g.has_w = d.has_w;
    g.alphamax = d.alphamax;
    std::vector<float2>& tw = j->tw;
    tw.clear();
    const int len[3] = {d.nx, d.ny, d.nz};
    for (int a = 0; a < 3; ++a) {
        GenPlanDev& p = g.plan[a];
        bool blue = false;
        const int n = len[a], m = gen_fft_len(n, &blue);
        p.n = n;
        p.m = m;
        p.bluestein = blue ? 1 : 0;
        const std::vector<int> r = gen_radices(m);
        if ((int)r.size() > GEN_MAX_STAGES || 2 * (size_t)m * sizeof(float2) > kGenLds) return DWHIP_ERR_SIZE;
        p.nst = (int)r.size();
        int ns = 1;
        for (int s = 0; s < p.nst; ++s) {
            p.radix[s] = r[s];
            p.ns[s] = ns;
            p.twoff[s] = (int)tw.size();
            for (int k = 0; k < ns; ++k)
                for (int q = 1; q < r[s]; ++q) {
                    const double ang = -2.0 * M_PI * (double)(q * k) / (double)(ns * r[s]);
                    tw.push_back(make_float2((float)cos(ang), (float)sin(ang)));
                }
            ns *= r[s];
        }
        p.twlen = p.nst > 0 ? (int)tw.size() - p.twoff[0] : 0;
        // pencils per workgroup: as many as the LDS holds beside the stage
        // twiddles (read there rather than from memory in every butterfly:
        // the stages are latency-bound), at most 16 (adjacent x columns of
        // the y and z passes share their lines); without the twiddles when
        // they do not fit beside one pencil
        const size_t pen = 2 * (size_t)m * sizeof(float2), twb = (size_t)p.twlen * sizeof(float2);
        p.twlds = pen + twb <= kGenWgLds ? 1 : 0;
        const int bmax = (int)((kGenWgLds - (p.twlds ? twb : 0)) / pen);
        p.B = std::max(1, std::min(16, bmax));
        p.chirp = p.hf = p.hi = 0;
        if (blue) {
            // c_j = exp(-i pi j^2 / n), the angle from j^2 mod 2n (exact)
            std::vector<std::complex<double>> c(n);
            for (int jj = 0; jj < n; ++jj) {
                const double ang = -M_PI * (double)(((int64_t)jj * jj) % (2 * (int64_t)n)) / (double)n;
                c[jj] = std::complex<double>(cos(ang), sin(ang));
            }
            p.chirp = (int)tw.size();
            for (int jj = 0; jj < n; ++jj) tw.push_back(make_float2((float)c[jj].real(), (float)c[jj].imag()));
            // H- = FFT_m(conj(c) / m) and H+ = FFT_m(c / m), c circularly
            // extended to length m (c_{m-l} = c_l, l < n)
            for (int side = 0; side < 2; ++side) {
                std::vector<std::complex<double>> h(m, 0.0);
                for (int l = 0; l < n; ++l) {
                    const std::complex<double> v = (side == 0 ? std::conj(c[l]) : c[l]) / (double)m;
                    h[l] = v;
                    if (l > 0) h[m - l] = v;
                }
                host_fft(h, -1);
                (side == 0 ? p.hf : p.hi) = (int)tw.size();
                for (int k = 0; k < m; ++k) tw.push_back(make_float2((float)h[k].real(), (float)h[k].imag()));
            }
        }
    }
    return DWHIP_OK;
}

size_t dwhip_generic_lds(const DwhipJob* j, int axis)
{
    const GenPlanDev& p = j->g.plan[axis];
    return 2 * (size_t)p.B * p.m * sizeof(float2) + (p.twlds ? (size_t)p.twlen * sizeof(float2) : 0);
}

int dwhip_generic_ratio_grid() { return kRatioGrid; }

// Points the GenDev at the job's buffers (after dwhip_job_bind has laid
// them out) and queues its upload.
int dwhip_generic_bind(DwhipJob* j, char* gdev, hipStream_t s)
{
    std::call_once(g_lds_once, [] { g_lds_err = dwhip_gen_lds_limit(kGenLds); });
    if (g_lds_err != hipSuccess) return dwhip_hip_fail(g_lds_err, "LDS limit of the any-size kernels", __LINE__);
    GenDev& g = j->g;
    const JobDev& d = j->d;
    g.Z = d.C;
    g.K = d.K;
    g.U0 = d.U0;
    g.U1 = d.U1;
    g.W = d.W;
    g.im = d.im;
    g.psf = d.psf;
    g.part = d.part;
    g.st = d.st;
    g.tw = d.tw;
    j->gd = (GenDev*)gdev;
    const hipError_t e = hipMemcpyAsync(gdev, &g, sizeof(GenDev), hipMemcpyHostToDevice, s);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "any-size job upload", __LINE__);
}

namespace {

// tiles of an axis pass: B row pairs along x, B bins of a plane (y) or a
// row (z)
int ntiles_axis(const GenDev& g, int a)
{
    const int B = g.plan[a].B;
    if (a == 0) return ((g.ny * g.nz + 1) / 2 + B - 1) / B;
    const int nxb = (g.hc + B - 1) / B;
    return nxb * (a == 1 ? g.nz : g.ny);
}

// a: axis 0-2, or 3 / 4 for the z axis with the K / conj(K) junction
int axis(const DwhipJob* j, int a, int sign, hipStream_t s)
{
    const int ax = std::min(a, 2);
    const int grid = std::min(ntiles_axis(j->g, ax), 65536);
    // threads for about DWHIP_GEN_VALS values each of the tile's B*m (256
    // to 1024; a build tunable)
    const int vals = j->g.plan[ax].B * j->g.plan[ax].m;
    const int threads = vals >= DWHIP_GEN_VALS * 1024 ? 1024 : vals >= DWHIP_GEN_VALS * 512 ? 512 : 256;
    const hipError_t e = dwhip_gen_launch_axis(j->gd, a, sign, grid, dwhip_generic_lds(j, ax), threads, s);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "any-size axis pass", __LINE__);
}

int elem(const DwhipJob* j, int op, hipStream_t s)
{
    const int grid = (op == OP_RATIO_IDIV || op == OP_RATIO_MSE) ? kRatioGrid : kElemGrid;
    const hipError_t e = dwhip_gen_launch_elem(j->gd, op, grid, s);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "any-size elementwise pass", __LINE__);
}

#define GCHK(call)                     \
    do {                               \
        int r_ = (call);               \
        if (r_ != DWHIP_OK) return r_; \
    } while (0)

// Z <- IFFT(op(K) FFT(Z)), unnormalised (1/(nx ny nz) is in K); the z
// transforms and the K multiply in one pass
int convolve(const DwhipJob* j, bool conj, hipStream_t s)
{
    GCHK(axis(j, 0, -1, s));
    GCHK(axis(j, 1, -1, s));
    GCHK(axis(j, conj ? 4 : 3, 0, s));
    GCHK(axis(j, 1, +1, s));
    GCHK(axis(j, 0, +1, s));
    return DWHIP_OK;
}

}  // namespace

// OTF, Bertero weights and start guess (method_shb.c:108-294).
int dwhip_generic_setup(const DwhipJob* j, hipStream_t s)
{
    hipError_t e = dwhip_launch_stats(j->jd, 0, DWHIP_STAT_BLOCKS, s);
    if (e == hipSuccess) e = dwhip_launch_start(j->jd, 0, 1024, s);
    if (e != hipSuccess) return dwhip_hip_fail(e, "any-size set-up", __LINE__);
    GCHK(elem(j, OP_PSF, s));
    for (int a = 0; a < 3; ++a) GCHK(axis(j, a, -1, s));
    GCHK(elem(j, OP_OTF, s));
    if (j->d.has_w) {
        GCHK(elem(j, OP_ONES, s));
        GCHK(convolve(j, true, s));
        GCHK(elem(j, OP_WEIGHTS, s));
    }
    return DWHIP_OK;
}

// One iteration: update (finishing the previous one), y = H p, error and
// ratio, and with `full` the adjoint IFFT(conj(K) FFT(ratio)).
int dwhip_generic_round(DwhipJob* j, bool full, hipStream_t s)
{
    GCHK(elem(j, OP_UPDATE, s));
    GCHK(convolve(j, false, s));
    GCHK(elem(j, j->d.metric == DWHIP_DEV_METRIC_IDIV ? OP_RATIO_IDIV : OP_RATIO_MSE, s));
    if (full) GCHK(convolve(j, true, s));
    const hipError_t e = dwhip_launch_round_end(j->bd, j->jd, 1, 1, s);
    if (e != hipSuccess) return dwhip_hip_fail(e, "round end", __LINE__);
    return DWHIP_OK;
}

// RL's x_n after its last round (the update step alone).
int dwhip_generic_final_update(DwhipJob* j, hipStream_t s) { return elem(j, OP_UPDATE, s); }
