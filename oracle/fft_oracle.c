/*
 * CPU FFT used by the SHB oracle.  TEST INFRASTRUCTURE ONLY: this file
 * is the checker, never the product.  Only tests/, the smoke check and
 * the CPU-baseline leg of the benchmark may load it.
 *
 * It replaces the FFTW3-float calls of deconwolf's src/fft.c
 * (fftwf_plan_dft_r2c_3d / fftwf_plan_dft_c2r_3d, fft.c:236-290, executed at
 * fft.c:412,844,853).  Conventions follow FFTW's:
 *   - forward transform uses exp(-2 pi i j k / n), inverse exp(+2 pi i ...),
 *     both unnormalised (deconwolf divides by wM*wN*wP afterwards,
 *     fft.c:856-861);
 *   - the real-to-complex layout keeps (nx/2+1) complex values per x-row,
 *     x fastest (fft.c:129-133 `nch`).
 *
 * All arithmetic is carried out in double precision.  Complex FFTs of any
 * length: mixed radix Stockham passes for small prime factors, Bluestein's
 * chirp-z algorithm (L. Bluestein, 1970) when a prime factor exceeds 61.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fft_oracle.h"

#define MAXSTAGE 48

struct cfft_plan {
    int n;
    int nstage;
    int radix[MAXSTAGE];
    int ns[MAXSTAGE];          /* product of the radices of earlier stages */
    double *tw[MAXSTAGE];      /* ns*(R-1) complex twiddles, forward sign */
    double *root[MAXSTAGE];    /* R complex roots exp(-2 pi i q / R) */
    /* Bluestein */
    int blue;
    int L;
    struct cfft_plan *sub;
    double *chirp;             /* n complex, exp(+i pi j^2 / n) */
    double *bfft;              /* L complex, FFT of the extended chirp */
};

static void *xmalloc(size_t nbytes)
{
    void *p = NULL;
    if (nbytes == 0)
        nbytes = 16;
    if (posix_memalign(&p, 64, nbytes) != 0)
        abort();
    return p;
}

static int smallest_factor(int n)
{
    if (n % 4 == 0 && n != 8)
        return 4;
    for (int f = 2; f * f <= n; f++)
        if (n % f == 0)
            return f;
    return n;
}

static int max_prime_factor(int n)
{
    int m = 1;
    for (int f = 2; f * f <= n; f++)
        while (n % f == 0) {
            m = f > m ? f : m;
            n /= f;
        }
    return n > m ? n : m;
}

cfft_plan *oracle_cfft_plan(int n)
{
    cfft_plan *p = calloc(1, sizeof(cfft_plan));
    p->n = n;
    if (n > 1 && max_prime_factor(n) > 61) {
        /* Bluestein: X_k = conj(b_k) sum_j (x_j conj(b_j)) b_{k-j},
         * b_j = exp(i pi j^2 / n). */
        int L = 1;
        while (L < 2 * n - 1)
            L <<= 1;
        p->blue = 1;
        p->L = L;
        p->sub = oracle_cfft_plan(L);
        p->chirp = xmalloc(sizeof(double) * 2 * n);
        for (int64_t j = 0; j < n; j++) {
            int64_t q = (j * j) % (2 * (int64_t)n);
            double a = M_PI * (double)q / (double)n;
            p->chirp[2 * j] = cos(a);
            p->chirp[2 * j + 1] = sin(a);
        }
        p->bfft = xmalloc(sizeof(double) * 2 * L);
        memset(p->bfft, 0, sizeof(double) * 2 * L);
        for (int j = 0; j < n; j++) {
            p->bfft[2 * j] = p->chirp[2 * j];
            p->bfft[2 * j + 1] = p->chirp[2 * j + 1];
            if (j > 0) {
                p->bfft[2 * (L - j)] = p->chirp[2 * j];
                p->bfft[2 * (L - j) + 1] = p->chirp[2 * j + 1];
            }
        }
        double *work = xmalloc(sizeof(double) * 2 * L);
        oracle_cfft_exec(p->sub, p->bfft, work, -1);
        free(work);
        return p;
    }
    int rem = n, ns = 1, s = 0;
    while (rem > 1) {
        int R = smallest_factor(rem);
        p->radix[s] = R;
        p->ns[s] = ns;
        p->tw[s] = xmalloc(sizeof(double) * 2 * (size_t)ns * (R - 1) + 16);
        for (int k = 0; k < ns; k++)
            for (int r = 1; r < R; r++) {
                double a = -2.0 * M_PI * (double)(r * k) / (double)(ns * R);
                p->tw[s][2 * (k * (R - 1) + r - 1)] = cos(a);
                p->tw[s][2 * (k * (R - 1) + r - 1) + 1] = sin(a);
            }
        p->root[s] = xmalloc(sizeof(double) * 2 * R);
        for (int q = 0; q < R; q++) {
            double a = -2.0 * M_PI * (double)q / (double)R;
            p->root[s][2 * q] = cos(a);
            p->root[s][2 * q + 1] = sin(a);
        }
        ns *= R;
        rem /= R;
        s++;
    }
    p->nstage = s;
    return p;
}

void oracle_cfft_free(cfft_plan *p)
{
    if (!p)
        return;
    for (int s = 0; s < p->nstage; s++) {
        free(p->tw[s]);
        free(p->root[s]);
    }
    if (p->blue) {
        oracle_cfft_free(p->sub);
        free(p->chirp);
        free(p->bfft);
    }
    free(p);
}

size_t oracle_cfft_work_size(const cfft_plan *p)
{
    /* complex elements of scratch needed by oracle_cfft_exec */
    if (p->blue)
        return (size_t)p->L * 2 + (size_t)p->n;
    return (size_t)p->n;
}

/* One radix-R butterfly on v[0..R-1] (interleaved complex), sign s. */
static inline void dft_small(double *v, int R, const double *root, double s)
{
    if (R == 2) {
        double ar = v[0], ai = v[1], br = v[2], bi = v[3];
        v[0] = ar + br; v[1] = ai + bi;
        v[2] = ar - br; v[3] = ai - bi;
        return;
    }
    if (R == 4) {
        double x0r = v[0], x0i = v[1], x1r = v[2], x1i = v[3];
        double x2r = v[4], x2i = v[5], x3r = v[6], x3i = v[7];
        double t0r = x0r + x2r, t0i = x0i + x2i;
        double t1r = x0r - x2r, t1i = x0i - x2i;
        double t2r = x1r + x3r, t2i = x1i + x3i;
        double t3r = x1r - x3r, t3i = x1i - x3i;
        /* multiply t3 by -i*s ... forward (s=-1): -i */
        double ur = s < 0 ? t3i : -t3i;
        double ui = s < 0 ? -t3r : t3r;
        v[0] = t0r + t2r; v[1] = t0i + t2i;
        v[2] = t1r + ur;  v[3] = t1i + ui;
        v[4] = t0r - t2r; v[5] = t0i - t2i;
        v[6] = t1r - ur;  v[7] = t1i - ui;
        return;
    }
    double out[2 * 64];
    for (int m = 0; m < R; m++) {
        double sr = 0, si = 0;
        for (int q = 0; q < R; q++) {
            int e = (q * m) % R;
            double wr = root[2 * e], wi = s < 0 ? root[2 * e + 1] : -root[2 * e + 1];
            sr += v[2 * q] * wr - v[2 * q + 1] * wi;
            si += v[2 * q] * wi + v[2 * q + 1] * wr;
        }
        out[2 * m] = sr;
        out[2 * m + 1] = si;
    }
    memcpy(v, out, sizeof(double) * 2 * R);
}

void oracle_cfft_exec(const cfft_plan *p, double *data, double *work, int sign)
{
    const int n = p->n;
    if (n <= 1)
        return;
    const double s = sign < 0 ? -1.0 : 1.0;
    if (p->blue) {
        const int L = p->L;
        double *a = work;                 /* L complex */
        double *w2 = work + 2 * (size_t)L; /* L complex scratch for sub */
        memset(a, 0, sizeof(double) * 2 * L);
        for (int j = 0; j < n; j++) {
            double xr = data[2 * j], xi = s < 0 ? data[2 * j + 1] : -data[2 * j + 1];
            double cr = p->chirp[2 * j], ci = -p->chirp[2 * j + 1];
            a[2 * j] = xr * cr - xi * ci;
            a[2 * j + 1] = xr * ci + xi * cr;
        }
        oracle_cfft_exec(p->sub, a, w2, -1);
        for (int k = 0; k < L; k++) {
            double ar = a[2 * k], ai = a[2 * k + 1];
            double br = p->bfft[2 * k], bi = p->bfft[2 * k + 1];
            a[2 * k] = ar * br - ai * bi;
            a[2 * k + 1] = ar * bi + ai * br;
        }
        oracle_cfft_exec(p->sub, a, w2, +1);
        for (int k = 0; k < n; k++) {
            double ar = a[2 * k] / L, ai = a[2 * k + 1] / L;
            double cr = p->chirp[2 * k], ci = -p->chirp[2 * k + 1];
            double yr = ar * cr - ai * ci;
            double yi = ar * ci + ai * cr;
            data[2 * k] = yr;
            data[2 * k + 1] = s < 0 ? yi : -yi;
        }
        return;
    }
    double *in = data, *out = work;
    for (int st = 0; st < p->nstage; st++) {
        const int R = p->radix[st], ns = p->ns[st], nr = n / R;
        const double *tw = p->tw[st];
        double v[2 * 64];
        for (int j = 0; j < nr; j++) {
            const int k = j % ns;
            for (int r = 0; r < R; r++) {
                double xr = in[2 * (j + r * nr)], xi = in[2 * (j + r * nr) + 1];
                if (r > 0 && k > 0) {
                    double wr = tw[2 * (k * (R - 1) + r - 1)];
                    double wi = tw[2 * (k * (R - 1) + r - 1) + 1] * (s < 0 ? 1.0 : -1.0);
                    double tr = xr * wr - xi * wi;
                    xi = xr * wi + xi * wr;
                    xr = tr;
                }
                v[2 * r] = xr;
                v[2 * r + 1] = xi;
            }
            dft_small(v, R, p->root[st], s);
            const int base = (j / ns) * ns * R + k;
            for (int r = 0; r < R; r++) {
                out[2 * (base + r * ns)] = v[2 * r];
                out[2 * (base + r * ns) + 1] = v[2 * r + 1];
            }
        }
        double *t = in;
        in = out;
        out = t;
    }
    if (in != data)
        memcpy(data, in, sizeof(double) * 2 * n);
}

void oracle_cfft(double *data, int n, int sign)
{
    cfft_plan *p = oracle_cfft_plan(n);
    double *w = xmalloc(sizeof(double) * 2 * oracle_cfft_work_size(p));
    oracle_cfft_exec(p, data, w, sign);
    free(w);
    oracle_cfft_free(p);
}

/*
 * 3-D real-to-complex forward transform, like fftwf_plan_dft_r2c_3d(P, N, M)
 * on an x-fastest array: out has (nx/2+1)*ny*nz complex values.
 */
void oracle_rfft3(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out)
{
    const int64_t hx = nx / 2 + 1;
    cfft_plan *px = oracle_cfft_plan((int)nx);
    cfft_plan *py = oracle_cfft_plan((int)ny);
    cfft_plan *pz = oracle_cfft_plan((int)nz);
    size_t wmax = oracle_cfft_work_size(px);
    if (oracle_cfft_work_size(py) > wmax) wmax = oracle_cfft_work_size(py);
    if (oracle_cfft_work_size(pz) > wmax) wmax = oracle_cfft_work_size(pz);
    int64_t lmax = nx > ny ? nx : ny;
    lmax = lmax > nz ? lmax : nz;

#pragma omp parallel
    {
        double *buf = xmalloc(sizeof(double) * 2 * lmax);
        double *w = xmalloc(sizeof(double) * 2 * wmax);
#pragma omp for schedule(static)
        for (int64_t r = 0; r < ny * nz; r++) {
            const double *row = in + r * nx;
            for (int64_t i = 0; i < nx; i++) {
                buf[2 * i] = row[i];
                buf[2 * i + 1] = 0;
            }
            oracle_cfft_exec(px, buf, w, -1);
            memcpy(out + 2 * r * hx, buf, sizeof(double) * 2 * hx);
        }
#pragma omp for schedule(static)
        for (int64_t c = 0; c < hx * nz; c++) {
            const int64_t kx = c % hx, z = c / hx;
            double *base = out + 2 * (kx + z * hx * ny);
            for (int64_t y = 0; y < ny; y++) {
                buf[2 * y] = base[2 * y * hx];
                buf[2 * y + 1] = base[2 * y * hx + 1];
            }
            oracle_cfft_exec(py, buf, w, -1);
            for (int64_t y = 0; y < ny; y++) {
                base[2 * y * hx] = buf[2 * y];
                base[2 * y * hx + 1] = buf[2 * y + 1];
            }
        }
#pragma omp for schedule(static)
        for (int64_t c = 0; c < hx * ny; c++) {
            double *base = out + 2 * c;
            const int64_t st = hx * ny;
            for (int64_t z = 0; z < nz; z++) {
                buf[2 * z] = base[2 * z * st];
                buf[2 * z + 1] = base[2 * z * st + 1];
            }
            oracle_cfft_exec(pz, buf, w, -1);
            for (int64_t z = 0; z < nz; z++) {
                base[2 * z * st] = buf[2 * z];
                base[2 * z * st + 1] = buf[2 * z + 1];
            }
        }
        free(buf);
        free(w);
    }
    oracle_cfft_free(px);
    oracle_cfft_free(py);
    oracle_cfft_free(pz);
}

/*
 * 3-D complex-to-real inverse transform, unnormalised, like
 * fftwf_plan_dft_c2r_3d(P, N, M).  `in` is left untouched.  The x-axis
 * step builds the Hermitian extension X[nx-k] = conj(X[k]); the imaginary
 * parts of the self-conjugate bins are dropped by taking the real part.
 */
void oracle_irfft3(const double *in, int64_t nx, int64_t ny, int64_t nz, double *out)
{
    const int64_t hx = nx / 2 + 1;
    const size_t nc = (size_t)hx * ny * nz;
    double *tmp = xmalloc(sizeof(double) * 2 * nc);
    memcpy(tmp, in, sizeof(double) * 2 * nc);
    cfft_plan *px = oracle_cfft_plan((int)nx);
    cfft_plan *py = oracle_cfft_plan((int)ny);
    cfft_plan *pz = oracle_cfft_plan((int)nz);
    size_t wmax = oracle_cfft_work_size(px);
    if (oracle_cfft_work_size(py) > wmax) wmax = oracle_cfft_work_size(py);
    if (oracle_cfft_work_size(pz) > wmax) wmax = oracle_cfft_work_size(pz);
    int64_t lmax = nx > ny ? nx : ny;
    lmax = lmax > nz ? lmax : nz;

#pragma omp parallel
    {
        double *buf = xmalloc(sizeof(double) * 2 * lmax);
        double *w = xmalloc(sizeof(double) * 2 * wmax);
#pragma omp for schedule(static)
        for (int64_t c = 0; c < hx * ny; c++) {
            double *base = tmp + 2 * c;
            const int64_t st = hx * ny;
            for (int64_t z = 0; z < nz; z++) {
                buf[2 * z] = base[2 * z * st];
                buf[2 * z + 1] = base[2 * z * st + 1];
            }
            oracle_cfft_exec(pz, buf, w, +1);
            for (int64_t z = 0; z < nz; z++) {
                base[2 * z * st] = buf[2 * z];
                base[2 * z * st + 1] = buf[2 * z + 1];
            }
        }
#pragma omp for schedule(static)
        for (int64_t c = 0; c < hx * nz; c++) {
            const int64_t kx = c % hx, z = c / hx;
            double *base = tmp + 2 * (kx + z * hx * ny);
            for (int64_t y = 0; y < ny; y++) {
                buf[2 * y] = base[2 * y * hx];
                buf[2 * y + 1] = base[2 * y * hx + 1];
            }
            oracle_cfft_exec(py, buf, w, +1);
            for (int64_t y = 0; y < ny; y++) {
                base[2 * y * hx] = buf[2 * y];
                base[2 * y * hx + 1] = buf[2 * y + 1];
            }
        }
#pragma omp for schedule(static)
        for (int64_t r = 0; r < ny * nz; r++) {
            const double *row = tmp + 2 * r * hx;
            for (int64_t k = 0; k < hx; k++) {
                buf[2 * k] = row[2 * k];
                buf[2 * k + 1] = row[2 * k + 1];
            }
            for (int64_t k = hx; k < nx; k++) {
                buf[2 * k] = row[2 * (nx - k)];
                buf[2 * k + 1] = -row[2 * (nx - k) + 1];
            }
            oracle_cfft_exec(px, buf, w, +1);
            for (int64_t i = 0; i < nx; i++)
                out[r * nx + i] = buf[2 * i];
        }
        free(buf);
        free(w);
    }
    free(tmp);
    oracle_cfft_free(px);
    oracle_cfft_free(py);
    oracle_cfft_free(pz);
}
