// Host side of libdwhip: transform plans, job set-up and the per-image SHB
// driver behind the C-ABI in include/dwhip.h.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "shb_kernels.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
const char* dwhip_strerror(int code)
{
    switch (code) {
    case DWHIP_OK: return "ok";
    case DWHIP_ERR_ARG: return "invalid argument";
    case DWHIP_ERR_SIZE: return "transform size beyond the kernels' limits";
    case DWHIP_ERR_HIP: return "HIP runtime error";
    case DWHIP_ERR_NOMEM: return "device memory allocation failed";
    case DWHIP_ERR_PSF: return "the PSF maximum could not be moved to the origin";
    case DWHIP_ERR_NODEVICE: return "no usable GPU device";
    case DWHIP_ERR_STATE: return "call not valid in this state";
    case DWHIP_ERR_SCRATCH: return "job larger than the library's static scratch block (DWHIP_SCRATCH_BYTES)";
    default: return "unknown error";
    }
}

int dwhip_hip_fail(hipError_t e, const char* what, int line)
{
    fprintf(stderr, "dwhip: %s failed (%s) at dwhip.cpp:%d\n", what, hipGetErrorString(e), line);
    return e == hipErrorOutOfMemory ? DWHIP_ERR_NOMEM : DWHIP_ERR_HIP;
}

static int g_init = 0;

int dwhip_init(void)
{
    if (g_init == 1) return DWHIP_OK;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return DWHIP_ERR_NODEVICE;
    if (hipSetDevice(0) != hipSuccess) return DWHIP_ERR_NODEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return DWHIP_ERR_NODEVICE;
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        fprintf(stderr, "dwhip: device is %s, libdwhip is compiled for gfx950\n", prop.gcnArchName);
        return DWHIP_ERR_NODEVICE;
    }
    g_init = 1;
    return DWHIP_OK;
}

void dwhip_params_default(dwhip_params* p, int niter)
{
    memset(p, 0, sizeof(*p));
    // dw_opts_new, src/dw.c:340-413
    p->niter = niter;
    p->iter_type = DWHIP_ITER_FIXED;
    p->maxiter = 250;
    p->err_rel = 0.02f;
    p->err_abs = 1.0f;
    p->bg = 1e-2f;
    p->bg_auto = 1;
    p->border_quality = 2;
    p->positivity = 1;
    p->alphamax = 1.0f;
    p->metric = DWHIP_METRIC_IDIV;
    p->start = DWHIP_START_FLAT;
    p->method = DWHIP_METHOD_SHB;
    p->planning = DWHIP_PLAN_MEASURE;  // FFTW_MEASURE, src/dw.c:400
}

// ---------------------------------------------------------------------------
// sizes and plans
// ---------------------------------------------------------------------------
static bool is_smooth(int64_t n)
{
    if (n < 1) return false;
    for (int f : {2, 3, 5, 7})
        while (n % f == 0) n /= f;
    return n == 1;
}

std::vector<int> dwhip_radices(int n, int axis);

// smallest length >= n with a plan along axis 1 (y) or 2 (z)
static int64_t next_smooth(int64_t n, int axis)
{
    while (!(is_smooth(n) && (n == 1 || !dwhip_radices((int)n, axis).empty()))) ++n;
    return n;
}

// smallest even length >= n whose half has a row plan
static int64_t next_smooth_even(int64_t n)
{
    int64_t h = (n + 1) / 2;
    while (!(is_smooth(h) && (h == 1 || !dwhip_radices((int)h, 0).empty()))) ++h;
    return 2 * h;
}

static void ref_work_size(int bq, int64_t M, int64_t N, int64_t P, int64_t pM, int64_t pN, int64_t pP,
                          int64_t* w)
{
    // method_shb.c:137-153
    w[0] = M + pM - 1;
    w[1] = N + pN - 1;
    w[2] = P + pP - 1;
    if (bq == 1) {
        w[0] = M + (pM + 1) / 2;
        w[1] = N + (pN + 1) / 2;
        w[2] = P + (pP + 1) / 2;
    }
    if (bq == 0) {
        w[0] = std::max(M, pM);
        w[1] = std::max(N, pN);
        w[2] = std::max(P, pP);
    }
}

// whether the fused passes transform a work size w
bool dwhip_fast_size_ok(const int64_t* w)
{
    if (w[0] % 2 != 0 || !is_smooth(w[0] / 2) || !is_smooth(w[1]) || !is_smooth(w[2])) return false;
    // kernel limits: at most DWHIP_XMAXE*NTHR complex per x-row tile,
    // DWHIP_YMAXE / DWHIP_ZMAXE * NTHR per column tile
    if (w[0] / 2 > DWHIP_XMAXE * DWHIP_XTHR) return false;
    if (w[1] > DWHIP_YMAXE * DWHIP_CTHR || w[2] > DWHIP_ZMAXE * DWHIP_ZTHR) return false;
    // and one pencil per column tile must fit in LDS
    if (std::max(w[1], w[2]) * (int64_t)sizeof(float2) > DWHIP_MAX_LDS_BYTES) return false;
    if (w[0] < 2 || w[1] < 1 || w[2] < 1) return false;
    if (w[0] / 2 > 1 && dwhip_radices((int)(w[0] / 2), 0).empty()) return false;
    if (w[1] > 1 && dwhip_radices((int)w[1], 1).empty()) return false;
    if (w[2] > 1 && dwhip_radices((int)w[2], 2).empty()) return false;
    return true;
}

int dwhip_work_size(int64_t M, int64_t N, int64_t P, int64_t pM, int64_t pN, int64_t pP,
                    int border_quality, int64_t* work, int64_t* ref_work)
{
    if (M < 1 || N < 1 || P < 1 || pM < 1 || pN < 1 || pP < 1 || border_quality < 0 || border_quality > 2)
        return DWHIP_ERR_ARG;
    int64_t r[3];
    ref_work_size(border_quality, M, N, P, pM, pN, pP, r);
    if (ref_work)
        for (int d = 0; d < 3; ++d) ref_work[d] = r[d];
    int64_t w[3];
    if (border_quality == 2) {
        w[0] = next_smooth_even(r[0]);
        w[1] = next_smooth(r[1], 1);
        w[2] = next_smooth(r[2], 2);
    } else {
        for (int d = 0; d < 3; ++d) w[d] = r[d];
    }
    if (work)
        for (int d = 0; d < 3; ++d) work[d] = w[d];
    // the fused passes, or else the any-size path (any factors, odd x)
    return dwhip_fast_size_ok(w) || dwhip_generic_size_ok(w) ? DWHIP_OK : DWHIP_ERR_SIZE;
}

// Radices for a 2,3,5,7-smooth n: fewest stages, then the smallest largest
// radix (fewer registers per butterfly), from the radices the kernels
// instantiate (DWHIP_FOR_RADICES in fft_engine.h).
static const int kRadices[] = {25, 20, 16, 15, 9, 8, 7, 6, 5, 4, 3, 2};
static int axis_rmax(int axis)
{
    return axis == 0 ? DWHIP_RMAX_X : axis == 1 ? DWHIP_RMAX_Y : DWHIP_RMAX_Z;
}

static void radix_search(int n, int rmax, std::vector<int>& cur, std::vector<int>& best)
{
    if (n == 1) {
        if (best.empty() || cur.size() < best.size() ||
            (cur.size() == best.size() && *std::max_element(cur.begin(), cur.end()) <
                                              *std::max_element(best.begin(), best.end())))
            best = cur;
        return;
    }
    if (cur.size() >= DWHIP_MAX_FFT_STAGES) return;
    if (!best.empty() && cur.size() + 1 > best.size()) return;
    const int nr = (int)(sizeof(kRadices) / sizeof(int));
    for (int i = 0; i < nr; ++i) {
        const int r = kRadices[i];
        if (r > rmax) continue;
        // non-increasing radices: each factorisation is visited once
        if (!cur.empty() && r > cur.back()) continue;
        if (n % r != 0) continue;
        cur.push_back(r);
        radix_search(n / r, rmax, cur, best);
        cur.pop_back();
    }
}

// Radices for a 2,3,5,7-smooth n along axis 0 (x rows), 1 (y) or 2 (z), at
// most DWHIP_MAX_FFT_STAGES of them; empty when n has no such plan.
std::vector<int> dwhip_radices(int n, int axis)
{
    std::vector<int> cur, best;
    if (n > 1) radix_search(n, axis_rmax(axis), cur, best);
    return best;
}

// Builds a plan with the given radix order and appends its twiddles
// (forward sign) to tw, offsets relative to tw's start.
int dwhip_make_plan(int n, const std::vector<int>& r, FftPlanDev* p, std::vector<float2>& tw)
{
    memset(p, 0, sizeof(*p));
    p->n = n;
    if ((int)r.size() > DWHIP_MAX_STAGES) return DWHIP_ERR_SIZE;
    p->nst = (int)r.size();
    int ns = 1;
    for (int s = 0; s < p->nst; ++s) {
        const int R = r[s];
        p->radix[s] = R;
        p->ns[s] = ns;
        p->nsmagic[s] = fdiv_magic((uint32_t)ns);
        p->twoff[s] = (int)tw.size();
        for (int k = 0; k < ns; ++k)
            for (int q = 1; q < R; ++q) {
                const double a = -2.0 * M_PI * (double)(q * k) / (double)(ns * R);
                tw.push_back(make_float2((float)cos(a), (float)sin(a)));
            }
        ns *= R;
    }
    return ns == n || n == 1 ? DWHIP_OK : DWHIP_ERR_SIZE;
}

// All radix lists for n along an axis, at most DWHIP_MAX_FFT_STAGES stages
// of radices the axis instantiates: each multiset in non-increasing and in
// non-decreasing order, multisets ranked as the rule planner ranks them
// (fewest stages, then smallest largest radix); the rule's plan first.
static void all_factorisations(int n, int rmax, std::vector<int>& cur, std::vector<std::vector<int>>& out)
{
    if (n == 1) {
        if (!cur.empty()) out.push_back(cur);
        return;
    }
    if ((int)cur.size() >= DWHIP_MAX_FFT_STAGES) return;
    for (int r : kRadices) {
        if (r > rmax || (!cur.empty() && r > cur.back()) || n % r != 0) continue;
        cur.push_back(r);
        all_factorisations(n / r, rmax, cur, out);
        cur.pop_back();
    }
}

int dwhip_radix_candidates(int n, int axis, std::vector<std::vector<int>>* out)
{
    out->clear();
    if (n <= 1) return DWHIP_OK;
    std::vector<std::vector<int>> ms;
    std::vector<int> cur;
    all_factorisations(n, axis_rmax(axis), cur, ms);
    if (ms.empty()) return DWHIP_ERR_SIZE;
    std::stable_sort(ms.begin(), ms.end(), [](const std::vector<int>& a, const std::vector<int>& b) {
        if (a.size() != b.size()) return a.size() < b.size();
        return *std::max_element(a.begin(), a.end()) < *std::max_element(b.begin(), b.end());
    });
    for (const auto& m : ms) {
        out->push_back(m);
        std::vector<int> rev(m.rbegin(), m.rend());
        if (rev != m) out->push_back(rev);
    }
    return DWHIP_OK;
}

// forward (radices R1..Rk) and inverse (Rk..R1) plans of length n, with
// the given radices (empty: the rule planner's)
static int make_plan_pair(int n, int axis, FftPlanDev* f, FftPlanDev* i, std::vector<float2>& tw,
                          const std::vector<int>* given = nullptr)
{
    std::vector<int> r = given && !given->empty() ? *given : dwhip_radices(n, axis);
    if (n > 1 && r.empty()) return DWHIP_ERR_SIZE;
    std::vector<int> rr(r.rbegin(), r.rend());
    int rc = dwhip_make_plan(n, r, f, tw);
    if (rc == DWHIP_OK) rc = dwhip_make_plan(n, rr, i, tw);
    return rc;
}

static int pow2_floor(int v)
{
    int p = 1;
    while (p * 2 <= v) p *= 2;
    return p;
}

// x tiles of pass E for SHB: ceil(N / xrows) tiles in each of the P planes
static int ratio_tiles(int64_t N, int64_t P, int xrows)
{
    return (int)(((N + xrows - 1) / xrows) * P);
}

// Tile shapes and LDS bytes for a transform size; pc (optional) gives
// smaller tiles than the rule's largest ones (a power of two for y and z).
void dwhip_tile_config(const int64_t* w, DwhipTiles* t, const DwhipPlanChoice* pc)
{
    const int h = (int)(w[0] / 2), ny = (int)w[1], nz = (int)w[2];
    const int hc = h + 1;
    t->xrows = std::max(1, std::min(DWHIP_XROWS_MAX, DWHIP_XMAXE * DWHIP_XTHR / h));
    // a column tile holds B pencils of n complex values in LDS
    const int ylim = std::min(DWHIP_YMAXE * DWHIP_CTHR, (int)(DWHIP_MAX_LDS_BYTES / sizeof(float2)));
    const int zlim = std::min(DWHIP_ZMAXE * DWHIP_ZTHR, (int)(DWHIP_MAX_LDS_BYTES / sizeof(float2)));
    t->ycols = std::max(1, std::min(64, pow2_floor(ylim / ny)));
    t->zcols = std::max(1, std::min(64, pow2_floor(zlim / nz)));
    if (pc) {
        if (pc->xrows > 0) t->xrows = std::min(t->xrows, pc->xrows);
        if (pc->ycols > 0) t->ycols = std::min(t->ycols, pow2_floor(pc->ycols));
        if (pc->zcols > 0) t->zcols = std::min(t->zcols, pow2_floor(pc->zcols));
    }
    t->ntile[KIND_X] = (ny * nz + t->xrows - 1) / t->xrows;
    t->ntile[KIND_Y] = ((hc + t->ycols - 1) / t->ycols) * nz;
    t->ntile[KIND_Z] = ((hc + t->zcols - 1) / t->zcols) * ny;
    // x tiles hold xrows rows of h values, or of hc values while the
    // half-spectrum rows are staged
    t->lds[KIND_X] = (size_t)t->xrows * (DWHIP_X_STAGE_LDS ? hc : h) * sizeof(float2);
    t->lds[KIND_Y] = (size_t)t->ycols * ny * sizeof(float2);
    t->lds[KIND_Z] = (size_t)t->zcols * nz * sizeof(float2);
    // the pruned y passes: all slabs until the job's image depth is known
    t->ntile[KIND_YD] = t->ntile[KIND_YF] = t->ntile[KIND_Y];
    t->lds[KIND_YD] = t->lds[KIND_YF] = t->lds[KIND_Y];
    t->ntile[KIND_XR] = t->ntile[KIND_X];
    t->lds[KIND_XR] = t->lds[KIND_X];
}

static size_t round_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

size_t dwhip_job_bytes(const int64_t* work, int64_t M, int64_t N, int64_t P, int64_t pM, int64_t pN,
                       int64_t pP, int border_quality, int niter)
{
    if (!dwhip_fast_size_ok(work)) {
        // any-size path: half spectra Z and K, U0, U1 (and W), partials, tables
        const size_t V = (size_t)work[0] * work[1] * work[2];
        const size_t Vh = round_up(work[0] / 2 + 1, 16) * work[1] * work[2];
        size_t b = 2 * Vh * sizeof(float2) + (border_quality > 0 ? 3 : 2) * V * sizeof(float);
        b += (size_t)dwhip_generic_ratio_grid() * sizeof(double);
        b += (size_t)(niter + 1) * (sizeof(float) + sizeof(double));
        b += DWHIP_STAT_BLOCKS * 4 * sizeof(double) + 4096 + sizeof(GenDev) + sizeof(JobDev);
        for (int d = 0; d < 3; ++d) b += 8 * (size_t)work[d] * sizeof(float2);  // twiddles, chirps, filters
        return b;
    }
    const size_t h = work[0] / 2, hcs = round_up(h + 1, 16);
    const size_t V = (size_t)work[0] * work[1] * work[2];
    const size_t Vc = hcs * work[1] * work[2];
    DwhipTiles t;
    dwhip_tile_config(work, &t);
    size_t b = 2 * Vc * sizeof(float2);                  // C, K
    b += (border_quality > 0 ? 3 : 2) * V * sizeof(float); // U0, U1, W
    b += (size_t)std::max(t.ntile[KIND_X], ratio_tiles(N, P, t.xrows)) * sizeof(double);
    b += (size_t)(niter + 1) * (sizeof(float) + sizeof(double));
    b += DWHIP_STAT_BLOCKS * 4 * sizeof(double) + 4096;
    (void)M; (void)N; (void)P; (void)pM; (void)pN; (void)pP;
    return b;
}

// ---------------------------------------------------------------------------
// jobs
// ---------------------------------------------------------------------------
#define HIPCHK(call)                                                   \
    do {                                                               \
        hipError_t e_ = (call);                                        \
        if (e_ != hipSuccess) return dwhip_hip_fail(e_, #call, __LINE__); \
    } while (0)

// Device blocks of the per-image path are kept between calls: a repeated
// job reuses the block instead of a hipMalloc/hipFree of its 1-5 GB each
// time.  One cached block per device, grown to the largest job seen and
// held until the process ends ($DWHIP_JOB_CACHE=0 turns this off); a call
// that finds it in use (another host thread's job) allocates its own.
namespace {
struct CachedBlock {
    void* p = nullptr;
    size_t bytes = 0;
    bool busy = false;
};
constexpr int kCacheDevices = 64;
std::mutex g_cache_mu;
CachedBlock g_cache[kCacheDevices];

bool job_cache_on()
{
    static const bool on = [] {
        const char* v = getenv("DWHIP_JOB_CACHE");
        return !(v && v[0] == '0');
    }();
    return on;
}
}  // namespace

static hipError_t job_mem_acquire(size_t bytes, void** out)
{
    int dev = 0;
    if (job_cache_on() && hipGetDevice(&dev) == hipSuccess && dev >= 0 && dev < kCacheDevices) {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        CachedBlock& c = g_cache[dev];
        if (!c.busy) {
            if (c.bytes < bytes) {
                if (c.p) (void)hipFree(c.p);
                c.p = nullptr;
                c.bytes = 0;
                const hipError_t e = hipMalloc(&c.p, bytes);
                if (e != hipSuccess) {
                    c.p = nullptr;
                    return e;
                }
                c.bytes = bytes;
            }
            c.busy = true;
            *out = c.p;
            return hipSuccess;
        }
    }
    return hipMalloc(out, bytes);
}

static void job_mem_release(void* p)
{
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        for (CachedBlock& c : g_cache)
            if (c.p == p && c.busy) {
                c.busy = false;
                return;
            }
    }
    (void)hipFree(p);
}

void dwhip_job_free(DwhipJob* j)
{
    if (!j) return;
    if (j->mem) job_mem_release(j->mem);
    j->mem = nullptr;
}

// Plans a job for device-resident im/psf: sizes, tiles, FFT plans and
// twiddles (into j->tw), and the byte layout of its device block
// (j->bytes).  Allocates nothing.
int dwhip_job_plan(DwhipJob* j, const float* im_dev, int64_t M, int64_t N, int64_t P, const float* psf_dev,
                   int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm, const float* start_dev)
{
    j->mem = nullptr;
    j->host_iter = 0;
    memset(&j->d, 0, sizeof(j->d));
    j->tw.clear();
    int64_t ref[3], w[3];
    int rc = dwhip_work_size(M, N, P, pM, pN, pP, prm->border_quality, w, ref);
    if (rc == DWHIP_ERR_ARG) return rc;
    bool user_work = prm->work[0] > 0 || prm->work[1] > 0 || prm->work[2] > 0;
    if (user_work) {
        for (int d = 0; d < 3; ++d) w[d] = prm->work[d] > 0 ? prm->work[d] : w[d];
        for (int d = 0; d < 3; ++d)
            if (w[d] < ref[d]) return DWHIP_ERR_ARG;
        if (prm->border_quality < 2)
            for (int d = 0; d < 3; ++d)
                if (w[d] != ref[d]) return DWHIP_ERR_ARG;  // padding changes bq 0/1 results
        if (!dwhip_fast_size_ok(w) && !dwhip_generic_size_ok(w)) return DWHIP_ERR_SIZE;
    } else if (rc != DWHIP_OK) {
        return rc;
    }
    j->generic = !dwhip_fast_size_ok(w);
    if (M * N * P > (int64_t)1 << 31 || w[0] * w[1] * w[2] > (int64_t)1 << 34) return DWHIP_ERR_SIZE;
    for (int d = 0; d < 3; ++d) {
        j->work[d] = w[d];
        j->ref_work[d] = ref[d];
    }
    const int niter = std::max(1, prm->iter_type == DWHIP_ITER_FIXED ? prm->niter : prm->maxiter);

    JobDev& d = j->d;
    d.nx = (int)w[0];
    d.ny = (int)w[1];
    d.nz = (int)w[2];
    d.h = d.nx / 2;
    d.hc = d.h + 1;
    d.hcs = (int)round_up(d.hc, 16);  // rows start on 128-byte lines
    d.M = (int)M; d.N = (int)N; d.P = (int)P;
    d.pM = (int)pM; d.pN = (int)pN; d.pP = (int)pP;
    d.refMNP = (double)ref[0] * (double)ref[1] * (double)ref[2];
    // plan: the autotuner's candidate, else a measured plan from the cache,
    // else the rule planner's
    DwhipPlanChoice cached;
    const DwhipPlanChoice* pc = j->force_plan;
    if (!pc && !j->generic && dwhip_plan_cached(w, &cached)) pc = &cached;
    dwhip_tile_config(w, &j->tiles, pc);
    {
        // pass D transforms slabs z < P; pass F too, except for RL, whose
        // ratio is 1e-6 outside the image
        const int nct = (d.hc + j->tiles.ycols - 1) / j->tiles.ycols;
        j->tiles.ntile[KIND_YD] = nct * (int)P;
        j->tiles.ntile[KIND_YF] = prm->method == DWHIP_METHOD_RL ? nct * (int)w[2] : nct * (int)P;
        // pass E: rows y < N of planes z < P for SHB (whose ratio is 0
        // elsewhere), all rows for RL
        if (prm->method != DWHIP_METHOD_RL) j->tiles.ntile[KIND_XR] = ratio_tiles(N, P, j->tiles.xrows);
    }
    for (int k = 0; k < KIND_N; ++k) d.ntile[k] = j->tiles.ntile[k];
    d.xrows = j->tiles.xrows;
    d.ycols = j->tiles.ycols;
    d.zcols = j->tiles.zcols;
    d.positivity = prm->positivity;
    d.metric = prm->metric;
    d.has_w = prm->border_quality > 0;
    d.bg_auto = prm->bg_auto;
    d.start_identity = prm->start != DWHIP_START_FLAT;
    d.niter = niter;
    d.rl = prm->method == DWHIP_METHOD_RL;
    d.iter_type = prm->iter_type;
    d.err_rel = prm->err_rel;
    d.err_abs = prm->err_abs;
    d.dev_stop = 0;
    d.alphamax = prm->alphamax;
    d.bg_user = prm->bg;
    d.im = im_dev;
    d.psf = psf_dev;
    d.start_img = start_dev;

    if (j->generic) {
        // any-size path: half spectra Z (in C) and K of hcs*ny*nz values, Z
        // also holding real rows of 2*hcs floats; k_round_end adds up the
        // ratio kernel's one partial per workgroup
        const int grc = dwhip_generic_plan(j);
        if (grc != DWHIP_OK) return grc;
        j->rtw_off = 0;
        for (int k = 0; k < KIND_N; ++k) d.ntile[k] = 0;
        d.ntile[KIND_XR] = dwhip_generic_ratio_grid();
        const size_t V = (size_t)d.nx * d.ny * d.nz;
        size_t off = 0;
        auto take = [&](size_t bytes) { size_t o = off; off = round_up(off + bytes, 256); return o; };
        DwhipLayout& L = j->lay;
        const size_t Vh = (size_t)d.hcs * d.ny * d.nz;
        L.C = take(Vh * sizeof(float2));
        L.K = take(Vh * sizeof(float2));
        L.U0 = take(V * sizeof(float));
        L.U1 = take(V * sizeof(float));
        L.W = d.has_w ? take(V * sizeof(float)) : 0;
        L.part = take((size_t)d.ntile[KIND_XR] * sizeof(double));
        L.stat = take(DWHIP_STAT_BLOCKS * 4 * sizeof(double));
        L.err = take((size_t)niter * sizeof(float));
        L.errd = take((size_t)niter * sizeof(double));
        L.tw = take(j->tw.size() * sizeof(float2));
        L.st = take(sizeof(JobState));
        L.tab = take(64 + round_up(sizeof(JobDev), 64) + sizeof(GenDev));
        j->bytes = off;
        return DWHIP_OK;
    }

    std::vector<float2>& tw = j->tw;
    if (make_plan_pair(d.h, 0, &d.pxf, &d.pxi, tw, pc ? &pc->r[0] : nullptr) ||
        make_plan_pair(d.ny, 1, &d.pyf, &d.pyi, tw, pc ? &pc->r[1] : nullptr) ||
        make_plan_pair(d.nz, 2, &d.pzf, &d.pzi, tw, pc ? &pc->r[2] : nullptr))
        return DWHIP_ERR_SIZE;
    j->rtw_off = tw.size();
    for (int k = 0; k < d.h; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)d.nx;
        tw.push_back(make_float2((float)cos(a), (float)sin(a)));
    }

    // layout of the job's block
    const size_t V = (size_t)d.nx * d.ny * d.nz;
    const size_t Vc = (size_t)d.hcs * d.ny * d.nz;
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t o = off; off = round_up(off + bytes, 256); return o; };
    DwhipLayout& L = j->lay;
    L.C = take(Vc * sizeof(float2));
    L.K = take(Vc * sizeof(float2));
    L.U0 = take(V * sizeof(float));
    L.U1 = take(V * sizeof(float));
    L.W = d.has_w ? take(V * sizeof(float)) : 0;
    L.part = take((size_t)std::max(d.ntile[KIND_X], d.ntile[KIND_XR]) * sizeof(double));
    L.stat = take(DWHIP_STAT_BLOCKS * 4 * sizeof(double));
    L.err = take((size_t)niter * sizeof(float));
    L.errd = take((size_t)niter * sizeof(double));
    L.tw = take(tw.size() * sizeof(float2));
    L.st = take(sizeof(JobState));
    L.tab = take(sizeof(BatchDev) + 64 + sizeof(JobDev));
    j->bytes = off;
    return DWHIP_OK;
}

// Points the job at its device block and queues the uploads (twiddles, the
// one-job table) on stream s.  The host copies must stay alive until s
// has run them (they live in the job).
int dwhip_job_bind(DwhipJob* j, void* mem, hipStream_t s)
{
    j->mem = mem;
    char* base = (char*)mem;
    const DwhipLayout& L = j->lay;
    JobDev& d = j->d;
    d.C = (float2*)(base + L.C);
    d.K = (float2*)(base + L.K);
    d.U0 = (float*)(base + L.U0);
    d.U1 = (float*)(base + L.U1);
    d.W = d.has_w ? (float*)(base + L.W) : nullptr;
    d.part = (double*)(base + L.part);
    d.stat = (double*)(base + L.stat);
    d.err = (float*)(base + L.err);
    d.errd = (double*)(base + L.errd);
    d.tw = (const float2*)(base + L.tw);
    d.rtw = d.tw + j->rtw_off;
    d.st = (JobState*)(base + L.st);
    j->bd = (BatchDev*)(base + L.tab);
    j->jd = (JobDev*)(base + L.tab + 64);
    memset(&j->bh, 0, sizeof(j->bh));
    j->bh.njobs = 1;
    for (int k = 0; k < KIND_N; ++k) {
        d.tile0[k] = 0;
        j->bh.ntile[k] = d.ntile[k];
    }
    j->dcopy = d;
    HIPCHK(hipMemcpyAsync(base + L.tw, j->tw.data(), j->tw.size() * sizeof(float2), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(j->bd, &j->bh, sizeof(j->bh), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(j->jd, &j->dcopy, sizeof(j->dcopy), hipMemcpyHostToDevice, s));
    if (j->generic) return dwhip_generic_bind(j, base + L.tab + 64 + round_up(sizeof(JobDev), 64), s);
    return DWHIP_OK;
}

// Plans, allocates and binds a job for the per-image path.
int dwhip_job_create(DwhipJob* j, const float* im_dev, int64_t M, int64_t N, int64_t P,
                     const float* psf_dev, int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm,
                     const float* start_dev)
{
    int rc = dwhip_job_plan(j, im_dev, M, N, P, psf_dev, pM, pN, pP, prm, start_dev);
    if (rc != DWHIP_OK) return rc;
    void* mem = nullptr;
    hipError_t e = job_mem_acquire(j->bytes, &mem);
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? DWHIP_ERR_NOMEM : DWHIP_ERR_HIP;
    rc = dwhip_job_bind(j, mem, nullptr);
    if (rc == DWHIP_OK) {
        e = hipStreamSynchronize(nullptr);
        if (e != hipSuccess) rc = dwhip_hip_fail(e, "job upload", __LINE__);
    }
    return rc;
}

// Algorithmic HBM bytes of one pass over job d: each operand read or
// written once, complex volumes counted at (nx/2+1)*ny*nz values (row
// padding excluded).  xkind: 1 for X passes, 0 for column passes; iter
// matters for the update pass (iteration 0 has no inverse transform).
double dwhip_pass_bytes(int xkind, int mode, const JobDev& d, int iter)
{
    const double row = (double)d.hc * sizeof(float2);          // one complex x-row
    const double Vc = row * d.ny * d.nz;
    const double V = (double)d.nx * d.ny * d.nz * sizeof(float);
    const double img = (double)d.M * d.N * d.P * sizeof(float);
    const double slabP = row * d.ny * d.P;                     // z < P
    const double imgrows = row * d.N * d.P;                    // y < N, z < P
    if (!xkind) {
        switch (mode) {
        case CM_ZCONV: return Vc + Vc + slabP;                 // read C, K; write z < P
        case CM_ZCONVJ: return d.rl ? 3 * Vc : slabP + Vc + Vc; // read z < P (RL: all), K; write C
        case CM_YINV_IMG: return slabP + imgrows;              // slabs z < P, write y < N
        case CM_YFWD_IMG: return d.rl ? 2 * Vc : imgrows + slabP;
        default: return 2 * Vc;
        }
    }
    switch (mode) {
    case XM_PSF: return Vc + (double)d.pM * d.pN * d.pP * sizeof(float);
    case XM_ONES: return Vc;
    case XM_WEIGHTS: return Vc + V;
    case XM_UPDATE:
        // SHB reads x_{k-1}, x_{k-2} (and W), writes x_k; RL reads x_{k-1}
        // (and W), writes x_k
        if (iter == 0) return (d.rl ? 1 : 2) * V + Vc;
        return 2 * Vc + ((d.rl ? 2 : 3) + (d.has_w ? 1 : 0)) * V;
    case XM_RATIO: return d.rl ? imgrows + Vc + img : 2 * imgrows + img;
    case XM_LOAD: return Vc + V;
    case XM_STORE: return Vc + V;
    default: return 0;
    }
}

static double host_iter(const DwhipJob* j, int mode)
{
    return mode == XM_UPDATE ? (double)j->host_iter : 0;
}

int dwhip_pass_x(int mode, const DwhipJob* j, const BatchDev* bd, const JobDev* jd, int ntile, size_t lds,
                 hipStream_t s)
{
    if (ntile <= 0) return DWHIP_OK;
    const int slot = dwhip_prof_begin(s);
    hipError_t e = dwhip_launch_xpass(mode, bd, jd, ntile, lds, s);
    if (slot >= 0) dwhip_prof_end(s, slot, mode, dwhip_pass_bytes(1, mode, j->d, (int)host_iter(j, mode)));
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "x pass", __LINE__);
}

int dwhip_pass_c(int mode, const DwhipJob* j, const BatchDev* bd, const JobDev* jd, int ntile, size_t lds,
                 hipStream_t s)
{
    if (ntile <= 0) return DWHIP_OK;
    const int slot = dwhip_prof_begin(s);
    hipError_t e = dwhip_launch_colpass(mode, bd, jd, ntile, lds, s);
    if (slot >= 0) dwhip_prof_end(s, slot, 7 + mode, dwhip_pass_bytes(0, mode, j->d, 0));
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "column pass", __LINE__);
}

#define RCHK(call)                      \
    do {                                \
        int r_ = (call);                \
        if (r_ != DWHIP_OK) return r_;  \
    } while (0)

// OTF, Bertero weights and start guess for a single job (method_shb.c:108-294).
int dwhip_job_setup(const DwhipJob* j, hipStream_t s)
{
    if (j->generic) return dwhip_generic_setup(j, s);
    const DwhipTiles& t = j->tiles;
    const int nblk = DWHIP_STAT_BLOCKS;
    HIPCHK(dwhip_launch_stats(j->jd, 0, nblk, s));
    HIPCHK(dwhip_launch_start(j->jd, 0, 1024, s));
    // OTF: shifted PSF -> R2C -> K / (nx*ny*nz)
    RCHK(dwhip_pass_x(XM_PSF, j, j->bd, j->jd, t.ntile[KIND_X], t.lds[KIND_X], s));
    RCHK(dwhip_pass_c(CM_YFWD, j, j->bd, j->jd, t.ntile[KIND_Y], t.lds[KIND_Y], s));
    RCHK(dwhip_pass_c(CM_ZOTF, j, j->bd, j->jd, t.ntile[KIND_Z], t.lds[KIND_Z], s));
    if (j->d.has_w) {
        // P1 = IFFT(conj(K) FFT(1_box)); W = P1 > 0.01 ? 1/P1 : 0
        RCHK(dwhip_pass_x(XM_ONES, j, j->bd, j->jd, t.ntile[KIND_X], t.lds[KIND_X], s));
        RCHK(dwhip_pass_c(CM_YFWD, j, j->bd, j->jd, t.ntile[KIND_Y], t.lds[KIND_Y], s));
        RCHK(dwhip_pass_c(CM_ZCONVJ, j, j->bd, j->jd, t.ntile[KIND_Z], t.lds[KIND_Z], s));
        RCHK(dwhip_pass_c(CM_YINV, j, j->bd, j->jd, t.ntile[KIND_Y], t.lds[KIND_Y], s));
        RCHK(dwhip_pass_x(XM_WEIGHTS, j, j->bd, j->jd, t.ntile[KIND_X], t.lds[KIND_X], s));
    }
    return DWHIP_OK;
}

// One round (iteration k) of a single job; `full` = also run the adjoint
// half (F, G, H), which only the next iteration's update consumes.
int dwhip_job_round(DwhipJob* j, bool full, hipStream_t s)
{
    if (j->generic) {
        RCHK(dwhip_generic_round(j, full, s));
        j->host_iter += 1;
        return DWHIP_OK;
    }
    const DwhipTiles& t = j->tiles;
    RCHK(dwhip_pass_x(XM_UPDATE, j, j->bd, j->jd, t.ntile[KIND_X], t.lds[KIND_X], s));
    RCHK(dwhip_pass_c(CM_YFWD, j, j->bd, j->jd, t.ntile[KIND_Y], t.lds[KIND_Y], s));
    RCHK(dwhip_pass_c(CM_ZCONV, j, j->bd, j->jd, t.ntile[KIND_Z], t.lds[KIND_Z], s));
    RCHK(dwhip_pass_c(CM_YINV_IMG, j, j->bd, j->jd, t.ntile[KIND_YD], t.lds[KIND_YD], s));
    RCHK(dwhip_pass_x(XM_RATIO, j, j->bd, j->jd, t.ntile[KIND_XR], t.lds[KIND_XR], s));
    if (full) {
        RCHK(dwhip_pass_c(CM_YFWD_IMG, j, j->bd, j->jd, t.ntile[KIND_YF], t.lds[KIND_YF], s));
        RCHK(dwhip_pass_c(CM_ZCONVJ, j, j->bd, j->jd, t.ntile[KIND_Z], t.lds[KIND_Z], s));
        RCHK(dwhip_pass_c(CM_YINV, j, j->bd, j->jd, t.ntile[KIND_Y], t.lds[KIND_Y], s));
    }
    HIPCHK(dwhip_launch_round_end(j->bd, j->jd, 1, 1, s));
    j->host_iter += 1;
    return DWHIP_OK;
}

// x_n = max(H^T(ratio_{n-1}) * x_{n-1} * W, bg) of Richardson-Lucy after
// its last round (method_rl.c:85-97, 357-366): the update half of pass A.
// Its forward transform is written to C and not used.
int dwhip_job_final_update(DwhipJob* j, const BatchDev* bd, const JobDev* jd, hipStream_t s)
{
    if (j->generic) return dwhip_generic_final_update(j, s);
    return dwhip_pass_x(XM_UPDATE, j, bd, jd, j->tiles.ntile[KIND_X], j->tiles.lds[KIND_X], s);
}

// dw_iterator_next for REL / ABS (src/dw.c:235-268), given the error of the
// last two iterations run.
static bool iterator_continue(const dwhip_params* p, int iter, float error, float lasterror)
{
    if (p->iter_type == DWHIP_ITER_REL) {
        if (iter < 2) lasterror = 2 * error * p->err_rel;
        // C's fabs: the float difference, then double arithmetic
        if (fabs((double)(error - lasterror)) / (double)error < (double)p->err_rel) return false;
    } else if (p->iter_type == DWHIP_ITER_ABS) {
        if (iter > 0 && error < p->err_abs) return false;
    }
    const int niter = p->iter_type == DWHIP_ITER_FIXED ? p->niter : p->maxiter;
    return iter < niter;
}

// One job on stream s.  mem: a device block of cap bytes for the job's
// state, or nullptr to allocate one here for the call.  With a given block
// nothing is allocated on the device (no timing events either: setup_ms
// and iter_ms read 0), and a job needing more than cap bytes is refused
// with DWHIP_ERR_SCRATCH before anything is queued.
// FIXED iterations of a job on stream s: niter rounds, the last one of SHB
// without its adjoint half (SHB returns x_{n-1}, so that half feeds
// nothing).  Graph mode 1 replays one captured full round for every round
// but SHB's last; mode 2 captures the whole loop; mode 0, a job bound to a
// caller's block (nothing may be allocated, and instantiating a graph
// allocates) and profiled runs (events per pass) launch every pass.
struct LoopCtx {
    DwhipJob* job;
    int niter;
    int full_only;  // capture one full round
};

static int loop_body(void* ctx, hipStream_t cs)
{
    LoopCtx* c = (LoopCtx*)ctx;
    if (c->full_only) return dwhip_job_round(c->job, true, cs);
    for (int k = 0; k < c->niter; ++k) RCHK(dwhip_job_round(c->job, c->job->d.rl || k + 1 < c->niter, cs));
    return DWHIP_OK;
}

// *exec receives the graph launched (destroyed by the caller once s has
// run it), or nullptr.
static int fixed_loop(DwhipJob* job, int niter, bool may_graph, hipStream_t s, hipGraphExec_t* exec_out)
{
    *exec_out = nullptr;
    const int mode = may_graph && !dwhip_prof_on() && niter > 1 ? dwhip_graph_mode() : 0;
    if (mode == 0) {
        for (int k = 0; k < niter; ++k) RCHK(dwhip_job_round(job, job->d.rl || k + 1 < niter, s));
        return DWHIP_OK;
    }
    const int nfull = mode == 2 ? 1 : (job->d.rl ? niter : niter - 1);
    LoopCtx c{job, niter, mode == 1};
    hipGraphExec_t exec = nullptr;
    RCHK(dwhip_capture(loop_body, &c, &exec));
    *exec_out = exec;
    hipError_t e = hipSuccess;
    for (int k = 0; k < nfull && e == hipSuccess; ++k) e = hipGraphLaunch(exec, s);
    if (e != hipSuccess) return dwhip_hip_fail(e, "hipGraphLaunch", __LINE__);
    job->host_iter = mode == 2 ? niter : nfull;
    if (mode == 1 && !job->d.rl) RCHK(dwhip_job_round(job, false, s));
    return DWHIP_OK;
}

// The PSF inserted at the transform size w and shifted so that its
// maximum (lowest linear index on ties) is at the origin: the volume Z of
// method_shb.c:184-209 (fim_insert, fim_argmax_max, fim_circshift).
static void shifted_psf_host(const float* psf, int64_t pM, int64_t pN, int64_t pP, const int64_t* w, float* Z)
{
    const size_t np = (size_t)pM * pN * pP;
    size_t a = 0;
    for (size_t i = 1; i < np; ++i)
        if (psf[i] > psf[a]) a = i;
    const int64_t mz = (int64_t)(a / ((size_t)pM * pN)), my = (int64_t)((a / pM) % pN), mx = (int64_t)(a % pM);
    for (int64_t z = 0; z < w[2]; ++z)
        for (int64_t y = 0; y < w[1]; ++y)
            for (int64_t x = 0; x < w[0]; ++x) {
                const int64_t sx = (x + mx) % w[0], sy = (y + my) % w[1], sz = (z + mz) % w[2];
                Z[x + w[0] * (y + w[1] * z)] =
                    (sx < pM && sy < pN && sz < pP) ? psf[sx + pM * (sy + pN * sz)] : 0.f;
            }
}

// The loop with the caller's per-iteration hook: every round launched
// directly, the host waits for it, reads its error and, every hook_every
// iterations, the iterate x_k cropped into `out` (its scratch until the
// final crop) and copied to the host; then the optional full-size calls.
static int hooked_loop(DwhipJob* job, const dwhip_params* prm, float* out, hipStream_t s, int* nrun)
{
    const bool fixed = prm->iter_type == DWHIP_ITER_FIXED;
    const JobDev& d = job->d;
    const size_t n_img = (size_t)d.M * d.N * d.P;
    std::vector<float> hx(prm->hook_every > 0 ? n_img : 0);
    float error = 1.f, lasterror = 1.f;
    *nrun = 0;
    for (int k = 0;; ++k) {
        if (fixed ? k >= prm->niter : !iterator_continue(prm, k, error, lasterror)) break;
        RCHK(dwhip_job_round(job, job->d.rl || !fixed || k + 1 < prm->niter, s));
        *nrun = k + 1;
        float ek = 0;
        HIPCHK(hipMemcpyAsync(&ek, d.err + k, sizeof(float), hipMemcpyDeviceToHost, s));
        const float* xh = nullptr;
        if (prm->hook_every > 0 && k % prm->hook_every == 0) {
            // x_k: RL's rounds end with x_k in U[k & 1] as SHB's do
            HIPCHK(dwhip_launch_crop(job->jd, 0, k, out, 1024, s));
            HIPCHK(hipMemcpyAsync(hx.data(), out, n_img * sizeof(float), hipMemcpyDeviceToHost, s));
            xh = hx.data();
        }
        HIPCHK(hipStreamSynchronize(s));
        prm->hook(prm->hook_user, k, ek, xh, d.M, d.N, d.P);
        lasterror = error;
        error = ek;
    }
    if (prm->hook_full) {
        const int64_t w[3] = {d.nx, d.ny, d.nz};
        const size_t V = (size_t)w[0] * w[1] * w[2];
        std::vector<float> full(V);
        if (job->d.rl && *nrun > 0) RCHK(dwhip_job_final_update(job, job->bd, job->jd, s));
        const int k = dwhip_result_iterate(job->d.rl != 0, *nrun);
        const float* U = (k & 1) ? d.U1 : d.U0;
        HIPCHK(hipMemcpyAsync(full.data(), U, V * sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        prm->hook(prm->hook_user, -1, 0.f, full.data(), w[0], w[1], w[2]);
        std::vector<float> hp((size_t)d.pM * d.pN * d.pP);
        HIPCHK(hipMemcpy(hp.data(), d.psf, hp.size() * sizeof(float), hipMemcpyDeviceToHost));
        shifted_psf_host(hp.data(), d.pM, d.pN, d.pP, w, full.data());
        prm->hook(prm->hook_user, -2, 0.f, full.data(), w[0], w[1], w[2]);
    }
    return DWHIP_OK;
}

static int run_job(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                   int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st, hipStream_t s, void* mem,
                   size_t cap)
{
    DwhipJob job;
    const float* start = (prm->start != DWHIP_START_FLAT) ? prm->start_image : nullptr;
    int rc;
    if (!mem && prm->planning == DWHIP_PLAN_MEASURE) {
        // measure this transform size's plans on its first use (a job bound
        // to a caller's block allocates nothing, so plans by rule or cache)
        int64_t w[3], ref[3];
        if (dwhip_work_size(M, N, P, pM, pN, pP, prm->border_quality, w, ref) == DWHIP_OK) {
            for (int d = 0; d < 3; ++d)
                if (prm->work[d] > 0) w[d] = prm->work[d];
            rc = dwhip_plan_ensure(w, prm->method);
            if (rc != DWHIP_OK) return rc;
        }
    }
    if (mem) {
        rc = dwhip_job_plan(&job, im, M, N, P, psf, pM, pN, pP, prm, start);
        if (rc == DWHIP_OK && job.bytes > cap) rc = DWHIP_ERR_SCRATCH;
        if (rc == DWHIP_OK) rc = dwhip_job_bind(&job, mem, s);
        if (rc != DWHIP_OK) {
            job.mem = nullptr;  // not ours to free
            return rc;
        }
    } else {
        rc = dwhip_job_create(&job, im, M, N, P, psf, pM, pN, pP, prm, start);
        if (rc != DWHIP_OK) {
            dwhip_job_free(&job);
            return rc;
        }
    }
    const bool timed = mem == nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr;
    if (timed) {
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        hipEventCreate(&e2);
        hipEventRecord(e0, s);
    }
    rc = dwhip_job_setup(&job, s);
    if (timed) hipEventRecord(e1, s);

    const bool fixed = prm->iter_type == DWHIP_ITER_FIXED;
    int nrun = 0;
    hipGraphExec_t loop_graph = nullptr;
    if (rc == DWHIP_OK && prm->hook) {
        rc = hooked_loop(&job, prm, out, s, &nrun);
    } else if (rc == DWHIP_OK) {
        if (fixed) {
            rc = fixed_loop(&job, prm->niter, mem == nullptr, s, &loop_graph);
            if (rc == DWHIP_OK) nrun = prm->niter;
        } else {
            // the error of iteration k decides whether k+1 runs
            float error = 1.f, lasterror = 1.f;
            for (int k = 0; rc == DWHIP_OK; ++k) {
                if (!iterator_continue(prm, k, error, lasterror)) break;
                rc = dwhip_job_round(&job, true, s);
                if (rc != DWHIP_OK) break;
                nrun = k + 1;
                float ek = 0;
                hipError_t e = hipMemcpyAsync(&ek, job.d.err + k, sizeof(float), hipMemcpyDeviceToHost, s);
                if (e == hipSuccess) e = hipStreamSynchronize(s);
                if (e != hipSuccess) rc = dwhip_hip_fail(e, "error readback", __LINE__);
                lasterror = error;
                error = ek;
            }
        }
    }
    // RL returns x_n: one more update pass turns the last adjoint into x_n
    // (the hooked loop has run it when it handed x_n out at full size)
    if (rc == DWHIP_OK && nrun > 0 && job.d.rl && !(prm->hook && prm->hook_full))
        rc = dwhip_job_final_update(&job, job.bd, job.jd, s);
    if (timed) hipEventRecord(e2, s);
    if (rc == DWHIP_OK) {
        // with no iteration run (REL stopping before iteration 0) the result
        // is the start guess x_0, which the swap-back of method_shb.c:385-390
        // hands back (deconvolve_rl would crop a NULL iterate there)
        hipError_t e = dwhip_launch_crop(job.jd, 0, dwhip_result_iterate(job.d.rl != 0, nrun), out, 1024, s);
        if (e != hipSuccess) rc = dwhip_hip_fail(e, "crop", __LINE__);
    }
    hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess && rc == DWHIP_OK) rc = dwhip_hip_fail(e, "synchronize", __LINE__);
    if (loop_graph) (void)hipGraphExecDestroy(loop_graph);
    if (dwhip_prof_on()) dwhip_prof_collect();
    if (rc == DWHIP_OK && st) {
        JobState hs;
        hipMemcpy(&hs, job.d.st, sizeof(hs), hipMemcpyDeviceToHost);
        st->bg = hs.bg;
        st->iterations = nrun;
        for (int d = 0; d < 3; ++d) {
            st->work[d] = job.work[d];
            st->ref_work[d] = job.ref_work[d];
        }
        st->setup_ms = st->iter_ms = 0;
        if (timed) {
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            st->setup_ms = ms;
            hipEventElapsedTime(&ms, e1, e2);
            st->iter_ms = ms;
        }
        if (st->err_trace && st->err_trace_len > 0 && nrun > 0) {
            const int n = std::min(nrun, st->err_trace_len);
            hipMemcpy(st->err_trace, job.d.err, n * sizeof(float), hipMemcpyDeviceToHost);
        }
        if (prm->verbosity > 0)
            printf("dwhip: image %lldx%lldx%lld psf %lldx%lldx%lld job %lldx%lldx%lld (ref %lldx%lldx%lld), "
                   "%d iterations, setup %.2f ms, iterations %.2f ms\n",
                   (long long)M, (long long)N, (long long)P, (long long)pM, (long long)pN, (long long)pP,
                   (long long)job.work[0], (long long)job.work[1], (long long)job.work[2],
                   (long long)job.ref_work[0], (long long)job.ref_work[1], (long long)job.ref_work[2], nrun,
                   st->setup_ms, st->iter_ms);
    }
    if (timed) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        hipEventDestroy(e2);
        dwhip_job_free(&job);
    }
    return rc;
}

static int check_args(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                      int64_t pP, const dwhip_params* prm, const float* out)
{
    if (!im || !psf || !prm || !out) return DWHIP_ERR_ARG;
    if (M < 1 || N < 1 || P < 1 || pM < 1 || pN < 1 || pP < 1) return DWHIP_ERR_ARG;
    if (prm->iter_type < DWHIP_ITER_ABS || prm->iter_type > DWHIP_ITER_FIXED) return DWHIP_ERR_ARG;
    if (prm->metric != DWHIP_METRIC_MSE && prm->metric != DWHIP_METRIC_IDIV) return DWHIP_ERR_ARG;
    if (prm->start < DWHIP_START_FLAT || prm->start > DWHIP_START_IDENTITY) return DWHIP_ERR_ARG;
    if (prm->method != DWHIP_METHOD_SHB && prm->method != DWHIP_METHOD_RL) return DWHIP_ERR_ARG;
    if (prm->niter < 0) return DWHIP_ERR_ARG;
    if (prm->iter_type != DWHIP_ITER_FIXED && prm->maxiter < 0) return DWHIP_ERR_ARG;
    return DWHIP_OK;
}

int dwhip_run_device_in(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM,
                        int64_t pN, int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st,
                        void* stream, void* mem, size_t cap)
{
    int rc = check_args(im, M, N, P, psf, pM, pN, pP, prm, out);
    if (rc != DWHIP_OK) return rc;
    rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    if (prm->niter == 0) {  // any iter_type, as the reference tests nIter alone
        // method_shb.c:101-105, method_rl.c:120-123
        HIPCHK(hipMemcpyAsync(out, im, (size_t)M * N * P * sizeof(float), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        if (st) {
            st->bg = prm->bg;
            st->iterations = 0;
        }
        return DWHIP_OK;
    }
    return run_job(im, M, N, P, psf, pM, pN, pP, prm, out, st, s, mem, cap);
}

static int run_device(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                      int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st, void* stream)
{
    return dwhip_run_device_in(im, M, N, P, psf, pM, pN, pP, prm, out, st, stream, nullptr, 0);
}

static int run_host(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                  int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st)
{
    int rc = check_args(im, M, N, P, psf, pM, pN, pP, prm, out);
    if (rc != DWHIP_OK) return rc;
    const size_t n = (size_t)M * N * P, np = (size_t)pM * pN * pP;
    if (prm->niter == 0) {  // any iter_type, as the reference tests nIter alone
        memcpy(out, im, n * sizeof(float));  // method_shb.c:101-105, method_rl.c:120-123
        if (st) {
            st->bg = prm->bg;
            st->iterations = 0;
        }
        return DWHIP_OK;
    }
    rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    const bool has_start = prm->start != DWHIP_START_FLAT && prm->start_image;
    float *dim = nullptr, *dpsf = nullptr, *dout = nullptr, *dstart = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreate(&s);
    if (e == hipSuccess) e = hipMalloc(&dim, n * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dpsf, np * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&dout, n * sizeof(float));
    if (e == hipSuccess && has_start) e = hipMalloc(&dstart, n * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(dim, im, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dpsf, psf, np * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess && has_start) e = hipMemcpy(dstart, prm->start_image, n * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        rc = dwhip_hip_fail(e, "upload", __LINE__);
    } else {
        dwhip_params p2 = *prm;
        p2.start_image = dstart;
        rc = run_job(dim, M, N, P, dpsf, pM, pN, pP, &p2, dout, st, s, nullptr, 0);
        if (rc == DWHIP_OK) {
            e = hipMemcpy(out, dout, n * sizeof(float), hipMemcpyDeviceToHost);
            if (e != hipSuccess) rc = dwhip_hip_fail(e, "download", __LINE__);
        }
    }
    if (dim) hipFree(dim);
    if (dpsf) hipFree(dpsf);
    if (dout) hipFree(dout);
    if (dstart) hipFree(dstart);
    if (s) hipStreamDestroy(s);
    return rc;
}

static dwhip_params with_method(const dwhip_params* prm, int method)
{
    dwhip_params p = *prm;
    p.method = method;
    return p;
}

int dwhip_shb_run_device(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                         int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st, void* stream)
{
    if (!prm) return DWHIP_ERR_ARG;
    const dwhip_params p = with_method(prm, DWHIP_METHOD_SHB);
    return run_device(im, M, N, P, psf, pM, pN, pP, &p, out, st, stream);
}

int dwhip_rl_run_device(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                        int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st, void* stream)
{
    if (!prm) return DWHIP_ERR_ARG;
    const dwhip_params p = with_method(prm, DWHIP_METHOD_RL);
    return run_device(im, M, N, P, psf, pM, pN, pP, &p, out, st, stream);
}

int dwhip_shb_run(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                  int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st)
{
    if (!prm) return DWHIP_ERR_ARG;
    const dwhip_params p = with_method(prm, DWHIP_METHOD_SHB);
    return run_host(im, M, N, P, psf, pM, pN, pP, &p, out, st);
}

int dwhip_rl_run(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM, int64_t pN,
                 int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st)
{
    if (!prm) return DWHIP_ERR_ARG;
    const dwhip_params p = with_method(prm, DWHIP_METHOD_RL);
    return run_host(im, M, N, P, psf, pM, pN, pP, &p, out, st);
}

// ---------------------------------------------------------------------------
// FFT check entry
// ---------------------------------------------------------------------------
int dwhip_fft3_device(float* vol, float* spec, int64_t nx, int64_t ny, int64_t nz, int inverse, void* stream)
{
    if (!vol || !spec) return DWHIP_ERR_ARG;
    int64_t w[3] = {nx, ny, nz};
    if (!dwhip_fast_size_ok(w)) return DWHIP_ERR_SIZE;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    // a job whose U0 is `vol`; C is its own buffer, copied to/from `spec`
    dwhip_params prm;
    dwhip_params_default(&prm, 1);
    prm.border_quality = 0;
    prm.work[0] = nx;
    prm.work[1] = ny;
    prm.work[2] = nz;
    DwhipJob job;
    rc = dwhip_job_create(&job, vol, nx, ny, nz, vol, 1, 1, 1, &prm, nullptr);
    if (rc != DWHIP_OK) {
        dwhip_job_free(&job);
        return rc;
    }
    const JobDev& d = job.d;
    // point U0 at the caller's volume
    JobDev d2 = d;
    d2.U0 = vol;
    hipError_t e = hipMemcpy(job.jd, &d2, sizeof(d2), hipMemcpyHostToDevice);
    const DwhipTiles& t = job.tiles;
    const size_t rowb = (size_t)d.hc * sizeof(float2), pitch = (size_t)d.hcs * sizeof(float2);
    if (e == hipSuccess) {
        if (!inverse) {
            rc = dwhip_pass_x(XM_LOAD, &job, job.bd, job.jd, t.ntile[KIND_X], t.lds[KIND_X], s);
            if (!rc) rc = dwhip_pass_c(CM_YFWD, &job, job.bd, job.jd, t.ntile[KIND_Y], t.lds[KIND_Y], s);
            if (!rc) rc = dwhip_pass_c(CM_ZFWD, &job, job.bd, job.jd, t.ntile[KIND_Z], t.lds[KIND_Z], s);
            if (!rc) e = hipMemcpy2DAsync(spec, rowb, d.C, pitch, rowb, (size_t)ny * nz, hipMemcpyDeviceToDevice, s);
        } else {
            e = hipMemcpy2DAsync(d.C, pitch, spec, rowb, rowb, (size_t)ny * nz, hipMemcpyDeviceToDevice, s);
            if (e == hipSuccess) {
                rc = dwhip_pass_c(CM_ZINV, &job, job.bd, job.jd, t.ntile[KIND_Z], t.lds[KIND_Z], s);
                if (!rc) rc = dwhip_pass_c(CM_YINV, &job, job.bd, job.jd, t.ntile[KIND_Y], t.lds[KIND_Y], s);
                if (!rc) rc = dwhip_pass_x(XM_STORE, &job, job.bd, job.jd, t.ntile[KIND_X], t.lds[KIND_X], s);
            }
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess && rc == DWHIP_OK) rc = dwhip_hip_fail(e, "fft3", __LINE__);
    dwhip_job_free(&job);
    return rc;
}
