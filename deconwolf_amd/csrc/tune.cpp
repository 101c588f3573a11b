// Measured planning of the fused passes and its plan cache: the analogue of
// FFTW's planner with FFTW_MEASURE and its wisdom file, which deconwolf
// uses for its CPU transforms (src/fft.c:135-172,206-319; the planning
// flag is dw_opts.fftw3_planning, src/dw.c:400).
//
// For a transform size the rule planner (dwhip_radices, dwhip_tile_config)
// gives one plan; the tuner times, axis by axis, the other radix orders the
// kernels can run (up to four stages of the axis's radices, each multiset
// in both monotone orders) and the next smaller tile shapes, on the passes
// that depend on that axis (x: the update pass; y: forward and inverse
// column passes; z: the OTF junction pass), and keeps a candidate only if
// it beats the rule's plan by more than 3 %.  Plans are kept per process
// and in a text file, one line per size:
//   gfx950 <nx> <ny> <nz> x=<r,r,..> xrows=<n> y=<r,..> ycols=<n> z=<r,..> zcols=<n>
// at $DWHIP_PLAN_CACHE if set (empty: no file), else
// $HOME/.config/deconwolf/dwhip_plans_gfx950.txt when $HOME/.config exists,
// where deconwolf keeps its FFTW wisdom.
#include <hip/hip_runtime.h>

#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "shb_kernels.h"

namespace {

using Key = std::array<int64_t, 3>;

struct Cache {
    std::mutex mu;
    bool loaded = false;
    bool path_set = false;
    std::string path;
    std::map<Key, DwhipPlanChoice> plans;
};
Cache g_cache;

std::string default_path()
{
    const char* env = getenv("DWHIP_PLAN_CACHE");
    if (env) return env;
    const char* home = getenv("HOME");
    if (!home || !*home) return "";
    std::string cfg = std::string(home) + "/.config";
    struct stat sb;
    if (stat(cfg.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode)) return "";
    cfg += "/deconwolf";
    if (stat(cfg.c_str(), &sb) != 0 && mkdir(cfg.c_str(), 0755) != 0 && errno != EEXIST) return "";
    return cfg + "/dwhip_plans_gfx950.txt";
}

std::string join(const std::vector<int>& r)
{
    std::string s;
    for (size_t i = 0; i < r.size(); ++i) s += (i ? "," : "") + std::to_string(r[i]);
    return s;
}

std::vector<int> split(const char* s)
{
    std::vector<int> r;
    while (*s) {
        char* end = nullptr;
        const long v = strtol(s, &end, 10);
        if (end == s) break;
        r.push_back((int)v);
        s = *end == ',' ? end + 1 : end;
    }
    return r;
}

// whether a plan can run: each radix list a factorisation of its length the
// axis kernels instantiate (the candidate enumeration holds all of them)
bool plan_valid(const Key& w, const DwhipPlanChoice& c)
{
    const int len[3] = {(int)(w[0] / 2), (int)w[1], (int)w[2]};
    for (int a = 0; a < 3; ++a) {
        if (c.r[a].empty()) continue;
        std::vector<std::vector<int>> cand;
        if (dwhip_radix_candidates(len[a], a, &cand) != DWHIP_OK) return false;
        if (std::find(cand.begin(), cand.end(), c.r[a]) == cand.end()) return false;
    }
    return c.xrows >= 0 && c.ycols >= 0 && c.zcols >= 0;
}

void load_locked()
{
    if (g_cache.loaded) return;
    g_cache.loaded = true;
    if (!g_cache.path_set) {
        g_cache.path = default_path();
        g_cache.path_set = true;
    }
    if (g_cache.path.empty()) return;
    FILE* f = fopen(g_cache.path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        char arch[32], xs[96], ys[96], zs[96];
        long long nx, ny, nz;
        int xr, yc, zc;
        if (sscanf(line, "%31s %lld %lld %lld x=%95s xrows=%d y=%95s ycols=%d z=%95s zcols=%d", arch, &nx, &ny, &nz,
                   xs, &xr, ys, &yc, zs, &zc) != 10 ||
            strcmp(arch, "gfx950") != 0)
            continue;
        DwhipPlanChoice c;
        c.r[0] = split(xs);
        c.r[1] = split(ys);
        c.r[2] = split(zs);
        c.xrows = xr;
        c.ycols = yc;
        c.zcols = zc;
        const Key k{nx, ny, nz};
        if (plan_valid(k, c)) g_cache.plans[k] = c;
    }
    fclose(f);
}

void save_locked()
{
    if (g_cache.path.empty()) return;
    const std::string tmp = g_cache.path + ".tmp";
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return;
    for (const auto& kv : g_cache.plans) {
        const DwhipPlanChoice& c = kv.second;
        // radices written out in full (an axis that kept the rule's plan
        // has an empty list in memory)
        const int len[3] = {(int)(kv.first[0] / 2), (int)kv.first[1], (int)kv.first[2]};
        std::string r[3];
        for (int a = 0; a < 3; ++a) r[a] = join(c.r[a].empty() ? dwhip_radices(len[a], a) : c.r[a]);
        if (r[0].empty() || r[1].empty() || r[2].empty()) continue;  // length-1 axes: nothing to choose
        fprintf(f, "gfx950 %lld %lld %lld x=%s xrows=%d y=%s ycols=%d z=%s zcols=%d\n", (long long)kv.first[0],
                (long long)kv.first[1], (long long)kv.first[2], r[0].c_str(), c.xrows, r[1].c_str(), c.ycols,
                r[2].c_str(), c.zcols);
    }
    fclose(f);
    rename(tmp.c_str(), g_cache.path.c_str());
}

#define TCHK(call)                                                        \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return dwhip_hip_fail(e_, #call, __LINE__); \
    } while (0)

// Mean time (ms) of the passes of `axis` for plan c on the block mem.
int time_axis(const Key& w, int method, const DwhipPlanChoice& c, int axis, void* mem, size_t cap, hipStream_t s,
              hipEvent_t e0, hipEvent_t e1, double* ms)
{
    *ms = 1e30;
    DwhipJob j;
    j.force_plan = &c;
    dwhip_params prm;
    dwhip_params_default(&prm, 2);
    prm.method = method;
    for (int d = 0; d < 3; ++d) prm.work[d] = w[d];
    // the image and PSF are not read by the timed passes; the job sizes are
    // those of a job whose work size is w
    const float* none = reinterpret_cast<const float*>(mem);
    int rc = dwhip_job_plan(&j, none, std::max<int64_t>(1, w[0] / 2), std::max<int64_t>(1, w[1] / 2),
                            std::max<int64_t>(1, w[2] / 2), none, 1, 1, 1, &prm, nullptr);
    if (rc != DWHIP_OK) return rc;
    if (j.bytes > cap) return DWHIP_ERR_NOMEM;
    rc = dwhip_job_bind(&j, mem, s);
    if (rc != DWHIP_OK) return rc;
    JobState st;
    memset(&st, 0, sizeof st);
    st.bg = 1.f;
    st.c0 = 1.f;
    st.iter = 1;  // an update pass with both of its transforms
    TCHK(hipMemcpyAsync(j.d.st, &st, sizeof st, hipMemcpyHostToDevice, s));
    auto launch = [&]() -> int {
        const DwhipTiles& t = j.tiles;
        if (axis == 0) return dwhip_pass_x(XM_UPDATE, &j, j.bd, j.jd, t.ntile[KIND_X], t.lds[KIND_X], s);
        if (axis == 1) {
            const int r = dwhip_pass_c(CM_YFWD, &j, j.bd, j.jd, t.ntile[KIND_Y], t.lds[KIND_Y], s);
            return r != DWHIP_OK ? r : dwhip_pass_c(CM_YINV, &j, j.bd, j.jd, t.ntile[KIND_Y], t.lds[KIND_Y], s);
        }
        return dwhip_pass_c(CM_ZCONV, &j, j.bd, j.jd, t.ntile[KIND_Z], t.lds[KIND_Z], s);
    };
    rc = launch();  // warm-up
    if (rc != DWHIP_OK) return rc;
    const int reps = 3;
    TCHK(hipEventRecord(e0, s));
    for (int k = 0; k < reps && rc == DWHIP_OK; ++k) rc = launch();
    TCHK(hipEventRecord(e1, s));
    TCHK(hipEventSynchronize(e1));
    if (rc != DWHIP_OK) return rc;
    float t = 0;
    TCHK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return DWHIP_OK;
}

// Measures the plans of size w; *tuned / *rule: the summed pass times of
// the chosen and of the rule's plan.
int tune_size(const Key& w, int method, DwhipPlanChoice* best, double* tuned, double* rule)
{
    const int len[3] = {(int)(w[0] / 2), (int)w[1], (int)w[2]};
    std::vector<std::vector<int>> rad[3];
    for (int a = 0; a < 3; ++a)
        if (dwhip_radix_candidates(len[a], a, &rad[a]) != DWHIP_OK) return DWHIP_ERR_SIZE;
    DwhipTiles rt;
    const int64_t wv[3] = {w[0], w[1], w[2]};
    dwhip_tile_config(wv, &rt);
    const int rule_tile[3] = {rt.xrows, rt.ycols, rt.zcols};

    // one block for every candidate (the layout depends on the plan only
    // through the twiddle table and the partial sums)
    size_t cap = dwhip_job_bytes(wv, std::max<int64_t>(1, w[0] / 2), std::max<int64_t>(1, w[1] / 2),
                                 std::max<int64_t>(1, w[2] / 2), 1, 1, 1, 2, 2) + ((size_t)4 << 20);
    void* mem = nullptr;
    TCHK(hipMalloc(&mem, cap));
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = DWHIP_OK;
    hipError_t he = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (he == hipSuccess) he = hipEventCreate(&e0);
    if (he == hipSuccess) he = hipEventCreate(&e1);
    if (he == hipSuccess) he = hipMemsetAsync(mem, 0, cap, s);
    if (he != hipSuccess) rc = dwhip_hip_fail(he, "tuner set-up", __LINE__);

    DwhipPlanChoice cur;  // rule plan: empty lists, zero tiles
    double sum_best = 0, sum_rule = 0;
    for (int a = 0; a < 3 && rc == DWHIP_OK; ++a) {
        double t_rule = 0;
        rc = time_axis(w, method, cur, a, mem, cap, s, e0, e1, &t_rule);
        if (rc != DWHIP_OK) break;
        double t_best = t_rule;
        DwhipPlanChoice pick = cur;
        // radix orders at the rule's tiles, then smaller tiles for the best
        for (size_t i = 1; i < rad[a].size() && i < 8 && rc == DWHIP_OK; ++i) {
            DwhipPlanChoice c = cur;
            c.r[a] = rad[a][i];
            double t = 0;
            rc = time_axis(w, method, c, a, mem, cap, s, e0, e1, &t);
            if (rc == DWHIP_OK && t < t_best * 0.97) {
                t_best = t;
                pick = c;
            }
        }
        for (int div = 2; div <= 4 && rc == DWHIP_OK; div *= 2) {
            if (rule_tile[a] / div < 1) break;
            DwhipPlanChoice c = pick;
            (a == 0 ? c.xrows : a == 1 ? c.ycols : c.zcols) = rule_tile[a] / div;
            double t = 0;
            rc = time_axis(w, method, c, a, mem, cap, s, e0, e1, &t);
            if (rc == DWHIP_OK && t < t_best * 0.97) {
                t_best = t;
                pick = c;
            }
        }
        cur = pick;
        sum_best += t_best;
        sum_rule += t_rule;
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    (void)hipFree(mem);
    if (rc != DWHIP_OK) return rc;
    *best = cur;
    if (tuned) *tuned = sum_best;
    if (rule) *rule = sum_rule;
    return DWHIP_OK;
}

}  // namespace

bool dwhip_plan_cached(const int64_t* w, DwhipPlanChoice* pc)
{
    std::lock_guard<std::mutex> lock(g_cache.mu);
    load_locked();
    auto it = g_cache.plans.find(Key{w[0], w[1], w[2]});
    if (it == g_cache.plans.end()) return false;
    *pc = it->second;
    return true;
}

// Tunes w unless a plan for it is cached (used by entry points whose
// params ask for DWHIP_PLAN_MEASURE).  Measuring only chooses among plans
// that all run: when it cannot run (no memory for its trial block beside
// the caller's, a size with nothing to choose), the rule's plan is cached
// for the size and the job goes on with it.
int dwhip_plan_ensure(const int64_t* w, int method)
{
    (void)method;
    if (!dwhip_fast_size_ok(w)) return DWHIP_OK;  // the any-size path has one plan
    DwhipPlanChoice c;
    if (dwhip_plan_cached(w, &c)) return DWHIP_OK;
    const int rc = dwhip_plan_tune(w, 0, nullptr, nullptr);
    if (rc != DWHIP_OK) {
        (void)hipGetLastError();  // a failed trial allocation leaves its error behind
        std::lock_guard<std::mutex> lock(g_cache.mu);
        load_locked();
        g_cache.plans[Key{w[0], w[1], w[2]}] = DwhipPlanChoice();
    }
    return DWHIP_OK;
}

extern "C" int dwhip_plan_tune(const int64_t* work, int retune, double* tuned_ms, double* rule_ms)
{
    if (!work || work[0] < 2 || work[1] < 1 || work[2] < 1) return DWHIP_ERR_ARG;
    if (!dwhip_fast_size_ok(work)) return DWHIP_ERR_SIZE;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    const Key k{work[0], work[1], work[2]};
    if (!retune) {
        DwhipPlanChoice c;
        if (dwhip_plan_cached(work, &c) && !tuned_ms && !rule_ms) return DWHIP_OK;
    }
    DwhipPlanChoice best;
    double tb = 0, tr = 0;
    // SHB's passes; RL transforms the same sizes with the same kernels
    rc = tune_size(k, DWHIP_METHOD_SHB, &best, &tb, &tr);
    if (rc != DWHIP_OK) return rc;
    if (tuned_ms) *tuned_ms = tb;
    if (rule_ms) *rule_ms = tr;
    std::lock_guard<std::mutex> lock(g_cache.mu);
    load_locked();
    g_cache.plans[k] = best;
    save_locked();
    return DWHIP_OK;
}

extern "C" int dwhip_plan_cache_path(const char* path)
{
    std::lock_guard<std::mutex> lock(g_cache.mu);
    g_cache.path = path ? path : default_path();
    g_cache.path_set = true;
    g_cache.loaded = false;
    g_cache.plans.clear();
    load_locked();
    return DWHIP_OK;
}

extern "C" int dwhip_plan_get(const int64_t* work, int* radices, int* tiles)
{
    if (!work || !radices || !tiles) return DWHIP_ERR_ARG;
    if (!dwhip_fast_size_ok(work)) return DWHIP_ERR_SIZE;
    DwhipPlanChoice c;
    const bool cached = dwhip_plan_cached(work, &c);
    DwhipTiles t;
    dwhip_tile_config(work, &t, cached ? &c : nullptr);
    const int len[3] = {(int)(work[0] / 2), (int)work[1], (int)work[2]};
    for (int a = 0; a < 3; ++a) {
        std::vector<int> r = cached && !c.r[a].empty() ? c.r[a] : dwhip_radices(len[a], a);
        for (int i = 0; i < 4; ++i) radices[4 * a + i] = i < (int)r.size() ? r[i] : 0;
    }
    tiles[0] = t.xrows;
    tiles[1] = t.ycols;
    tiles[2] = t.zcols;
    return cached ? 1 : 0;
}
