// Batched and served SHB: many deconvolution jobs of different shapes run
// together.  Each round runs one SHB iteration of every active job, one
// launch per pass over a table of all of them (the pass kernels walk the
// jobs' tiles); jobs are admitted from a FIFO queue while their device
// block fits in the engine's HBM budget, and leave as soon as they finish,
// so new requests join between rounds (continuous batching).  A job's
// arithmetic does not depend on the other jobs in its tables: results
// are the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "shb_kernels.h"

namespace {

#define ECHK(call)                                                        \
    do {                                                                  \
        hipError_t e_ = (call);                                           \
        if (e_ != hipSuccess) return dwhip_hip_fail(e_, #call, __LINE__); \
    } while (0)
#define RCHK(call)                     \
    do {                               \
        int r_ = (call);               \
        if (r_ != DWHIP_OK) return r_; \
    } while (0)

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// First-fit allocator over [0, size) with coalescing.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    size_t used = 0;
    void init(size_t n)
    {
        size = n;
        free_.clear();
        free_[0] = n;
        used = 0;
    }
    bool alloc(size_t n, size_t* off)
    {
        n = align_up(n, 4096);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n;
                const size_t o = it->first + n;
                free_.erase(it);
                if (rest) free_[o] = rest;
                used += n;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align_up(n, 4096);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

struct EngJob {
    DwhipJob j;
    long long serial = 0;    // unique per engine: identifies the job in graph_jobs
    dwhip_params prm;
    int64_t tag = 0;
    float* out = nullptr;
    size_t off = 0;
    bool stop = false;       // REL/ABS: the iterator said stop
    int nrun = 0;            // iterations run
    float last_err = 1, prev_err = 1;
    std::vector<float> err;  // error trace, filled at completion
    float bg = 0;
    double submitted = 0;    // host clock (s) at submission, for aging
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A device table: BatchDev at the start, the JobDev array 64 bytes in.
struct Table {
    char* dev = nullptr;
    char* host = nullptr;  // pinned
    int cap = 0;
    int n = 0;
    size_t lds[KIND_N] = {};
    int ntile[KIND_N] = {};
    std::vector<EngJob*> jobs;
    BatchDev* bd() const { return (BatchDev*)dev; }
    JobDev* jd() const { return (JobDev*)(dev + 64); }
    static size_t bytes(int cap) { return 64 + (size_t)cap * sizeof(JobDev); }
};

}  // namespace

struct dwhip_engine {
    dwhip_engine_config cfg;
    hipStream_t s = nullptr;
    bool own_stream = false;
    hipStream_t s2 = nullptr;      // second group's stream (groups == 2)
    hipEvent_t ev_setup = nullptr;  // set-up of fresh jobs done on s
    char* arena_mem = nullptr;
    Arena arena;
    std::deque<EngJob*> queue;
    std::vector<EngJob*> active;
    std::map<int64_t, EngJob*> done;
    Table tsetup, tweights, tae, tfh, tae2, tfh2;
    // graph of one full round of the current active set (both groups),
    // replayed for the rounds of a step before its last; its tables stay
    // as captured until the active set changes
    Table gae, gae2;
    hipGraphExec_t graph = nullptr;
    std::vector<long long> graph_jobs[2];  // serials of the jobs the graph was captured for
    long long next_serial = 1;
    hipStream_t cs2 = nullptr;     // second branch of a two-group capture
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    long long rounds = 0;
    long long graph_launches = 0;
};

namespace {

int table_alloc(Table* t, int cap)
{
    t->cap = cap;
    ECHK(hipMalloc((void**)&t->dev, Table::bytes(cap)));
    ECHK(hipHostMalloc((void**)&t->host, Table::bytes(cap), hipHostMallocDefault));
    return DWHIP_OK;
}

void table_free(Table* t)
{
    if (t->dev) (void)hipFree(t->dev);
    if (t->host) (void)hipHostFree(t->host);
    t->dev = t->host = nullptr;
}

// Fills t with jobs and queues its upload on s.
// REL/ABS jobs of the fused passes stop on the device: the end-of-round
// kernel applies dw_iterator_next and the passes skip a stopped job, so
// such jobs join multi-round steps; any-size jobs decide on the host
bool dev_stop_job(const EngJob* j)
{
    return j->prm.iter_type != DWHIP_ITER_FIXED && !j->j.generic;
}

int table_build(Table* t, const std::vector<EngJob*>& jobs, hipStream_t s)
{
    if ((int)jobs.size() > t->cap) return DWHIP_ERR_ARG;
    t->jobs = jobs;
    t->n = (int)jobs.size();
    BatchDev* bh = (BatchDev*)t->host;
    JobDev* jh = (JobDev*)(t->host + 64);
    memset(bh, 0, sizeof(*bh));
    bh->njobs = t->n;
    for (int k = 0; k < KIND_N; ++k) {
        t->ntile[k] = 0;
        t->lds[k] = 0;
    }
    for (int i = 0; i < t->n; ++i) {
        jh[i] = jobs[i]->j.d;
        jh[i].dev_stop = dev_stop_job(jobs[i]) ? 1 : 0;
        for (int k = 0; k < KIND_N; ++k) {
            jh[i].tile0[k] = t->ntile[k];
            t->ntile[k] += jh[i].ntile[k];
            t->lds[k] = std::max(t->lds[k], jobs[i]->j.tiles.lds[k]);
        }
    }
    for (int k = 0; k < KIND_N; ++k) bh->ntile[k] = t->ntile[k];
    if (t->n == 0) return DWHIP_OK;
    ECHK(hipMemcpyAsync(t->dev, t->host, Table::bytes(t->n), hipMemcpyHostToDevice, s));
    return DWHIP_OK;
}

// One pass over a table, with the profiling hooks of the per-image path.
int table_pass(const Table& t, bool xkind, int mode, hipStream_t s)
{
    const int kind = xkind ? xmode_kind(mode) : cmode_kind(mode);
    if (t.n == 0 || t.ntile[kind] == 0) return DWHIP_OK;
    const int slot = dwhip_prof_begin(s);
    hipError_t e = xkind ? dwhip_launch_xpass(mode, t.bd(), t.jd(), t.ntile[kind], t.lds[kind], s)
                         : dwhip_launch_colpass(mode, t.bd(), t.jd(), t.ntile[kind], t.lds[kind], s);
    if (slot >= 0) {
        double by = 0;
        for (EngJob* j : t.jobs) by += dwhip_pass_bytes(xkind ? 1 : 0, mode, j->j.d, j->j.host_iter);
        dwhip_prof_end(s, slot, (xkind ? 0 : 7) + mode, by);
    }
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, xkind ? "batched x pass" : "batched column pass", __LINE__);
}

int iter_target(const EngJob* j)
{
    return j->prm.iter_type == DWHIP_ITER_FIXED ? j->prm.niter : j->prm.maxiter;
}

// dw_iterator_next (src/dw.c:235-268) before iteration `iter`
bool keep_going(const EngJob* j, int iter)
{
    const dwhip_params* p = &j->prm;
    if (p->iter_type == DWHIP_ITER_REL) {
        float last = j->prev_err;
        if (iter < 2) last = 2 * j->last_err * p->err_rel;
        // C's fabs: the float difference, then double arithmetic
        if (fabs((double)(j->last_err - last)) / (double)j->last_err < (double)p->err_rel) return false;
    } else if (p->iter_type == DWHIP_ITER_ABS) {
        if (iter > 0 && j->last_err < p->err_abs) return false;
    }
    return iter < iter_target(j);
}

double job_voxels(const EngJob* j)
{
    return (double)j->j.work[0] * (double)j->j.work[1] * (double)j->j.work[2];
}

double job_work(const EngJob* j)
{
    return job_voxels(j) * (double)iter_target(j);
}

// Admission: while the active set has room (max_active, the voxel cap and
// the arena), the next job in the configured order that fits.  FIFO stops
// at the first job that does not fit, so later jobs never overtake it.
// Shortest-first takes the least work among those that fit, except that a
// job that has waited aging_ms or more goes before every younger job
// (oldest first) and, when it does not fit yet, keeps younger jobs from
// taking the room the running jobs free (aging_ms 0: least work first
// only).  The bound keeps a large job from being passed over for ever.
int admit(dwhip_engine* E, std::vector<EngJob*>& fresh)
{
    const double cap = (double)E->cfg.active_mvox * 1e6;
    double vox = 0;
    for (EngJob* j : E->active) vox += job_voxels(j);
    const double t = now_s();
    const double aging = E->cfg.aging_ms > 0 ? 1e-3 * E->cfg.aging_ms : 0;
    auto aged = [&](const EngJob* j) { return aging > 0 && t - j->submitted >= aging; };
    while (!E->queue.empty() && (int)E->active.size() < E->cfg.max_active) {
        std::vector<size_t> order(E->queue.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        if (E->cfg.order == DWHIP_ORDER_SHORTEST)
            std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
                const EngJob *ja = E->queue[a], *jb = E->queue[b];
                const bool oa = aged(ja), ob = aged(jb);
                if (oa != ob) return oa;                              // aged jobs first
                if (oa) return ja->submitted < jb->submitted;         // oldest first among them
                return job_work(ja) < job_work(jb);                   // then least work
            });
        bool taken = false;
        for (size_t i : order) {
            EngJob* j = E->queue[i];
            const bool fits_cap = cap <= 0 || E->active.empty() || vox + job_voxels(j) <= cap;
            size_t off = 0;
            if (fits_cap && E->arena.alloc(j->j.bytes, &off)) {
                E->queue.erase(E->queue.begin() + i);
                j->off = off;
                RCHK(dwhip_job_bind(&j->j, E->arena_mem + off, E->s));
                E->active.push_back(j);
                fresh.push_back(j);
                vox += job_voxels(j);
                taken = true;
                break;
            }
            if (E->cfg.order == DWHIP_ORDER_FIFO || aged(j)) break;  // nobody overtakes it
        }
        if (!taken) break;
    }
    return DWHIP_OK;
}

// Set-up of newly admitted jobs (method_shb.c:108-294), batched by table.
int setup(dwhip_engine* E, const std::vector<EngJob*>& fresh)
{
    if (fresh.empty()) return DWHIP_OK;
    hipStream_t s = E->s;
    // jobs on the any-size path set up one by one; the others batched
    std::vector<EngJob*> fast;
    for (EngJob* j : fresh) {
        if (j->j.generic)
            RCHK(dwhip_job_setup(&j->j, s));
        else
            fast.push_back(j);
    }
    if (fast.empty()) return DWHIP_OK;
    RCHK(table_build(&E->tsetup, fast, s));
    for (int i = 0; i < (int)fast.size(); ++i) {
        ECHK(dwhip_launch_stats(E->tsetup.jd(), i, DWHIP_STAT_BLOCKS, s));
        ECHK(dwhip_launch_start(E->tsetup.jd(), i, 1024, s));
    }
    RCHK(table_pass(E->tsetup, true, XM_PSF, s));
    RCHK(table_pass(E->tsetup, false, CM_YFWD, s));
    RCHK(table_pass(E->tsetup, false, CM_ZOTF, s));
    std::vector<EngJob*> wj;
    for (EngJob* j : fast)
        if (j->j.d.has_w) wj.push_back(j);
    if (!wj.empty()) {
        RCHK(table_build(&E->tweights, wj, s));
        RCHK(table_pass(E->tweights, true, XM_ONES, s));
        RCHK(table_pass(E->tweights, false, CM_YFWD, s));
        RCHK(table_pass(E->tweights, false, CM_ZCONVJ, s));
        RCHK(table_pass(E->tweights, false, CM_YINV, s));
        RCHK(table_pass(E->tweights, true, XM_WEIGHTS, s));
    }
    return DWHIP_OK;
}

}  // namespace

extern "C" {

void dwhip_engine_config_default(dwhip_engine_config* c)
{
    memset(c, 0, sizeof(*c));
    c->hbm_budget = (size_t)32 << 30;
    c->max_active = 64;
    c->groups = 1;
    c->rounds_per_step = 8;
}

int dwhip_engine_create(const dwhip_engine_config* cfg, void* stream, dwhip_engine** out)
{
    if (!cfg || !out || cfg->hbm_budget < ((size_t)1 << 20) || cfg->max_active < 1 || cfg->max_active > 4096)
        return DWHIP_ERR_ARG;
    if ((cfg->order != DWHIP_ORDER_FIFO && cfg->order != DWHIP_ORDER_SHORTEST) || cfg->active_mvox < 0)
        return DWHIP_ERR_ARG;
    if (cfg->groups != 0 && cfg->groups != 1 && cfg->groups != 2) return DWHIP_ERR_ARG;  // 0 reads as 1
    if (cfg->rounds_per_step < 0 || cfg->aging_ms < 0) return DWHIP_ERR_ARG;             // 0 reads as 1
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    dwhip_engine* E = new dwhip_engine();
    E->cfg = *cfg;
    if (stream) {
        E->s = (hipStream_t)stream;
    } else {
        if (hipStreamCreate(&E->s) != hipSuccess) {
            delete E;
            return DWHIP_ERR_HIP;
        }
        E->own_stream = true;
    }
    hipError_t e = hipMalloc((void**)&E->arena_mem, cfg->hbm_budget);
    if (e != hipSuccess) {
        if (E->own_stream) (void)hipStreamDestroy(E->s);
        delete E;
        return e == hipErrorOutOfMemory ? DWHIP_ERR_NOMEM : DWHIP_ERR_HIP;
    }
    E->arena.init(cfg->hbm_budget);
    const int cap = cfg->max_active;
    if (table_alloc(&E->tsetup, cap) || table_alloc(&E->tweights, cap) || table_alloc(&E->tae, cap) ||
        table_alloc(&E->tfh, cap) || table_alloc(&E->gae, cap)) {
        dwhip_engine_destroy(E);
        return DWHIP_ERR_NOMEM;
    }
    if (hipEventCreateWithFlags(&E->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&E->ev_join, hipEventDisableTiming) != hipSuccess) {
        dwhip_engine_destroy(E);
        return DWHIP_ERR_HIP;
    }
    if (cfg->groups == 2) {
        if (table_alloc(&E->tae2, cap) || table_alloc(&E->tfh2, cap) || table_alloc(&E->gae2, cap)) {
            dwhip_engine_destroy(E);
            return DWHIP_ERR_NOMEM;
        }
        if (hipStreamCreateWithFlags(&E->cs2, hipStreamNonBlocking) != hipSuccess) {
            dwhip_engine_destroy(E);
            return DWHIP_ERR_HIP;
        }
        // default priority, like every stream here
        if (hipStreamCreateWithFlags(&E->s2, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&E->ev_setup, hipEventDisableTiming) != hipSuccess) {
            dwhip_engine_destroy(E);
            return DWHIP_ERR_HIP;
        }
    }
    *out = E;
    return DWHIP_OK;
}

void dwhip_engine_destroy(dwhip_engine* E)
{
    if (!E) return;
    (void)hipStreamSynchronize(E->s);
    if (E->s2) (void)hipStreamSynchronize(E->s2);
    for (EngJob* j : E->queue) delete j;
    for (EngJob* j : E->active) delete j;
    for (auto& kv : E->done) delete kv.second;
    table_free(&E->tsetup);
    table_free(&E->tweights);
    table_free(&E->tae);
    table_free(&E->tfh);
    table_free(&E->tae2);
    table_free(&E->tfh2);
    table_free(&E->gae);
    table_free(&E->gae2);
    if (E->graph) (void)hipGraphExecDestroy(E->graph);
    if (E->cs2) (void)hipStreamDestroy(E->cs2);
    if (E->ev_fork) (void)hipEventDestroy(E->ev_fork);
    if (E->ev_join) (void)hipEventDestroy(E->ev_join);
    if (E->ev_setup) (void)hipEventDestroy(E->ev_setup);
    if (E->s2) (void)hipStreamDestroy(E->s2);
    if (E->arena_mem) (void)hipFree(E->arena_mem);
    if (E->own_stream) (void)hipStreamDestroy(E->s);
    delete E;
}

int dwhip_engine_submit(dwhip_engine* E, const float* im, int64_t M, int64_t N, int64_t P, const float* psf,
                        int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm, float* out, int64_t tag)
{
    if (!E || !im || !psf || !prm || !out) return DWHIP_ERR_ARG;
    if (M < 1 || N < 1 || P < 1 || pM < 1 || pN < 1 || pP < 1) return DWHIP_ERR_ARG;
    if (E->done.count(tag)) return DWHIP_ERR_ARG;
    for (EngJob* j : E->queue)
        if (j->tag == tag) return DWHIP_ERR_ARG;
    for (EngJob* j : E->active)
        if (j->tag == tag) return DWHIP_ERR_ARG;
    const bool fixed = prm->iter_type == DWHIP_ITER_FIXED;
    if (prm->niter < 1 || (!fixed && prm->maxiter < 1)) return DWHIP_ERR_ARG;
    if (prm->method != DWHIP_METHOD_SHB && prm->method != DWHIP_METHOD_RL) return DWHIP_ERR_ARG;
    if (prm->planning == DWHIP_PLAN_MEASURE) {
        // the transform size's measured plan, on its first use
        int64_t w[3], ref[3];
        if (dwhip_work_size(M, N, P, pM, pN, pP, prm->border_quality, w, ref) == DWHIP_OK) {
            for (int d = 0; d < 3; ++d)
                if (prm->work[d] > 0) w[d] = prm->work[d];
            const int trc = dwhip_plan_ensure(w, prm->method);
            if (trc != DWHIP_OK) return trc;
        }
    }
    EngJob* j = new EngJob();
    j->serial = E->next_serial++;
    j->submitted = now_s();
    j->prm = *prm;
    j->tag = tag;
    j->out = out;
    const float* start = prm->start != DWHIP_START_FLAT ? prm->start_image : nullptr;
    int rc = dwhip_job_plan(&j->j, im, M, N, P, psf, pM, pN, pP, prm, start);
    if (rc == DWHIP_OK && align_up(j->j.bytes, 4096) > E->cfg.hbm_budget) rc = DWHIP_ERR_NOMEM;
    if (rc != DWHIP_OK) {
        delete j;
        return rc;
    }
    E->queue.push_back(j);
    return DWHIP_OK;
}

// Completed jobs (their crops queued and the streams synchronised): error
// trace and bg read back, memory returned to the arena, tags reported.
static int finish_jobs(dwhip_engine* E, const std::vector<EngJob*>& finished, int64_t* done_tags, int max_done,
                       int* n_done)
{
    for (EngJob* j : finished) {
        j->err.resize(j->nrun);
        if (j->nrun > 0)
            ECHK(hipMemcpy(j->err.data(), j->j.d.err, j->nrun * sizeof(float), hipMemcpyDeviceToHost));
        JobState hs;
        ECHK(hipMemcpy(&hs, j->j.d.st, sizeof(hs), hipMemcpyDeviceToHost));
        j->bg = hs.bg;
        E->arena.release(j->off, j->j.bytes);
        j->j.mem = nullptr;
        E->done[j->tag] = j;
        if (done_tags && n_done && *n_done < max_done) done_tags[(*n_done)++] = j->tag;
    }
    return DWHIP_OK;
}

// Splits the active jobs into two groups of about equal transform voxels
// (largest first onto the lighter group); group membership keeps the
// active order within each group.
static void split_groups(const std::vector<EngJob*>& act, std::vector<EngJob*>* g)
{
    g[0].clear();
    g[1].clear();
    std::vector<size_t> idx(act.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](size_t x, size_t y) { return job_voxels(act[x]) > job_voxels(act[y]); });
    std::vector<int> side(act.size(), 0);
    double load[2] = {0, 0};
    for (size_t i : idx) {
        const int k = load[1] < load[0] ? 1 : 0;
        side[i] = k;
        load[k] += job_voxels(act[i]);
    }
    for (size_t i = 0; i < act.size(); ++i) g[side[i]].push_back(act[i]);
}

// One iteration of every job of a group on stream s; the adjoint half
// (F, G, H) only for jobs with a further iteration to feed.
static int group_round(Table* tae, Table* tfh, const std::vector<EngJob*>& all, hipStream_t s)
{
    if (all.empty()) return DWHIP_OK;
    std::vector<EngJob*> jobs, fh;
    for (EngJob* j : all) {
        const bool fixed = j->prm.iter_type == DWHIP_ITER_FIXED;
        const bool full = !fixed || j->j.d.rl || j->j.host_iter + 1 < j->prm.niter;
        if (j->j.generic) {
            // the any-size path runs its own passes, one job at a time
            RCHK(dwhip_generic_round(&j->j, full, s));
            continue;
        }
        jobs.push_back(j);
        if (full) fh.push_back(j);
    }
    if (jobs.empty()) return DWHIP_OK;
    RCHK(table_build(tae, jobs, s));
    RCHK(table_build(tfh, fh, s));
    RCHK(table_pass(*tae, true, XM_UPDATE, s));
    RCHK(table_pass(*tae, false, CM_YFWD, s));
    RCHK(table_pass(*tae, false, CM_ZCONV, s));
    RCHK(table_pass(*tae, false, CM_YINV_IMG, s));
    RCHK(table_pass(*tae, true, XM_RATIO, s));
    RCHK(table_pass(*tfh, false, CM_YFWD_IMG, s));
    RCHK(table_pass(*tfh, false, CM_ZCONVJ, s));
    RCHK(table_pass(*tfh, false, CM_YINV, s));
    ECHK(dwhip_launch_round_end(tae->bd(), tae->jd(), tae->n, 1, s));
    return DWHIP_OK;
}

// The passes of one full round (every job transforms its adjoint half) over
// table t, which serves as both the A-E and the F-H table.
static int full_round(const Table& t, hipStream_t s)
{
    RCHK(table_pass(t, true, XM_UPDATE, s));
    RCHK(table_pass(t, false, CM_YFWD, s));
    RCHK(table_pass(t, false, CM_ZCONV, s));
    RCHK(table_pass(t, false, CM_YINV_IMG, s));
    RCHK(table_pass(t, true, XM_RATIO, s));
    RCHK(table_pass(t, false, CM_YFWD_IMG, s));
    RCHK(table_pass(t, false, CM_ZCONVJ, s));
    RCHK(table_pass(t, false, CM_YINV, s));
    ECHK(dwhip_launch_round_end(t.bd(), t.jd(), t.n, 1, s));
    return DWHIP_OK;
}

// Capture body: group 0 on the capture stream, group 1 (if any) on a
// branch forked from it and joined back.
static int capture_round(void* ctx, hipStream_t cs)
{
    dwhip_engine* E = (dwhip_engine*)ctx;
    const bool two = E->gae2.n > 0;
    if (two) {
        ECHK(hipEventRecord(E->ev_fork, cs));
        ECHK(hipStreamWaitEvent(E->cs2, E->ev_fork, 0));
    }
    RCHK(full_round(E->gae, cs));
    if (two) {
        RCHK(full_round(E->gae2, E->cs2));
        ECHK(hipEventRecord(E->ev_join, E->cs2));
        ECHK(hipStreamWaitEvent(cs, E->ev_join, 0));
    }
    return DWHIP_OK;
}

// Rounds this step runs: up to the next known completion (the fewest
// rounds any job has left before nIter, or maxiter for REL/ABS jobs, which
// may stop earlier on the device), at most rounds_per_step; one round with
// an any-size job active (its REL/ABS rule runs on the host after each
// round) and while kernels are profiled (events per pass).
static int step_rounds(const dwhip_engine* E)
{
    const int cap = std::max(1, E->cfg.rounds_per_step);
    if (cap == 1 || dwhip_prof_on() || dwhip_graph_mode() == 0) return 1;
    int r = cap;
    for (const EngJob* j : E->active) {
        if (j->j.generic) return 1;
        r = std::min(r, iter_target(j) - j->j.host_iter);
    }
    return std::max(1, r);
}

// Replays `n` full rounds of the active groups on E->s from one captured
// round, capturing it again when the groups changed.
static int graph_rounds(dwhip_engine* E, const std::vector<EngJob*>* grp, int n)
{
    if (n <= 0) return DWHIP_OK;
    std::vector<long long> ids[2];
    for (int g = 0; g < 2; ++g)
        for (const EngJob* j : grp[g]) ids[g].push_back(j->serial);
    if (!E->graph || ids[0] != E->graph_jobs[0] || ids[1] != E->graph_jobs[1]) {
        if (E->graph) {
            ECHK(hipStreamSynchronize(E->s));  // a previous step's replays may still read gae
            (void)hipGraphExecDestroy(E->graph);
            E->graph = nullptr;
        }
        RCHK(table_build(&E->gae, grp[0], E->s));
        if (E->gae2.cap) RCHK(table_build(&E->gae2, grp[1], E->s));
        RCHK(dwhip_capture(capture_round, E, &E->graph));
        E->graph_jobs[0] = ids[0];
        E->graph_jobs[1] = ids[1];
    }
    for (int k = 0; k < n; ++k) ECHK(hipGraphLaunch(E->graph, E->s));
    E->graph_launches += n;
    for (int g = 0; g < 2; ++g)
        for (EngJob* j : grp[g]) j->j.host_iter += n;
    E->rounds += n;
    return DWHIP_OK;
}

int dwhip_engine_step(dwhip_engine* E, int64_t* done_tags, int max_done, int* n_done)
{
    if (!E) return DWHIP_ERR_ARG;
    if (n_done) *n_done = 0;
    hipStream_t s = E->s;
    std::vector<EngJob*> fresh;
    RCHK(admit(E, fresh));
    RCHK(setup(E, fresh));
    // REL/ABS jobs whose stopping rule ends them before iteration 0 (REL with
    // err_rel in (1/3, 1)): no round, the start guess is their result
    std::vector<EngJob*> finished;
    {
        std::vector<EngJob*> run;
        for (EngJob* j : E->active) {
            if (j->nrun == 0 && j->prm.iter_type != DWHIP_ITER_FIXED && !keep_going(j, 0)) {
                ECHK(dwhip_launch_crop(j->j.jd, 0, dwhip_result_iterate(j->j.d.rl != 0, 0), j->out, 1024, s));
                finished.push_back(j);
            } else {
                run.push_back(j);
            }
        }
        E->active = run;
    }
    if (E->active.empty()) {
        ECHK(hipStreamSynchronize(s));
        return finish_jobs(E, finished, done_tags, max_done, n_done);
    }
    // the jobs of this step, in one group on s or two on s and s2
    std::vector<EngJob*> grp[2];
    const bool two = E->s2 && E->active.size() >= 2;
    if (two)
        split_groups(E->active, grp);
    else
        grp[0] = E->active;
    // all rounds but the step's last from the round graph, on s
    RCHK(graph_rounds(E, grp, step_rounds(E) - 1));
    if (two) {
        // the second group's jobs may have been set up (or replayed) on s
        ECHK(hipEventRecord(E->ev_setup, s));
        ECHK(hipStreamWaitEvent(E->s2, E->ev_setup, 0));
    }
    hipStream_t gs[2] = {s, two ? E->s2 : s};
    RCHK(group_round(&E->tae, &E->tfh, grp[0], gs[0]));
    if (two) RCHK(group_round(&E->tae2, &E->tfh2, grp[1], gs[1]));
    std::map<EngJob*, hipStream_t> js;  // each job's stream this round
    for (int g = 0; g < 2; ++g)
        for (EngJob* j : grp[g]) js[j] = gs[g];
    for (EngJob* j : E->active) {
        j->j.host_iter += 1;
        j->nrun = j->j.host_iter;
    }
    E->rounds += 1;
    // REL/ABS jobs: the rounds run and the stop flag (fused passes), or the
    // error of this round (any-size jobs, whose rule runs here)
    std::vector<float> errs(E->active.size(), 0.f);
    std::vector<int> iter_stop(2 * E->active.size(), 0);
    bool need_err = false;
    for (size_t i = 0; i < E->active.size(); ++i) {
        EngJob* j = E->active[i];
        if (dev_stop_job(j)) {
            ECHK(hipMemcpyAsync(&iter_stop[2 * i], &j->j.d.st->iter, 2 * sizeof(int), hipMemcpyDeviceToHost, js[j]));
            need_err = true;
        } else if (j->prm.iter_type != DWHIP_ITER_FIXED) {
            ECHK(hipMemcpyAsync(&errs[i], j->j.d.err + (j->nrun - 1), sizeof(float), hipMemcpyDeviceToHost, js[j]));
            need_err = true;
        }
    }
    if (need_err) {
        ECHK(hipStreamSynchronize(s));
        if (two) ECHK(hipStreamSynchronize(E->s2));
    }
    // finished jobs: crop the result into out
    std::vector<EngJob*> still;
    for (size_t i = 0; i < E->active.size(); ++i) {
        EngJob* j = E->active[i];
        bool fin;
        if (j->prm.iter_type == DWHIP_ITER_FIXED) {
            fin = j->nrun >= j->prm.niter;
        } else if (dev_stop_job(j)) {
            j->j.host_iter = j->nrun = iter_stop[2 * i];
            fin = iter_stop[2 * i + 1] != 0;
        } else {
            j->prev_err = j->last_err;
            j->last_err = errs[i];
            fin = !keep_going(j, j->nrun);
        }
        if (fin) {
            // SHB returns x_{n-1}; RL returns x_n, after one more update pass
            if (j->j.d.rl) RCHK(dwhip_job_final_update(&j->j, j->j.bd, j->j.jd, js[j]));
            ECHK(dwhip_launch_crop(j->j.jd, 0, dwhip_result_iterate(j->j.d.rl != 0, j->nrun), j->out, 1024, js[j]));
            finished.push_back(j);
        } else {
            still.push_back(j);
        }
    }
    ECHK(hipStreamSynchronize(s));
    if (two) ECHK(hipStreamSynchronize(E->s2));
    if (dwhip_prof_on()) dwhip_prof_collect();
    E->active = still;
    return finish_jobs(E, finished, done_tags, max_done, n_done);
}

int dwhip_engine_status(const dwhip_engine* E, int* queued, int* active, size_t* bytes_in_use)
{
    if (!E) return DWHIP_ERR_ARG;
    if (queued) *queued = (int)E->queue.size();
    if (active) *active = (int)E->active.size();
    if (bytes_in_use) *bytes_in_use = E->arena.used;
    return DWHIP_OK;
}

int dwhip_engine_result(dwhip_engine* E, int64_t tag, dwhip_stats* st, int release)
{
    if (!E) return DWHIP_ERR_ARG;
    auto it = E->done.find(tag);
    if (it == E->done.end()) return DWHIP_ERR_STATE;
    EngJob* j = it->second;
    if (st) {
        st->bg = j->bg;
        st->iterations = j->nrun;
        for (int d = 0; d < 3; ++d) {
            st->work[d] = j->j.work[d];
            st->ref_work[d] = j->j.ref_work[d];
        }
        st->setup_ms = 0;
        st->iter_ms = 0;
        if (st->err_trace && st->err_trace_len > 0) {
            const int n = std::min((int)j->err.size(), st->err_trace_len);
            memcpy(st->err_trace, j->err.data(), n * sizeof(float));
        }
    }
    if (release) {
        delete j;
        E->done.erase(it);
    }
    return DWHIP_OK;
}

}  // extern "C"
