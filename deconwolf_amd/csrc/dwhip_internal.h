// Internal host-side types of libdwhip shared by the per-image driver and
// the batch/serve engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "dwhip.h"
#include "generic_dev.h"
#include "shb_dev.h"

#define DWHIP_STAT_BLOCKS 256

struct DwhipTiles {
    int xrows, ycols, zcols;
    int ntile[KIND_N];
    size_t lds[KIND_N];
};

// A plan of the fused passes for one transform size: forward radices of the
// x-row (length nx/2), y and z transforms and the tile shapes; empty radix
// lists and zero tiles stand for the rule planner's choice.
struct DwhipPlanChoice {
    std::vector<int> r[3];
    int xrows = 0, ycols = 0, zcols = 0;
};

// Byte offsets of a job's buffers inside its device block.
struct DwhipLayout {
    size_t C, K, U0, U1, W, part, stat, err, errd, tw, st, tab;
};

// One job: host copy of its descriptor plus its device block, which also
// holds a one-job table (bd, jd) for the per-image path.
struct DwhipJob {
    JobDev d;
    DwhipTiles tiles;
    DwhipLayout lay;
    int64_t work[3];
    int64_t ref_work[3];
    void* mem = nullptr;
    size_t bytes = 0;
    BatchDev* bd = nullptr;
    JobDev* jd = nullptr;
    int host_iter = 0;        // iteration the next round runs (host copy of st->iter)
    std::vector<float2> tw;   // twiddles, uploaded by dwhip_job_bind
    size_t rtw_off = 0;
    BatchDev bh;              // host copies of the one-job table being uploaded
    JobDev dcopy;
    // any-size path (generic.cpp): per-axis FFTs of any length; C holds
    // the work volume Z (half spectra, or real rows in place), K the OTF
    bool generic = false;
    GenDev g;                 // host copy
    GenDev* gd = nullptr;     // device copy (in the job block)
    // plan to use instead of the cached or rule plan (the autotuner's candidates)
    const DwhipPlanChoice* force_plan = nullptr;
};

int dwhip_hip_fail(hipError_t e, const char* what, int line);

// The iterate a method returns after nrun iterations: SHB x_{n-1} (swap-back,
// method_shb.c:385-390), RL x_n; the start guess x_0 when nothing ran.
inline int dwhip_result_iterate(bool rl, int nrun)
{
    if (nrun <= 0) return 0;
    return rl ? nrun : nrun - 1;
}

// profiling of pass launches (profile.cpp); kernel ids: X modes 0..6, then
// column modes 7..15
#define DWHIP_PROF_KERNELS 16
bool dwhip_prof_on();
int dwhip_prof_begin(hipStream_t s);
void dwhip_prof_end(hipStream_t s, int slot, int kernel, double bytes);
void dwhip_prof_collect();
double dwhip_pass_bytes(int xkind, int mode, const JobDev& d, int iter);
int dwhip_make_plan(int n, const std::vector<int>& radices, FftPlanDev* p, std::vector<float2>& tw);
std::vector<int> dwhip_radices(int n, int axis);
void dwhip_tile_config(const int64_t* w, DwhipTiles* t, const DwhipPlanChoice* pc = nullptr);
// plan cache and measured planning (tune.cpp)
bool dwhip_plan_cached(const int64_t* w, DwhipPlanChoice* pc);
int dwhip_plan_ensure(const int64_t* w, int method);
int dwhip_radix_candidates(int n, int axis, std::vector<std::vector<int>>* out);
int dwhip_job_plan(DwhipJob* j, const float* im_dev, int64_t M, int64_t N, int64_t P, const float* psf_dev,
                   int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm, const float* start_dev);
int dwhip_job_bind(DwhipJob* j, void* mem, hipStream_t s);
int dwhip_job_create(DwhipJob* j, const float* im_dev, int64_t M, int64_t N, int64_t P, const float* psf_dev,
                     int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm, const float* start_dev);
int dwhip_pass_x(int mode, const DwhipJob* j, const BatchDev* bd, const JobDev* jd, int ntile, size_t lds,
                 hipStream_t s);
int dwhip_pass_c(int mode, const DwhipJob* j, const BatchDev* bd, const JobDev* jd, int ntile, size_t lds,
                 hipStream_t s);
void dwhip_job_free(DwhipJob* j);
int dwhip_job_setup(const DwhipJob* j, hipStream_t s);
int dwhip_job_round(DwhipJob* j, bool full, hipStream_t s);
int dwhip_job_final_update(DwhipJob* j, const BatchDev* bd, const JobDev* jd, hipStream_t s);
// The per-image method on device pointers; mem/cap: the job's device block
// when the caller provides it (nothing allocated then), nullptr to allocate.
int dwhip_run_device_in(const float* im, int64_t M, int64_t N, int64_t P, const float* psf, int64_t pM,
                        int64_t pN, int64_t pP, const dwhip_params* prm, float* out, dwhip_stats* st,
                        void* stream, void* mem, size_t cap);

// HIP graphs of iteration rounds (graph.cpp).  Mode, from $DWHIP_GRAPH
// (read once): 0 launches every pass directly; 1 (default) captures one
// round as a graph and replays it; 2 captures a job's whole FIXED loop.
int dwhip_graph_mode();
// Captures what body(stream) queues on a library-owned stream (non-blocking,
// default priority, one per host thread) and instantiates it into *exec.
int dwhip_capture(int (*body)(void* ctx, hipStream_t cs), void* ctx, hipGraphExec_t* exec);

// any-size path (generic.cpp)
bool dwhip_generic_size_ok(const int64_t* w);
bool dwhip_fast_size_ok(const int64_t* w);
int dwhip_generic_plan(DwhipJob* j);
int dwhip_generic_bind(DwhipJob* j, char* gdev, hipStream_t s);
int dwhip_generic_setup(const DwhipJob* j, hipStream_t s);
int dwhip_generic_round(DwhipJob* j, bool full, hipStream_t s);
int dwhip_generic_final_update(DwhipJob* j, hipStream_t s);
int dwhip_generic_ratio_grid();
