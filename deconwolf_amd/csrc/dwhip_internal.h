// Internal host-side types of libdwhip shared by the per-image driver and
// the batch/serve engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "dwhip.h"
#include "shb_dev.h"

#define DWHIP_STAT_BLOCKS 256

struct DwhipTiles {
    int xrows, ycols, zcols;
    int ntile[KIND_N];
    size_t lds[KIND_N];
};

// One job: host copy of its descriptor plus its device allocation, which
// also holds a one-job table (bd, jd) for the per-image path.
struct DwhipJob {
    JobDev d;
    DwhipTiles tiles;
    int64_t work[3];
    int64_t ref_work[3];
    void* mem;
    size_t bytes;
    BatchDev* bd;
    JobDev* jd;
    int host_iter;   // iteration the next round runs (host copy of st->iter)
};

int dwhip_hip_fail(hipError_t e, const char* what, int line);

// profiling of pass launches (profile.cpp); kernel ids: X modes 0..6, then
// column modes 7..13
#define DWHIP_PROF_KERNELS 14
bool dwhip_prof_on();
int dwhip_prof_begin(hipStream_t s);
void dwhip_prof_end(hipStream_t s, int slot, int kernel, double bytes);
void dwhip_prof_collect();
double dwhip_pass_bytes(int xkind, int mode, const JobDev& d, int iter);
int dwhip_make_plan(int n, const std::vector<int>& radices, FftPlanDev* p, std::vector<float2>& tw);
std::vector<int> dwhip_radices(int n, bool rows);
void dwhip_tile_config(const int64_t* w, DwhipTiles* t);
int dwhip_job_create(DwhipJob* j, const float* im_dev, int64_t M, int64_t N, int64_t P, const float* psf_dev,
                     int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm, const float* start_dev);
void dwhip_job_free(DwhipJob* j);
int dwhip_job_setup(const DwhipJob* j, hipStream_t s);
int dwhip_job_round(DwhipJob* j, bool full, hipStream_t s);
