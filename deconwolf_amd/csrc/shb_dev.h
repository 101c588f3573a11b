// Device-side job table of the SHB engine.
//
// One JobDev describes one deconvolution job (a whole image or one tile).
// Every pass kernel receives the table and walks a shared tile space: the
// tiles of job j of pass kind k are [tile0[k], tile0[k] + ntile[k]).  A plain
// per-image run is a table with one job; the batched/served engine puts many
// jobs of different shapes in one table, so one launch per pass serves all
// of them.
#pragma once
#include <stdint.h>
#include "fft_engine.h"

// Tile spaces: x-row tiles, y column tiles of all slabs, z column tiles,
// and the tiles of the pruned passes — KIND_YD: slabs z < P (pass D);
// KIND_YF: pass F's slabs (z < P for SHB, all of them for RL); KIND_XR:
// pass E's rows (for SHB the rows y < N of the planes z < P, in tiles that
// stay within one plane; for RL all x tiles).  A pruned pass enumerates only
// the tiles it works on, so its grid carries no idle workgroups and every
// XCD gets a share of the work.
enum { KIND_X = 0, KIND_Y = 1, KIND_Z = 2, KIND_YD = 3, KIND_YF = 4, KIND_XR = 5, KIND_N = 6 };

// metric / start values mirror dw_metric and dw_start_condition (src/dw.h:64-77)
enum { DWHIP_DEV_METRIC_MSE = 0, DWHIP_DEV_METRIC_IDIV = 1 };
// stopping rules, the values of dw_iter_type (src/dw.h) and DWHIP_ITER_*
enum { DWHIP_DEV_ITER_ABS = 0, DWHIP_DEV_ITER_REL = 1, DWHIP_DEV_ITER_FIXED = 2 };

// State written on the device: by the setup statistics kernel (bg, c0,
// mid) and by the end-of-round kernel (iter, stop).  JobDev copies in
// several tables share it through the pointer.
struct JobState {
    float bg;           // background level (method_shb.c:108-117)
    float c0;           // flat start constant (method_shb.c:255-268)
    int mid[3];         // PSF argmax (method_shb.c:195)
    int iter;           // current iteration (rounds run)
    int stop;           // 1: the stopping rule ended the job (dev_stop tables)
    int pad;
};

struct JobDev {
    // transform (work) size; h = nx/2, hc = h+1 complex per x-row, hcs = row stride
    int nx, ny, nz, h, hc, hcs;
    // image and PSF sizes
    int M, N, P;
    int pM, pN, pP;
    // reference work size volume, divides the flat start (method_shb.c:266)
    double refMNP;
    // tile bookkeeping per pass kind
    int ntile[KIND_N];
    int tile0[KIND_N];
    int xrows;          // x-rows per X tile
    int ycols, zcols;   // columns per Y / Z tile
    // options
    int positivity;
    int metric;
    int has_w;          // Bertero weights in use (bq > 0)
    int bg_auto;
    int start_identity; // start from start_img (or im) inserted, instead of flat
    int niter;          // iterations requested (nIter, or maxiter for REL/ABS)
    int rl;             // 1: Richardson-Lucy (method_rl.c), 0: SHB (method_shb.c)
    // stopping rule (dw_iterator_next, src/dw.c:237-269)
    int iter_type;      // DWHIP_DEV_ITER_*
    float err_rel, err_abs;
    // 1 in the served engine's round tables: the end-of-round kernel applies
    // the stopping rule and sets st->stop, and the passes skip the tiles of
    // a stopped job (a job's own table keeps 0)
    int dev_stop;
    float alphamax;
    float bg_user;
    JobState* st;       // device-written state of this job
    // buffers
    float2* C;          // complex work volume, [nz][ny][hcs]
    float2* K;          // OTF / (nx*ny*nz), [nz][ny][hcs]
    float* U0;          // x_k lives in U[k & 1]
    float* U1;
    float* W;           // Bertero weights, nx*ny*nz (unused when has_w == 0)
    const float* im;    // image M*N*P (offset already added)
    const float* psf;   // PSF pM*pN*pP
    const float* start_img;  // M*N*P start image for identity/LP starts
    double* part;       // per X-tile partial sums (ntile[KIND_X])
    double* stat;       // setup reduction scratch
    float* err;         // per-iteration error, niter floats
    double* errd;       // per-iteration error in double, niter
    // FFT plans for the complex lengths h (x rows), ny and nz: forward with
    // radices R1..Rk, inverse with Rk..R1; twiddles relative to tw
    FftPlanDev pxf, pxi, pyf, pyi, pzf, pzi;
    const float2* tw;
    const float2* rtw;  // exp(-2 pi i k / nx), k < h
};

struct BatchDev {
    int njobs;
    int ntile[KIND_N];
};

// One tile of deconwolf's XY tiling (tiling.c): core box pos and the box
// grown by the overlap, xpos (inclusive bounds x0 x1 y0 y1 z0 z1), and the
// tile's deconvolved output (xsize volume, offset still added).
struct TileDev {
    int pos[6];
    int xpos[6];
    int xsize[3];
    int pad;
    const float* out;
};
