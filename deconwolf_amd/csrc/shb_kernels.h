// Host-side launchers of the SHB pass kernels (shb_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include "shb_dev.h"

// X-pass modes: PSF (shifted PSF rows, R2C), ONES (image box, R2C),
// WEIGHTS (C2R, Bertero weights), UPDATE (pass A), RATIO (pass E),
// LOAD (R2C of U0), STORE (C2R into U0, unnormalised).
enum { XM_PSF = 0, XM_ONES = 1, XM_WEIGHTS = 2, XM_UPDATE = 3, XM_RATIO = 4, XM_LOAD = 5, XM_STORE = 6 };
// Column-pass modes; ZOTF: forward, divide by nx*ny*nz, store into K.
// Pruned modes of the SHB iteration (the skipped values are exact zeros or
// never read): ZCONV writes only z < P; ZCONVJ reads only z < P; YINV_IMG
// (pass D) transforms only slabs z < P and writes only y < N; YFWD_IMG
// (pass F) reads only y < N of slabs z < P and skips the other slabs.
enum { CM_YFWD = 0, CM_YINV = 1, CM_ZOTF = 2, CM_ZCONV = 3, CM_ZCONVJ = 4, CM_ZFWD = 5, CM_ZINV = 6,
       CM_YINV_IMG = 7, CM_YFWD_IMG = 8 };
#define DWHIP_N_CMODES 9
__host__ __device__ constexpr bool cmode_is_y(int m)
{
    return m == CM_YFWD || m == CM_YINV || m == CM_YINV_IMG || m == CM_YFWD_IMG;
}
// tile space of an x mode
__host__ __device__ constexpr int xmode_kind(int m) { return m == XM_RATIO ? KIND_XR : KIND_X; }
// tile space of a column mode
__host__ __device__ constexpr int cmode_kind(int m)
{
    return m == CM_YINV_IMG ? KIND_YD : m == CM_YFWD_IMG ? KIND_YF : cmode_is_y(m) ? KIND_Y : KIND_Z;
}

#ifndef DWHIP_XTHR
#define DWHIP_XTHR 256   // threads per X-pass workgroup
#endif
#ifndef DWHIP_CTHR
#define DWHIP_CTHR 1024  // threads per workgroup, Y passes (long pencils)
#endif
#ifndef DWHIP_ZTHR
#define DWHIP_ZTHR 512   // threads per workgroup, Z passes
#endif
// dynamic LDS a workgroup may use: gfx950's 160 KiB less 1 KiB kept for
// static LDS (the limit the kernels are raised to)
#define DWHIP_MAX_LDS_BYTES (159 * 1024)
#ifndef DWHIP_XMAXE
#define DWHIP_XMAXE 25   // complex elements per thread per stage, X pass
#endif
#ifndef DWHIP_CMAXE
#define DWHIP_CMAXE 20   // complex elements per thread per stage, Y/Z passes
#endif
#ifndef DWHIP_YMAXE
#define DWHIP_YMAXE DWHIP_CMAXE  // Y passes
#endif
#ifndef DWHIP_ZMAXE
#define DWHIP_ZMAXE 10   // Z passes: short pencils, many waves per SIMD
#endif
#ifndef DWHIP_XROWS_MAX
#define DWHIP_XROWS_MAX 64  // rows per X tile at most
#endif
#define DWHIP_MAX_RADIX 25
#ifndef DWHIP_X_STAGE_LDS
#define DWHIP_X_STAGE_LDS 1  // X passes stage their half-spectrum rows in LDS before the C2R
#endif

extern "C" {
hipError_t dwhip_launch_xpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s);
hipError_t dwhip_launch_colpass(int mode, const BatchDev* bd, const JobDev* jobs, int grid, size_t lds, hipStream_t s);
hipError_t dwhip_launch_stats(const JobDev* jobs, int j, int nblk, hipStream_t s);
hipError_t dwhip_launch_start(const JobDev* jobs, int j, int grid, hipStream_t s);
hipError_t dwhip_launch_round_end(const BatchDev* bd, const JobDev* jobs, int njobs, int advance, hipStream_t s);
hipError_t dwhip_launch_crop(const JobDev* jobs, int j, int k, float* out, int grid, hipStream_t s);
hipError_t dwhip_launch_tile_extract(const float* im, int M, int N, int x0, int y0, int z0, int xm, int xn, int xp,
                                     float offset, float* out, hipStream_t s);
hipError_t dwhip_launch_tile_blend(const TileDev* tiles, int ntiles, int M, int N, int P, float offset, float* out,
                                   hipStream_t s);
}
