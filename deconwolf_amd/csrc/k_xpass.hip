// One X-pass mode (XMODE, set by the build) of the SHB loop.
#ifndef XMODE
#error "compile with -DXMODE=<mode>"
#endif

// Per-mode build tunables of the two iteration passes: occupancy target
// (waves per SIMD) and element chunk of the update (XM_UPDATE = 3) and
// ratio (XM_RATIO = 4) passes, over the x-pass-wide DWHIP_X_WAVES / XCHUNK.
// The ratio pass is held to three waves per SIMD (168 VGPRs; its LDS tiles
// allow three workgroups per CU): without the bound the compiler takes 170.
#if XMODE == 4 && !defined(DWHIP_XR_WAVES) && !defined(DWHIP_X_WAVES)
#define DWHIP_XR_WAVES 3
#endif
#if XMODE == 3 && defined(DWHIP_XU_WAVES)
#define DWHIP_X_WAVES DWHIP_XU_WAVES
#endif
#if XMODE == 4 && defined(DWHIP_XR_WAVES)
#define DWHIP_X_WAVES DWHIP_XR_WAVES
#endif
#if XMODE == 3 && defined(DWHIP_XU_CHUNK)
#define XCHUNK DWHIP_XU_CHUNK
#endif
#if XMODE == 4 && defined(DWHIP_XR_CHUNK)
#define XCHUNK DWHIP_XR_CHUNK
#endif

#include "shb_passes.h"
#define DWHIP_CAT2(a, b) a##b
#define DWHIP_CAT(a, b) DWHIP_CAT2(a, b)

extern "C" hipError_t DWHIP_CAT(dwhip_launch_x, XMODE)(const BatchDev* bd, const JobDev* jobs, int grid, size_t lds,
                                                      hipStream_t s)
{
    static bool once = false;
    if (!once) {
        once = true;
        dwhip_allow_big_lds(reinterpret_cast<const void*>(k_xpass<XMODE>));
    }
    hipLaunchKernelGGL(k_xpass<XMODE>, dim3(grid), dim3(XTHR), lds, s, bd, jobs);
    return hipGetLastError();
}
