// One column-pass mode (CMODE, set by the build) of the SHB loop.
#include "shb_passes.h"

#ifndef CMODE
#error "compile with -DCMODE=<mode>"
#endif
#define DWHIP_CAT2(a, b) a##b
#define DWHIP_CAT(a, b) DWHIP_CAT2(a, b)

extern "C" hipError_t DWHIP_CAT(dwhip_launch_c, CMODE)(const BatchDev* bd, const JobDev* jobs, int grid, size_t lds,
                                                      hipStream_t s)
{
    static bool once = false;
    if (!once) {
        once = true;
        dwhip_allow_big_lds(reinterpret_cast<const void*>(k_colpass<CMODE>));
    }
    hipLaunchKernelGGL(k_colpass<CMODE>, dim3(grid), dim3(col_threads<CMODE>()), lds, s, bd, jobs);
    return hipGetLastError();
}
