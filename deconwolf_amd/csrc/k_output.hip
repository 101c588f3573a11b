// Output path of dw_run on the device: the statistics deconwolf takes of
// the result before writing it (largest value for the 16-bit scaling,
// scaling_for_u16 src/dw.c:98-103 over fim_max; voxels at the background
// level, get_nbg src/dw.c:2406-2418) and the float -> uint16 conversions of
// its writers:
//   DWHIP_U16_ROUND  fim_tiff_write_opt (src/fim_tiff.c:1138-1150): v*s, values
//                    above 65535 -> 65535, non-finite -> 0, round half away
//   DWHIP_U16_TRUNC  fim_imwrite_u16 to .npy (src/fim.c:5142-5147) and
//                    fim_tiff_imwrite_u16_from_raw (src/fim_tiff.c:819-822):
//                    (uint16_t)(v*s), truncation
//   DWHIP_U16_RINT   raw_to_npio (src/dw.c:187-190): nearbyintf(v*s)
// The C conversions are undefined outside [0, 65536); here they saturate to
// [0, 65535] and map NaN to 0.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "dwhip.h"
#include "dwhip_internal.h"

#define OTHR 256

// Monotone map of floats onto unsigned ints (NaN excluded by the caller).
__device__ __forceinline__ unsigned int float_key(float f)
{
    const unsigned int b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__host__ __device__ __forceinline__ float key_float(unsigned int k)
{
    const unsigned int b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
#ifdef __HIP_DEVICE_COMPILE__
    return __uint_as_float(b);
#else
    float f;
    memcpy(&f, &b, sizeof(f));
    return f;
#endif
}

struct OutStats {
    unsigned int maxkey;
    unsigned int pad;
    unsigned long long nbg;
};

__global__ __launch_bounds__(OTHR) void k_out_stats(const float* __restrict__ v, int64_t n, float bg,
                                                    OutStats* __restrict__ st)
{
    __shared__ unsigned int smax[OTHR / 64];
    __shared__ unsigned long long scnt[OTHR / 64];
    unsigned int mk = 0;
    unsigned long long cnt = 0;
    for (int64_t i = (int64_t)blockIdx.x * OTHR + threadIdx.x; i < n; i += (int64_t)gridDim.x * OTHR) {
        const float x = v[i];
        if (x == x) mk = max(mk, float_key(x));
        cnt += (x == bg);
    }
    for (int o = 32; o > 0; o >>= 1) {
        mk = max(mk, __shfl_xor(mk, o));
        cnt += __shfl_xor(cnt, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smax[w] = mk;
        scnt[w] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < OTHR / 64; ++k) {
            mk = max(mk, smax[k]);
            cnt += scnt[k];
        }
        atomicMax(&st->maxkey, mk);
        atomicAdd(&st->nbg, cnt);
    }
}

template <int MODE>
__device__ __forceinline__ uint16_t to_u16(float x, float scaling)
{
    float v = x * scaling;
    if (MODE == DWHIP_U16_ROUND) {
        if (v > 65535.f) v = 65535.f;
        if (!isfinite(v)) v = 0.f;
        v = roundf(v);
    } else if (MODE == DWHIP_U16_TRUNC) {
        v = truncf(v);
    } else {
        v = rintf(v);
    }
    if (!(v > 0.f)) return 0;  // also NaN
    return v >= 65535.f ? (uint16_t)65535 : (uint16_t)v;
}

template <int MODE>
__global__ __launch_bounds__(OTHR) void k_to_u16(const float* __restrict__ v, int64_t n, float scaling,
                                                 uint16_t* __restrict__ out)
{
    for (int64_t i = (int64_t)blockIdx.x * OTHR + threadIdx.x; i < n; i += (int64_t)gridDim.x * OTHR)
        out[i] = to_u16<MODE>(v[i], scaling);
}

static int out_grid(int64_t n)
{
    const int64_t g = (n + OTHR - 1) / OTHR;
    return (int)(g < 4096 ? (g > 0 ? g : 1) : 4096);
}

extern "C" {

int dwhip_output_stats_device(const float* v, int64_t n, float bg, float* vmax, int64_t* n_bg, void* stream)
{
    if (!v || n < 1) return DWHIP_ERR_ARG;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    OutStats* d = nullptr;
    OutStats h;
    hipError_t e = hipMalloc((void**)&d, sizeof(OutStats));
    if (e != hipSuccess) return e == hipErrorOutOfMemory ? DWHIP_ERR_NOMEM : DWHIP_ERR_HIP;
    e = hipMemsetAsync(d, 0, sizeof(OutStats), s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_out_stats, dim3(out_grid(n)), dim3(OTHR), 0, s, v, n, bg, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(d);
    if (e != hipSuccess) return dwhip_hip_fail(e, "output statistics", __LINE__);
    // key 0 means every value was NaN
    if (vmax) *vmax = h.maxkey ? key_float(h.maxkey) : NAN;
    if (n_bg) *n_bg = (int64_t)h.nbg;
    return DWHIP_OK;
}

int dwhip_to_u16_device(const float* v, int64_t n, float scaling, int mode, uint16_t* out, void* stream)
{
    if (!v || !out || n < 0) return DWHIP_ERR_ARG;
    if (mode != DWHIP_U16_ROUND && mode != DWHIP_U16_TRUNC && mode != DWHIP_U16_RINT) return DWHIP_ERR_ARG;
    if (n == 0) return DWHIP_OK;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    const int g = out_grid(n);
    if (mode == DWHIP_U16_ROUND)
        hipLaunchKernelGGL(k_to_u16<DWHIP_U16_ROUND>, dim3(g), dim3(OTHR), 0, s, v, n, scaling, out);
    else if (mode == DWHIP_U16_TRUNC)
        hipLaunchKernelGGL(k_to_u16<DWHIP_U16_TRUNC>, dim3(g), dim3(OTHR), 0, s, v, n, scaling, out);
    else
        hipLaunchKernelGGL(k_to_u16<DWHIP_U16_RINT>, dim3(g), dim3(OTHR), 0, s, v, n, scaling, out);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "u16 conversion", __LINE__);
}

}  // extern "C"
