// dwhip_shb_run_device_noalloc: the per-image SHB method on device pointers
// with no device allocation, for callers that cannot allocate (code under
// HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
//
// A job's state (spectra, OTF, iterates, weights, partial sums, twiddles,
// its one-job table) lives in `scratch`, a static device array of
// DWHIP_SCRATCH_BYTES that each GPU the library runs on holds until the
// process ends.  Its device address is looked up once per device with
// hipGetSymbolAddress (the array's name in host code is a host-side stand-in,
// not a device address) and that address is what the job is bound to.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind; the set-up kernels write everything the loop reads before it is
// read (bg, the flat constant, the argmax and the iteration counter come
// from k_stats_final, the iterates from k_start, the partial sums from
// each round), so nothing needs clearing between calls.  Calls that overlap
// on one GPU (two host threads, or two streams) race on the array.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dwhip.h"
#include "dwhip_internal.h"

static __device__ __align__(256) unsigned char scratch[DWHIP_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// Device address of the current device's scratch array (looked up on the
// first call with that device current).
int scratch_base(void** base)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return DWHIP_ERR_NODEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return dwhip_hip_fail(e, "hipGetSymbolAddress(scratch)", __LINE__);
        g_scratch[dev] = p;
    }
    *base = g_scratch[dev];
    return DWHIP_OK;
}

}  // namespace

extern "C" int dwhip_shb_run_device_noalloc(const float* im, int64_t M, int64_t N, int64_t P, const float* psf,
                                            int64_t pM, int64_t pN, int64_t pP, const dwhip_params* prm,
                                            float* out, dwhip_stats* st, void* stream)
{
    if (!prm) return DWHIP_ERR_ARG;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    void* base = nullptr;
    rc = scratch_base(&base);
    if (rc != DWHIP_OK) return rc;
    dwhip_params p = *prm;
    p.method = DWHIP_METHOD_SHB;
    return dwhip_run_device_in(im, M, N, P, psf, pM, pN, pP, &p, out, st, stream, base, DWHIP_SCRATCH_BYTES);
}
