// Per-pass kernel timing for the benchmark's roofline figures: when enabled,
// every pass launch is bracketed by two HIP events on its stream, and the
// algorithmic bytes of the launch are recorded beside it.  Events are
// resolved at the next collection point (after a stream synchronisation).
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "shb_kernels.h"

namespace {

struct Pending {
    hipEvent_t a, b;
    int kernel;
    double bytes;
};

struct Totals {
    long long launches = 0;
    double ms = 0;
    double bytes = 0;
};

std::mutex g_mu;
bool g_on = false;
std::vector<hipEvent_t> g_pool;
std::vector<Pending> g_pending;
Totals g_tot[DWHIP_PROF_KERNELS];

hipEvent_t take_event()
{
    if (!g_pool.empty()) {
        hipEvent_t e = g_pool.back();
        g_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

}  // namespace

const char* dwhip_prof_kernel_name(int k)
{
    static const char* names[DWHIP_PROF_KERNELS] = {
        "x_psf", "x_ones", "x_weights", "x_update", "x_ratio", "x_load", "x_store",
        "y_fwd", "y_inv", "z_otf", "z_conv", "z_convj", "z_fwd", "z_inv", "y_inv_img", "y_fwd_img",
    };
    return (k >= 0 && k < DWHIP_PROF_KERNELS) ? names[k] : "?";
}

bool dwhip_prof_on() { return g_on; }

int dwhip_prof_begin(hipStream_t s)
{
    if (!g_on) return -1;
    std::lock_guard<std::mutex> lk(g_mu);
    Pending p;
    p.a = take_event();
    p.b = take_event();
    if (!p.a || !p.b) return -1;
    p.kernel = -1;
    p.bytes = 0;
    if (hipEventRecord(p.a, s) != hipSuccess) return -1;
    g_pending.push_back(p);
    return (int)g_pending.size() - 1;
}

void dwhip_prof_end(hipStream_t s, int slot, int kernel, double bytes)
{
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    if (slot >= (int)g_pending.size()) return;
    Pending& p = g_pending[slot];
    p.kernel = kernel;
    p.bytes = bytes;
    (void)hipEventRecord(p.b, s);
}

// Resolves the recorded launches; call after their streams have finished.
void dwhip_prof_collect()
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (Pending& p : g_pending) {
        float ms = 0;
        if (p.kernel >= 0 && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            Totals& t = g_tot[p.kernel];
            t.launches += 1;
            t.ms += ms;
            t.bytes += p.bytes;
        }
        g_pool.push_back(p.a);
        g_pool.push_back(p.b);
    }
    g_pending.clear();
}

extern "C" int dwhip_profile_enable(int on)
{
    dwhip_prof_collect();
    std::lock_guard<std::mutex> lk(g_mu);
    g_on = on != 0;
    for (int k = 0; k < DWHIP_PROF_KERNELS; ++k) g_tot[k] = Totals();
    return DWHIP_OK;
}

extern "C" int dwhip_profile_get(int k, const char** name, long long* launches, double* total_ms, double* bytes)
{
    if (k < 0 || k >= DWHIP_PROF_KERNELS) return DWHIP_ERR_ARG;
    dwhip_prof_collect();
    std::lock_guard<std::mutex> lk(g_mu);
    if (name) *name = dwhip_prof_kernel_name(k);
    if (launches) *launches = g_tot[k].launches;
    if (total_ms) *total_ms = g_tot[k].ms;
    if (bytes) *bytes = g_tot[k].bytes;
    return DWHIP_OK;
}
