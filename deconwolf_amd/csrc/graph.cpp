// HIP graphs of the SHB rounds.  A round's launch sequence depends only on
// the job table (the iteration counter, the heavy-ball weight and the
// buffer parity are read on the device from the job's state), so one
// captured round replays as every round of a FIXED job, or of a batch
// whose members all have rounds left.  Capture runs on a stream the
// library owns (the caller's may be the legacy default stream, which
// cannot be captured); the graph is launched on the caller's stream.
#include <hip/hip_runtime.h>

#include <stdlib.h>

#include "dwhip.h"
#include "dwhip_internal.h"

int dwhip_graph_mode()
{
    static const int mode = [] {
        const char* v = getenv("DWHIP_GRAPH");
        if (!v || !*v) return 1;
        const int m = atoi(v);
        return (m < 0 || m > 2) ? 1 : m;
    }();
    return mode;
}

static hipStream_t capture_stream()
{
    // one per host thread, kept for the thread's lifetime
    static thread_local hipStream_t cs = nullptr;
    if (!cs && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) cs = nullptr;
    return cs;
}

int dwhip_capture(int (*body)(void* ctx, hipStream_t cs), void* ctx, hipGraphExec_t* exec)
{
    *exec = nullptr;
    hipStream_t cs = capture_stream();
    if (!cs) return DWHIP_ERR_HIP;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return dwhip_hip_fail(e, "hipStreamBeginCapture", __LINE__);
    const int rc = body(ctx, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (rc != DWHIP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return dwhip_hip_fail(e, "hipStreamEndCapture", __LINE__);
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        *exec = nullptr;
        return dwhip_hip_fail(e, "hipGraphInstantiate", __LINE__);
    }
    return DWHIP_OK;
}
