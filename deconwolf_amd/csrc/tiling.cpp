// XY tiling of large images, the device form of deconwolf's tiling.c and of
// the tile loop in deconvolve_tiles (src/dw.c:2000-2229): the plan
// (tiling_create), extraction of the grown tile boxes and the weighted
// blend of the deconvolved tiles.  The tiles themselves run as jobs of the
// batched engine (or one by one through dwhip_shb_run_device).
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "dwhip.h"
#include "dwhip_internal.h"
#include "shb_kernels.h"

namespace {

// tiling_getDivision (tiling.c:22-58): [start, end] of each section
std::vector<int64_t> division(int64_t M, int64_t m)
{
    const float ns = (float)ceil((double)M / (double)m);
    const float width = (float)M / ns;
    const int n = (int)ns;
    std::vector<int64_t> d(2 * (size_t)n, 0);
    d[0] = 0;
    d[2 * n - 1] = M - 1;
    for (int64_t dd = 1; dd < n; ++dd) {
        d[dd * 2 - 1] = (int64_t)(int)ceil(width * (float)dd) - 1;
        d[dd * 2] = d[dd * 2 - 1] + 1;
    }
    return d;
}

}  // namespace

extern "C" int dwhip_tiling_create(int64_t M, int64_t N, int64_t P, int64_t max_size, int64_t overlap,
                                   dwhip_tile* tiles, int max_tiles, int* ntiles)
{
    if (M < 1 || N < 1 || P < 1 || max_size < 1 || overlap < 0 || !ntiles) return DWHIP_ERR_ARG;
    const std::vector<int64_t> dm = division(M, max_size), dn = division(N, max_size);
    const int nM = (int)dm.size() / 2, nN = (int)dn.size() / 2;
    *ntiles = nM * nN;
    if (!tiles) return DWHIP_OK;
    if (max_tiles < nM * nN) return DWHIP_ERR_ARG;
    int b = 0;
    for (int mm = 0; mm < nM; ++mm)
        for (int nn = 0; nn < nN; ++nn) {
            dwhip_tile& t = tiles[b++];
            t.pos[0] = dm[2 * mm];
            t.pos[1] = dm[2 * mm + 1];
            t.pos[2] = dn[2 * nn];
            t.pos[3] = dn[2 * nn + 1];
            t.pos[4] = 0;
            t.pos[5] = P - 1;
            t.xpos[0] = std::max<int64_t>(0, t.pos[0] - overlap);
            t.xpos[1] = std::min<int64_t>(t.pos[1] + overlap, M - 1);
            t.xpos[2] = std::max<int64_t>(0, t.pos[2] - overlap);
            t.xpos[3] = std::min<int64_t>(t.pos[3] + overlap, N - 1);
            t.xpos[4] = t.pos[4];
            t.xpos[5] = t.pos[5];
            t.xsize[0] = t.xpos[1] - t.xpos[0] + 1;
            t.xsize[1] = t.xpos[3] - t.xpos[2] + 1;
            t.xsize[2] = t.xpos[5] - t.xpos[4] + 1;
        }
    return DWHIP_OK;
}

extern "C" int dwhip_tile_extract_device(const float* im, int64_t M, int64_t N, int64_t P, const dwhip_tile* t,
                                         float offset, float* out, void* stream)
{
    if (!im || !t || !out) return DWHIP_ERR_ARG;
    if (t->xpos[0] < 0 || t->xpos[1] >= M || t->xpos[2] < 0 || t->xpos[3] >= N || t->xpos[4] < 0 ||
        t->xpos[5] >= P || t->xsize[0] != t->xpos[1] - t->xpos[0] + 1 || t->xsize[1] != t->xpos[3] - t->xpos[2] + 1 ||
        t->xsize[2] != t->xpos[5] - t->xpos[4] + 1)
        return DWHIP_ERR_ARG;
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipError_t e = dwhip_launch_tile_extract(im, (int)M, (int)N, (int)t->xpos[0], (int)t->xpos[2], (int)t->xpos[4],
                                             (int)t->xsize[0], (int)t->xsize[1], (int)t->xsize[2], offset, out,
                                             (hipStream_t)stream);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "tile extract", __LINE__);
}

extern "C" int dwhip_tiles_blend_device(const dwhip_tile* tiles, const float* const* outs, int ntiles, int64_t M,
                                        int64_t N, int64_t P, float offset, float* out, void* stream)
{
    if (!tiles || !outs || ntiles < 1 || !out || M < 1 || N < 1 || P < 1) return DWHIP_ERR_ARG;
    for (int i = 0; i < ntiles; ++i) {
        const dwhip_tile& t = tiles[i];
        if (!outs[i] || t.xpos[0] < 0 || t.xpos[1] >= M || t.xpos[2] < 0 || t.xpos[3] >= N || t.xpos[4] < 0 ||
            t.xpos[5] >= P || t.xsize[0] != t.xpos[1] - t.xpos[0] + 1 || t.xsize[1] != t.xpos[3] - t.xpos[2] + 1)
            return DWHIP_ERR_ARG;
    }
    int rc = dwhip_init();
    if (rc != DWHIP_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    std::vector<TileDev> h(ntiles);
    for (int i = 0; i < ntiles; ++i) {
        memset(&h[i], 0, sizeof(TileDev));
        for (int k = 0; k < 6; ++k) {
            h[i].pos[k] = (int)tiles[i].pos[k];
            h[i].xpos[k] = (int)tiles[i].xpos[k];
        }
        for (int k = 0; k < 3; ++k) h[i].xsize[k] = (int)tiles[i].xsize[k];
        h[i].out = outs[i];
    }
    // the tile table from the stream-ordered pool (no device-wide
    // allocation after the first call); the synchronisation keeps h alive
    // until its upload has run
    TileDev* d = nullptr;
    hipError_t e = hipMallocAsync((void**)&d, sizeof(TileDev) * ntiles, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d, h.data(), sizeof(TileDev) * ntiles, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = dwhip_launch_tile_blend(d, ntiles, (int)M, (int)N, (int)P, offset, out, s);
    if (d) {
        const hipError_t f = hipFreeAsync(d, s);
        if (e == hipSuccess) e = f;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? DWHIP_OK : dwhip_hip_fail(e, "tile blend", __LINE__);
}
