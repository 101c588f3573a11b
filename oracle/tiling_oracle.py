"""Restatement of deconwolf's XY tiling (src/tiling.c) and of the tile
loop of deconvolve_tiles (src/dw.c:2000-2229), numpy, float32 where the
reference computes in float.  TEST INFRASTRUCTURE ONLY: the checker of
the device tiling path.

Tiles are dicts with pos[6] (core box, inclusive bounds: x0 x1 y0 y1 z0 z1),
xpos[6] (box grown by the overlap, clamped to the image) and xsize[3].
"""
import math

import numpy as np

f32 = np.float32


def get_division(M, m):
    """tiling_getDivision (tiling.c:22-58): [start, end] pairs."""
    ns = f32(math.ceil(M / m))
    width = f32(f32(M) / ns)
    n = int(ns)
    divs = [0] * (2 * n)
    divs[2 * n - 1] = M - 1
    for dd in range(1, n):
        divs[dd * 2 - 1] = int(math.ceil(float(f32(width * f32(dd))))) - 1
        divs[dd * 2] = divs[dd * 2 - 1] + 1
    return divs


def create(M, N, P, max_size, overlap):
    """tiling_create (tiling.c:70-130): M-divisions outer, N-divisions inner."""
    dm = get_division(M, max_size)
    dn = get_division(N, max_size)
    tiles = []
    for mm in range(len(dm) // 2):
        for nn in range(len(dn) // 2):
            pos = [dm[2 * mm], dm[2 * mm + 1], dn[2 * nn], dn[2 * nn + 1], 0, P - 1]
            xpos = [max(0, pos[0] - overlap), min(pos[1] + overlap, M - 1),
                    max(0, pos[2] - overlap), min(pos[3] + overlap, N - 1), pos[4], pos[5]]
            xsize = [xpos[1] - xpos[0] + 1, xpos[3] - xpos[2] + 1, xpos[5] - xpos[4] + 1]
            tiles.append(dict(pos=pos, xpos=xpos, xsize=xsize))
    return tiles


def weight1d(a, b, c, d, x):
    """getWeight1d (tiling.c:169-195), float arithmetic; x an int array."""
    a, b, c, d = f32(a), f32(b), f32(c), f32(d)
    x = np.asarray(x)
    xf = x.astype(f32)
    w = np.zeros(x.shape, f32)
    ramp_up = ((xf - a) + f32(1)) / ((b - a) + f32(1))
    ramp_dn = f32(1) - (xf - c) / ((d - c) + f32(1))
    inside = (xf >= b) & (xf <= c)
    up = (xf >= a) & (xf <= b) & ~inside
    dn = (xf >= c) & (xf <= d) & ~inside & ~up
    w[inside] = 1
    w[up] = ramp_up[up]
    w[dn] = ramp_dn[dn]
    return w


def tile_weight(t, m, n):
    """tile_getWeight (tiling.c:197-213): min of the x and y ramps; m, n
    broadcastable int arrays."""
    wm = weight1d(t["xpos"][0], t["pos"][0], t["pos"][1], t["xpos"][1], m)
    wn = weight1d(t["xpos"][2], t["pos"][2], t["pos"][3], t["xpos"][3], n)
    return np.minimum(wm, wn)


def weight_sum(tiles, m, n):
    """tiling_getWeights (tiling.c:215-223), summed in tile order in float."""
    s = np.zeros(np.broadcast(np.asarray(m), np.asarray(n)).shape, f32)
    for t in tiles:
        s = (s + tile_weight(t, m, n)).astype(f32)
    return s


def get_tile(t, V):
    """tiling_get_tile (tiling.c:300-330); V shaped (P, N, M)."""
    x = t["xpos"]
    return np.ascontiguousarray(V[x[4]:x[5] + 1, x[2]:x[3] + 1, x[0]:x[1] + 1])


def put_tiles(tiles, outs, shape):
    """tiling_put_tile_raw (tiling.c:353-410) for every tile in order:
    target += (w / sum w) * tile, in float, starting from zeros."""
    P, N, M = shape
    target = np.zeros(shape, f32)
    yy, xx = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    wsum = weight_sum(tiles, xx, yy)
    for t, S in zip(tiles, outs):
        x = t["xpos"]
        m = xx[x[2]:x[3] + 1, x[0]:x[1] + 1]
        n = yy[x[2]:x[3] + 1, x[0]:x[1] + 1]
        w = (tile_weight(t, m, n) / wsum[x[2]:x[3] + 1, x[0]:x[1] + 1]).astype(f32)
        sub = target[x[4]:x[5] + 1, x[2]:x[3] + 1, x[0]:x[1] + 1]
        target[x[4]:x[5] + 1, x[2]:x[3] + 1, x[0]:x[1] + 1] = (sub + (w[None] * S.astype(f32))).astype(f32)
    return target


def finish_tile(out, offset):
    """deconvolve_tiles after s->fun (dw.c:2141-2145): -offset, then
    negatives to 0."""
    o = (out.astype(f32) - f32(offset)).astype(f32)
    o[o < 0] = 0
    return o
