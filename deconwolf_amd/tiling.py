"""Tiled deconvolution of large images on the GPU: the device form of
deconwolf's deconvolve_tiles (src/dw.c:2000-2229) and tiling.c.  The image
is split into overlapping XY tiles (tiling_create), each tile is
deconvolved with its own autocropped PSF and offset handling, and the tile
outputs are blended with deconwolf's linear weights.  Tiles run together
through the batched engine (one launch per pass for all of them) instead
of one after the other through raw temporary files.
"""
import ctypes

import numpy as np

from ._lib import DwhipTile, check, lib
from .method_shb import DwOpts, fim_normalize_sum1, psf_autocrop


def plan(M, N, P, max_size, overlap):
    """deconwolf's tiling (tiling_create, tiling.c:70-130) as a list of
    dwhip_tile structs."""
    n = ctypes.c_int()
    check(lib().dwhip_tiling_create(M, N, P, max_size, overlap, None, 0, ctypes.byref(n)), "dwhip_tiling_create")
    arr = (DwhipTile * n.value)()
    check(lib().dwhip_tiling_create(M, N, P, max_size, overlap, arr, n.value, ctypes.byref(n)), "dwhip_tiling_create")
    return list(arr)


def tile_dict(t):
    return dict(pos=list(t.pos), xpos=list(t.xpos), xsize=list(t.xsize))


def extract(im, t, offset):
    """Tile t of the device image im (torch (P,N,M)), offset added."""
    import torch
    out = torch.empty((t.xsize[2], t.xsize[1], t.xsize[0]), dtype=torch.float32, device=im.device)
    P, N, M = im.shape
    s = torch.cuda.current_stream().cuda_stream
    check(lib().dwhip_tile_extract_device(im.data_ptr(), M, N, P, ctypes.byref(t), float(offset), out.data_ptr(), s),
          "dwhip_tile_extract_device")
    return out


def blend(tiles, outs, shape, offset, out=None):
    """Weighted blend of the tiles' method outputs into an image (P,N,M)."""
    import torch
    P, N, M = shape
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=outs[0].device)
    arr = (DwhipTile * len(tiles))(*tiles)
    ptrs = (ctypes.c_void_p * len(outs))(*[o.data_ptr() for o in outs])
    s = torch.cuda.current_stream().cuda_stream
    check(lib().dwhip_tiles_blend_device(arr, ptrs, len(tiles), M, N, P, float(offset), out.data_ptr(), s),
          "dwhip_tiles_blend_device")
    return out


def tile_psf(psf, tshape, s: DwOpts):
    """The per-tile PSF of deconvolve_tiles (dw.c:2127-2129): normalise,
    autocrop for the tile, normalise."""
    p = fim_normalize_sum1(psf)
    p = psf_autocrop(p, tshape, s)
    return fim_normalize_sum1(p)


def deconvolve_tiles(im, psf, s: DwOpts, max_size, overlap, engine=None):
    """im: (P,N,M) float32 GPU tensor without offset; psf: numpy (pP,pN,pM),
    normalised and autocropped for the whole image (dw_run,
    src/dw.c:2799-2804).  Returns the blended (P,N,M) result, offset removed
    and clipped at 0 per tile as deconvolve_tiles does.  engine: a
    deconwolf_amd.serve.Engine to run the tiles together; None runs them
    one after the other."""
    import copy
    import torch
    P, N, M = im.shape
    tiles = plan(M, N, P, max_size, overlap)
    if len(tiles) < 2:
        raise ValueError("only one tile (src/dw.c:2016-2024): omit tiling or use a smaller tile size")
    ins, psfs, opts = [], [], []
    for t in tiles:
        ins.append(extract(im, t, s.offset))
        tp = tile_psf(psf, (t.xsize[2], t.xsize[1], t.xsize[0]), s)
        psfs.append(torch.as_tensor(tp, device=im.device).contiguous())
        opts.append(copy.copy(s))
    outs = []
    if engine is None:
        from ._lib import METHOD_RL
        from .method_shb import deconvolve_rl_hip_device, deconvolve_shb_hip_device
        run = deconvolve_rl_hip_device if int(s.method) == METHOD_RL else deconvolve_shb_hip_device
        for x, p, o in zip(ins, psfs, opts):
            outs.append(run(x, p, o))
    else:
        base = 1 << 40
        for i, (x, p, o) in enumerate(zip(ins, psfs, opts)):
            engine.submit(base + i, x, p, o)
        done = set()
        while len(done) < len(tiles):
            done.update(engine.step())
        for i, o in enumerate(opts):
            out, st = engine.result(base + i)
            o.err_trace = st["err_trace"]
            o.stats = st
            outs.append(out)
    return blend(tiles, outs, (P, N, M), s.offset), tiles, opts
