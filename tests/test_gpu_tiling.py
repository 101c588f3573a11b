"""GPU tiling (dwhip_tile_extract_device, dwhip_tiles_blend_device, and the
tiled SHB of deconwolf_amd.tiling) against the oracle restatement of
tiling.c and deconvolve_tiles: extraction and blend bit-exact, the tiled
deconvolution within rel-L2 1e-5 of the oracle's per-tile float32 SHB
blended by the oracle; tiles batched through the engine equal tiles run
one by one."""
import copy

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import synth, tiling
from deconwolf_amd.serve import Engine
from oracle import pyoracle as po
from oracle import tiling_oracle as to

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def test_extract_and_blend_exact():
    rng = np.random.default_rng(0)
    P, N, M = 5, 47, 61
    V = (rng.random((P, N, M)) * 100).astype(np.float32)
    tiles = tiling.plan(M, N, P, 20, 5)
    ref_tiles = to.create(M, N, P, 20, 5)
    dV = torch.from_numpy(V).to(DEV)
    outs, refs = [], []
    for t, rt in zip(tiles, ref_tiles):
        x = tiling.extract(dV, t, 5.0)
        assert np.array_equal(x.cpu().numpy(), to.get_tile(rt, V) + np.float32(5.0))
        # a fake method output around the offset, some values below it
        y = (x.cpu().numpy() * np.float32(0.9) - np.float32(3)).astype(np.float32)
        outs.append(torch.from_numpy(y).to(DEV))
        refs.append(to.finish_tile(y, 5.0))
    got = tiling.blend(tiles, outs, (P, N, M), 5.0).cpu().numpy()
    want = to.put_tiles(ref_tiles, refs, (P, N, M))
    assert np.array_equal(got, want)


def _image(seed=1, shape=(10, 70, 84)):
    rng = np.random.default_rng(seed)
    k = synth.gaussian_psf(11, 11, 7, 1.4, 1.7, 2.0)
    V = synth.point_sources(shape, 60, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), 60.0, rng=rng)
    return im, k


@pytest.mark.parametrize("use_engine,method", [(False, dw.METHOD_SHB), (True, dw.METHOD_SHB),
                                               (False, dw.METHOD_RL), (True, dw.METHOD_RL)])
def test_tiled_shb_matches_oracle(use_engine, method):
    print("tiled, engine =", use_engine, "method", method, flush=True)
    im, k = _image()
    P, N, M = im.shape
    s = dw.DwOpts(nIter=8, iter_type=dw.ITER_FIXED, method=method)
    psf = tiling.tile_psf(k, im.shape, s)     # dw_run's preparation for the whole image
    eng = Engine(hbm_budget_gb=1.0, max_active=16) if use_engine else None
    out, tiles, opts = tiling.deconvolve_tiles(torch.from_numpy(im).to(DEV), psf, s, 40, 8, engine=eng)
    out = out.cpu().numpy()
    # oracle: per tile, float32 SHB at the GPU's transform size, then the blend
    ref_tiles = to.create(M, N, P, 40, 8)
    assert len(ref_tiles) == 6   # ceil(84/40) x ceil(70/40)
    refs = []
    for t, rt in zip(tiles, ref_tiles):
        x = to.get_tile(rt, im) + np.float32(s.offset)
        tp = tiling.tile_psf(psf, x.shape, s)
        w, _, _ = dw.work_size(x.shape, tp.shape)
        run = po.rl if method == dw.METHOD_RL else po.shb
        o, _, _, _ = run(x, tp, niter=8, work=w)
        refs.append(to.finish_tile(o, s.offset))
    want = to.put_tiles(ref_tiles, refs, im.shape)
    rel = np.linalg.norm(out.astype(np.float64) - want) / np.linalg.norm(want)
    assert rel <= 1e-5, rel
    if eng is not None:
        eng.close()


def test_engine_tiles_equal_sequential_tiles():
    im, k = _image(seed=2)
    s = dw.DwOpts(nIter=5, iter_type=dw.ITER_FIXED)
    psf = tiling.tile_psf(k, im.shape, s)
    dim = torch.from_numpy(im).to(DEV)
    a, _, _ = tiling.deconvolve_tiles(dim, psf, copy.copy(s), 40, 8, engine=None)
    eng = Engine(hbm_budget_gb=1.0, max_active=16)
    b, _, _ = tiling.deconvolve_tiles(dim, psf, copy.copy(s), 40, 8, engine=eng)
    assert torch.equal(a, b)
    eng.close()
