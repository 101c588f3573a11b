"""The per-iteration hook behind deconwolf's --iterdump, --tsv and --fulldump
(method_shb.c:205-209,302-322,380,407-411), and measured planning.

* The hook sees every iteration's error (equal to the error trace of the
  same job run without a hook, whose FIXED loop replays a captured graph:
  the direct launches and the graph give the same bytes), and every k-th
  iteration the guess x_k at the top of that iteration, cropped: for SHB
  that is what a run of k + 1 iterations returns (x_{n-1}), for RL what a
  run of k iterations returns (x_n).
* With fulldump the hook gets the returned guess at the transform size
  (its crop is the result) and the PSF inserted at that size and shifted so
  its maximum (lowest index) is at the origin (numpy.roll of the insert).
* dwhip_plan_tune measures a size once and caches the plan; a job of that
  size matches the oracle with the measured plan.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _case(seed=5, shape=(10, 26, 30), psf_shape=(7, 9, 9)):
    rng = np.random.default_rng(seed)
    pz, py, px = psf_shape
    k = synth.gaussian_psf(px, py, pz, 1.2, 1.5, 1.9)
    k[1, 2, 3] = k.max()           # a tie with the centre: the lowest index wins
    k = (k / k.sum(dtype=np.float64)).astype(np.float32)
    V = synth.point_sources(shape, 30, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), 45.0, rng=rng) + np.float32(5.0)
    return im, k


@pytest.mark.parametrize("method", [dw.METHOD_SHB, dw.METHOD_RL])
def test_hook_errors_iterates_and_full_dumps(method):
    print("hook: method", method, flush=True)
    im, k = _case()
    run = dw.deconvolve_rl_hip if method == dw.METHOD_RL else dw.deconvolve_shb_hip
    n = 6
    plain = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED, method=method)
    ref = run(im, k, plain)
    seen = {}

    def hook(it, err, x):
        seen[it] = (err, None if x is None else x.copy())

    s = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED, method=method, iterdump=2, fulldump=1, iter_hook=hook)
    out = run(im, k, s)
    assert np.array_equal(out, ref)
    assert sorted(seen) == [-2, -1] + list(range(n))
    np.testing.assert_array_equal([seen[i][0] for i in range(n)], plain.err_trace)
    for it in range(n):
        x = seen[it][1]
        if it % 2:
            assert x is None
            continue
        if method == dw.METHOD_SHB:
            want = run(im, k, dw.DwOpts(nIter=it + 1, iter_type=dw.ITER_FIXED))
        elif it == 0:
            want = None
        else:
            want = run(im, k, dw.DwOpts(nIter=it, iter_type=dw.ITER_FIXED, method=method))
        assert x.shape == im.shape
        if want is not None:
            assert np.array_equal(x, want), it
    full = seen[-1][1]
    nx, ny, nz = s.stats["work"]
    P, N, M = im.shape
    assert full.shape == (nz, ny, nx) and np.array_equal(full[:P, :N, :M], out)
    Z = np.zeros((nz, ny, nx), np.float32)
    Z[:k.shape[0], :k.shape[1], :k.shape[2]] = k
    mid = np.unravel_index(int(np.argmax(k)), k.shape)   # first (lowest) index of the maximum
    assert np.array_equal(seen[-2][1], np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2)))


def test_plan_tune_cached_and_parity(tmp_path):
    print("planning: tune and cache", flush=True)
    dw.plan_cache_path(str(tmp_path / "plans.txt"))
    try:
        im, k = _case(seed=9, shape=(14, 40, 48), psf_shape=(7, 9, 9))
        w, _, _ = dw.work_size(im.shape, k.shape)
        assert not dw.plan_get(w)[0]
        tuned, rule = dw.plan_tune(w)
        assert 0 < tuned <= rule
        measured, rad, tiles = dw.plan_get(w)
        assert measured
        text = (tmp_path / "plans.txt").read_text()
        assert text.startswith("gfx950 %d %d %d " % w)
        s = dw.DwOpts(nIter=8, iter_type=dw.ITER_FIXED)
        out = dw.deconvolve_shb_hip(im, k, s)
        o32, e32, _, _ = po.shb(im, k, niter=8, work=s.stats["work"])
        assert np.linalg.norm(out.astype(np.float64) - o32) / np.linalg.norm(o32) <= 1e-5
        # a fresh load of the file gives the same plan
        dw.plan_cache_path(str(tmp_path / "plans.txt"))
        assert dw.plan_get(w) == (True, rad, tiles)
    finally:
        dw.plan_cache_path("")
