"""GPU tests of the any-size path: border quality 0 and 1 at deconwolf's
exact work size (method_shb.c:137-153) when that size has an odd x extent
or a prime factor above 7, which the fused passes cannot transform.  Each
axis is a complex FFT in LDS (Stockham, or Bluestein's chirp-z form for
other factors; x as a real transform of two rows per pencil, y and z over
its half spectrum).  Compared with the float32 oracle at the same size (its own
transforms, Bluestein above 61): relative L2 <= 1e-5 and error traces to
1e-4, SHB and RL, FIXED and REL; in the engine, bit-identical to solo
runs beside batched jobs.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import synth
from deconwolf_amd.serve import Engine
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(seed, shape, psf_shape):
    rng = np.random.default_rng(seed)
    pz, py, px = psf_shape
    k = synth.gaussian_psf(px, py, pz, 1.3, 1.6, 1.9)
    V = synth.point_sources(shape, 30, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), 50.0, rng=rng) + np.float32(5.0)
    return im, k


# (image (P,N,M), psf (pP,pN,pM), bq): work sizes with odd x, primes > 7
CASES = [
    ((12, 26, 37), (7, 9, 9), 0),     # 37 x 26 x 12: odd prime x
    ((12, 26, 37), (7, 9, 9), 1),     # 42 x 31 x 16: 31 prime
    ((9, 40, 52), (7, 11, 11), 1),    # 58 x 46 x 13: 29, 23, 13
    ((20, 61, 64), (9, 9, 9), 0),     # 64 x 61 x 20: 61 prime
    # an odd number of x-rows (ny * nz): the last row pair of the real x
    # transform holds one row
    ((9, 27, 35), (7, 9, 9), 0),      # 35 x 27 x 9: odd x, no Bluestein, 243 rows
    ((11, 23, 37), (7, 9, 9), 0),     # 37 x 23 x 11: primes, 253 rows
]


@pytest.mark.parametrize("shape,psf_shape,bq", CASES)
@pytest.mark.parametrize("method", [dw.METHOD_SHB, dw.METHOD_RL])
def test_exact_size_matches_oracle(shape, psf_shape, bq, method):
    print("any-size:", shape, psf_shape, "bq", bq, "method", method, flush=True)
    im, k = _case(7, shape, psf_shape)
    w, r, rc = dw.work_size(im.shape, k.shape, bq)
    assert rc == 0 and w == r
    s = dw.DwOpts(nIter=8, iter_type=dw.ITER_FIXED, borderQuality=bq, method=method)
    run = dw.deconvolve_rl_hip if method == dw.METHOD_RL else dw.deconvolve_shb_hip
    out = run(im, k, s)
    assert tuple(s.stats["work"]) == tuple(r)
    orun = po.rl if method == dw.METHOD_RL else po.shb
    o32, e32, bg, n = orun(im, k, niter=8, border_quality=bq)
    assert n == 8 and s.bg == np.float32(bg)
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_exact_size_rel_stopping_and_mse():
    print("any-size: REL stopping, MSE", flush=True)
    im, k = _case(8, (10, 22, 33), (5, 7, 7))
    s = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=0.02, maxiter=40, borderQuality=1, metric=dw.METRIC_MSE)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, e32, _, n = po.shb(im, k, iter_type=po.ITER_REL, err_rel=0.02, maxiter=40, border_quality=1,
                            metric=po.METRIC_MSE)
    assert s.stats["iterations"] == n > 1
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_exact_size_in_engine_equals_solo():
    """An any-size job and batched fused-pass jobs in one engine: each
    result bit-identical to the job run alone."""
    print("any-size: engine", flush=True)
    specs = [((12, 26, 37), (7, 9, 9), 1, 6), ((10, 24, 30), (5, 7, 7), 2, 5), ((8, 20, 24), (5, 5, 5), 0, 4)]
    jobs = []
    for i, (sh, ps, bq, n) in enumerate(specs):
        im, k = _case(20 + i, sh, ps)
        s = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED, borderQuality=bq)
        jobs.append((torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), s))
    eng = Engine(hbm_budget_gb=1.0, max_active=8)
    try:
        for t, (im, k, s) in enumerate(jobs):
            eng.submit(t, im, k, s)
        done = []
        while len(done) < len(jobs):
            done += eng.step()
        for t, (im, k, s) in enumerate(jobs):
            import copy
            s2 = copy.copy(s)
            ref = dw.deconvolve_shb_hip_device(im, k, s2)
            out, st = eng.result(t)
            assert st["iterations"] == specs[t][3]
            assert torch.equal(out, ref), t
            np.testing.assert_array_equal(st["err_trace"], s2.err_trace)
    finally:
        eng.close()


def test_rel_any_size_and_fused_jobs_in_engine():
    """REL jobs of both kinds in one engine: the any-size job's stopping
    rule runs on the host (one round per step while it is active), the
    fused-pass jobs' on the device; every job stops after its solo run's
    iterations with a bit-identical result, and a fused job left alone
    finishes in multi-round steps."""
    print("any-size: REL jobs of both kinds in one engine", flush=True)
    specs = [((12, 26, 37), (7, 9, 9), 1, 0.02, 12), ((10, 24, 30), (5, 7, 7), 2, 0.01, 40),
             ((8, 20, 24), (5, 5, 5), 2, 1e-9, 30)]
    jobs = []
    for i, (sh, ps, bq, rel, mx) in enumerate(specs):
        im, k = _case(40 + i, sh, ps)
        s = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=rel, maxiter=mx, borderQuality=bq)
        jobs.append((torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), s))
    eng = Engine(hbm_budget_gb=1.0, max_active=8, rounds_per_step=8)
    try:
        for t, (im, k, s) in enumerate(jobs):
            eng.submit(t, im, k, s)
        done, steps = [], 0
        while len(done) < len(jobs):
            done += eng.step()
            steps += 1
            assert steps < 200
        rounds = 0
        for t, (im, k, s) in enumerate(jobs):
            import copy
            s2 = copy.copy(s)
            ref = dw.deconvolve_shb_hip_device(im, k, s2)
            out, st = eng.result(t)
            assert st["iterations"] == s2.stats["iterations"], (t, st["iterations"], s2.stats["iterations"])
            assert torch.equal(out, ref), t
            rounds = max(rounds, st["iterations"])
        # the last job runs to its maxiter of 30, past the any-size job's end
        assert rounds == 30 and steps < rounds
    finally:
        eng.close()
