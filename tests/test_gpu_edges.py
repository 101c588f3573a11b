"""GPU tests of the loop's edges against closed forms and the oracle
(tests/edgecases.py): iteration-0 metrics, the ABS and REL stopping rules
(REL stopping before iteration 0 included, on the plain path and in the
engine), the copysign branch of the ratio, and PSF maxima ties.

Tolerances: float32 loop against the float32 oracle, relative L2 <= 1e-5;
closed forms of iteration 0 to 2e-6 relative; the branch's closed form to
1e-5 of the largest value (transform rounding).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd.serve import Engine
from oracle import pyoracle as po
from tests import edgecases as ec

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _fixed(n, **kw):
    s = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def test_iteration0_metrics_closed_form():
    print("edges: metrics", flush=True)
    im, k = ec.metric_case()
    mse, idiv, bg = ec.metric_closed_forms(im)
    # y_0 = IFFT(K FFT(bg)) comes out of float32 transforms as bg within a
    # few ulp (the oracle's rounds only once per axis pass and hits bg); allow
    # |y_0 - bg| <= 4 ulp(bg) = d and propagate it: |dMSE| <= 2 mean|g-bg| d + d^2,
    # |dIDIV| <= mean|1 - g/bg| d, on top of the 2e-6 of the float32 metric.
    d = 4 * float(np.spacing(np.float32(bg)))
    gm = im.astype(np.float64)
    s = _fixed(1, metric=dw.METRIC_MSE)
    dw.deconvolve_shb_hip(im, k, s)
    assert s.bg == np.float32(bg)
    tol = 2e-6 * mse + 2 * np.mean(np.abs(gm - bg)) * d + d * d
    assert abs(s.err_trace[0] - mse) <= tol, (s.err_trace[0], mse, tol)
    s = _fixed(1)
    dw.deconvolve_shb_hip(im, k, s)
    tol = 2e-6 * idiv + np.mean(np.abs(1 - gm / bg)) * d
    assert abs(s.err_trace[0] - idiv) <= tol, (s.err_trace[0], idiv, tol)


def _decreasing_case():
    rng = np.random.default_rng(21)
    im = (rng.random((8, 14, 16)) * 100 + 10).astype(np.float32)
    k = rng.random((5, 5, 5)).astype(np.float32)
    k[2, 2, 2] += 3
    k = (k / k.sum(dtype=np.float64)).astype(np.float32)
    return im, k


@pytest.mark.parametrize("metric", [dw.METRIC_IDIV, dw.METRIC_MSE])
def test_abs_stopping(metric):
    """ABS with err_abs between e_4 and e_5 of the GPU's own FIXED trace
    stops after 6 iterations, as the oracle's ABS run does."""
    print("edges: abs", metric, flush=True)
    im, k = _decreasing_case()
    s = _fixed(12, metric=metric)
    dw.deconvolve_shb_hip(im, k, s)
    e = s.err_trace
    thr = float(np.sqrt(float(e[4]) * float(e[5])))
    assert e[5] < thr < e[4]
    sa = dw.DwOpts(iter_type=dw.ITER_ABS, err_abs=thr, maxiter=40, metric=metric)
    out = dw.deconvolve_shb_hip(im, k, sa)
    assert sa.stats["iterations"] == 6
    o32, e32, _, n = po.shb(im, k, iter_type=po.ITER_ABS, err_abs=thr, maxiter=40, metric=metric,
                            work=sa.stats["work"])
    assert n == 6
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(sa.err_trace, e32, rtol=1e-4)


@pytest.mark.parametrize("method", [dw.METHOD_SHB, dw.METHOD_RL])
def test_rel_stop_before_iteration0(method):
    """err_rel 0.5: no iteration runs; the result is the flat start for
    deconwolf's work size, bit for bit, on the plain path and through the
    engine."""
    print("edges: rel stop before 0", method, flush=True)
    im, k = _decreasing_case()
    c = ec.flat_constant(im, po.work_size(2, im.shape, k.shape))
    s = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=0.5, maxiter=40, method=method)
    run = dw.deconvolve_rl_hip if method == dw.METHOD_RL else dw.deconvolve_shb_hip
    out = run(im, k, s)
    assert s.stats["iterations"] == 0 and s.err_trace.size == 0
    assert np.all(out == c)
    assert s.bg == np.float32(max(im.min(), np.float32(1e-2)))
    eng = Engine(hbm_budget_gb=0.5, max_active=4)
    try:
        s2 = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=0.5, maxiter=40, method=method)
        s3 = dw.DwOpts(nIter=3, iter_type=dw.ITER_FIXED, method=method)
        ti, tk = torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV)
        eng.submit(1, ti, tk, s2)
        eng.submit(2, ti, tk, s3)
        done = []
        for _ in range(10):
            done += eng.step()
            if len(done) == 2:
                break
        assert sorted(done) == [1, 2]
        o1, st1 = eng.result(1)
        assert st1["iterations"] == 0 and np.all(o1.cpu().numpy() == c)
        o2, st2 = eng.result(2)
        s4 = dw.DwOpts(nIter=3, iter_type=dw.ITER_FIXED, method=method)
        assert np.array_equal(o2.cpu().numpy(), run(im, k, s4))
    finally:
        eng.close()


def test_ratio_epsilon_branch():
    """bg 0 and a flat start below 1e-6: every image voxel takes the
    copysign branch at iteration 0; x_1 = (g / 1e-6) c (closed form), and
    the 6-iteration trajectory, which keeps y below 1e-6 in the image's
    faint voxels, matches the oracle."""
    print("edges: epsilon branch", flush=True)
    im, k = ec.eps_case()
    c = ec.flat_constant(im, po.work_size(2, im.shape, k.shape))
    assert 0 < c < 1e-6
    s = _fixed(2, bg_auto=0, bg=0.0)
    out = dw.deconvolve_shb_hip(im, k, s)
    ref = ec.eps_closed_form(im, c)
    assert np.abs(out - ref).max() <= 1e-5 * ref.max()
    s6 = _fixed(6, bg_auto=0, bg=0.0)
    out6 = dw.deconvolve_shb_hip(im, k, s6)
    o32, _, _, _ = po.shb(im, k, niter=6, bg_auto=0, bg=0.0, work=s6.stats["work"])
    assert rel_l2(out6, o32) <= 1e-5


def test_argmax_tie_uses_lowest_index():
    """A PSF with two equal maxima: the GPU matches the oracle (lowest
    linear index) and not the same PSF with the tie broken the other way."""
    print("edges: argmax tie", flush=True)
    k, _, (a, b) = ec.tie_psf()
    rng = np.random.default_rng(9)
    im = (rng.random((10, 24, 30)) * 80 + 20).astype(np.float32)
    s = _fixed(5)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, _, _, _ = po.shb(im, k, niter=5, work=s.stats["work"])
    assert rel_l2(out, o32) <= 1e-5
    # the other maximum, made the unique one by one ulp
    hi = a if np.ravel_multi_index(a, k.shape) > np.ravel_multi_index(b, k.shape) else b
    k2 = k.copy()
    k2[hi] = np.nextafter(k2[hi], np.float32(1))
    o_other, _, _, _ = po.shb(im, k2, niter=5, work=s.stats["work"])
    assert rel_l2(out, o_other) > 1e-3
