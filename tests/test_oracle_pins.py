"""Oracle pins for the parts the spatial restatement (test_oracle.py) leaves
open: the MSE metric, the ABS and REL stopping rules at their edges, the
copysign branch of the ratio, and PSF maxima ties.  Each expected value
comes from a closed form or from another run of the oracle under a
different rule (tests/edgecases.py), never from the oracle's own formula.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests import edgecases as ec


# the trace is float in both builds (getError returns float, dw.c:1436-1459)
@pytest.mark.parametrize("dtype,rtol", [(np.float32, 2e-6), (np.float64, 1.2e-7)])
def test_iteration0_metrics_closed_form(dtype, rtol):
    """y_0 == bg (delta PSF, flat start below bg): MSE_0 = mean((bg - g)^2)
    and IDIV_0 = mean over g > 0 of g log(g / bg) - (g - bg)."""
    im, k = ec.metric_case()
    mse, idiv, bg = ec.metric_closed_forms(im)
    c = ec.flat_constant(im, po.work_size(2, im.shape, k.shape))
    assert c < bg
    _, e, bgo, _ = po.shb(im, k, niter=1, metric=po.METRIC_MSE, dtype=dtype)
    assert bgo == np.float32(bg)
    assert abs(e[0] - mse) <= rtol * mse + 1e-30, (e[0], mse)
    _, e, _, _ = po.shb(im, k, niter=1, metric=po.METRIC_IDIV, dtype=dtype)
    assert abs(e[0] - idiv) <= max(rtol, 2e-6) * idiv, (e[0], idiv)


def test_mse_zero_at_the_fixed_point():
    """Identity start with a delta PSF: y == g from iteration 0, MSE 0."""
    im, k = ec.metric_case(seed=3)
    _, e, _, n = po.shb(im, k, niter=4, metric=po.METRIC_MSE, start=po.START_IDENTITY, dtype=np.float64)
    assert n == 4 and np.all(np.abs(e) < 1e-18)


def _decreasing_case():
    rng = np.random.default_rng(21)
    im = (rng.random((8, 14, 16)) * 100 + 10).astype(np.float32)
    k = rng.random((5, 5, 5)).astype(np.float32)
    k[2, 2, 2] += 3
    k = (k / k.sum(dtype=np.float64)).astype(np.float32)
    return im, k


@pytest.mark.parametrize("metric", [po.METRIC_IDIV, po.METRIC_MSE])
def test_abs_rule_from_fixed_trace(metric):
    """dw_iterator_next's ABS rule (dw.c:256-262): before iteration k > 0,
    stop when e_{k-1} < err_abs.  With err_abs between e_4 and e_5 of a
    FIXED run, the ABS run stops after 6 iterations with the FIXED run's
    6-iteration output."""
    im, k = _decreasing_case()
    o12, e, _, _ = po.shb(im, k, niter=12, metric=metric)
    assert np.all(np.diff(e[:7]) < 0), e
    thr = float(np.sqrt(float(e[4]) * float(e[5])))
    assert e[5] < thr < e[4]
    oabs, eabs, _, n = po.shb(im, k, iter_type=po.ITER_ABS, err_abs=thr, maxiter=40, metric=metric)
    assert n == 6
    o6, _, _, _ = po.shb(im, k, niter=6, metric=metric)
    assert np.array_equal(oabs, o6) and np.array_equal(eabs, e[:6])
    # iteration 0 always runs: the rule is tested for k > 0 only
    _, _, _, n = po.shb(im, k, iter_type=po.ITER_ABS, err_abs=1e30, maxiter=40, metric=metric)
    assert n == 1
    # maxiter caps it
    _, _, _, n = po.shb(im, k, iter_type=po.ITER_ABS, err_abs=0.0, maxiter=7, metric=metric)
    assert n == 7


@pytest.mark.parametrize("method", ["shb", "rl"])
def test_rel_rule_stops_before_iteration0(method):
    """err_rel in (1/3, 1): |e - 2 e err_rel| / e < err_rel already at
    iteration 0 (e = 1), so nothing runs and the result is the start
    guess (the flat constant) cropped to the image."""
    im, k = _decreasing_case()
    run = po.shb if method == "shb" else po.rl
    out, e, _, n = run(im, k, iter_type=po.ITER_REL, err_rel=0.5, maxiter=40)
    assert n == 0 and e.size == 0
    c = ec.flat_constant(im, po.work_size(2, im.shape, k.shape))
    assert np.all(out == c)
    # just outside the interval the first iteration runs
    _, _, _, n = run(im, k, iter_type=po.ITER_REL, err_rel=0.3, maxiter=1)
    assert n == 1


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ratio_epsilon_branch_closed_form(dtype):
    """bg 0 and a flat start c < 1e-6: y_0 = c on the whole image, every
    image voxel takes fabs(y) < 1e-6 -> y = copysign(1e-6, y)
    (method_shb.c:41-42), and x_1 = (g / 1e-6) c; the unclamped ratio g / c
    would give x_1 = g instead."""
    im, k = ec.eps_case()
    c = ec.flat_constant(im, po.work_size(2, im.shape, k.shape))
    assert 0 < c < 1e-6
    out, _, bg, n = po.shb(im, k, niter=2, bg_auto=0, bg=0.0, dtype=dtype)
    assert bg == 0.0 and n == 2
    ref = ec.eps_closed_form(im, c)
    assert np.abs(out - ref).max() <= 1e-5 * ref.max()
    assert np.abs(out - im).max() > 0.5 * im.max()


def test_argmax_tie_lowest_index():
    """Two equal PSF maxima: the OTF is the PSF shifted by the lower
    linear index (fim.c:307-319), by brute force."""
    k, (ilo, ihi), _ = ec.tie_psf()
    im = (np.random.default_rng(4).random((6, 9, 11)) * 50 + 20).astype(np.float32)
    cK, _, _, _, _ = po.shb_setup(im, k, dtype=np.float64)
    wM, wN, wP = po.work_size(2, im.shape, k.shape)
    Z = np.zeros((wP, wN, wM))
    Z[:k.shape[0], :k.shape[1], :k.shape[2]] = k
    for idx, expect in ((ilo, True), (ihi, False)):
        mid = np.unravel_index(idx, k.shape)
        K = np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2))
        close = np.allclose(cK, np.fft.rfftn(K), atol=1e-12)
        assert close == expect
