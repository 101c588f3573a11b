"""GPU parity of the HIP Richardson-Lucy path (deconvolve_rl,
src/method_rl.c:105-400) against the oracle's RL restatement.

Tolerances, as for the SHB path: relative L2 <= 1e-5 against the float32
oracle at the same transform size and <= 1e-4 against the float64 one;
error trace relative 1e-4 per iteration.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from oracle import pyoracle as po
from test_gpu_shb import DEV, case, rel_l2

pytestmark = pytest.mark.gpu


def _opts(n, **kw):
    s = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED, method=dw.METHOD_RL)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


@pytest.mark.parametrize("niter", [1, 2, 5, 15])
def test_rl_matches_oracle(niter):
    print("rl niter", niter, flush=True)
    im, k = case((16, 36, 40), (7, 9, 9), seed=21)
    s = _opts(niter)
    out = dw.deconvolve_rl_hip(im, k, s)
    work = s.stats["work"]
    o32, e32, bg32, n32 = po.rl(im, k, niter=niter, work=work)
    o64, _, _, _ = po.rl(im, k, niter=niter, work=work, dtype=np.float64)
    assert s.stats["iterations"] == niter == n32
    assert s.bg == bg32
    assert rel_l2(out, o32) <= 1e-5
    assert rel_l2(out, o64) <= 1e-4
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_rl_reference_work_size_vs_padded():
    im, k = case((13, 31, 35), (5, 7, 7), seed=22)
    s = _opts(6)
    out = dw.deconvolve_rl_hip(im, k, s)
    assert s.stats["work"] != s.stats["ref_work"]
    oref, _, _, _ = po.rl(im, k, niter=6, dtype=np.float64)
    assert rel_l2(out, oref) <= 1e-4


def test_rl_iter0_and_delta_psf():
    im, k = case((8, 20, 24), (5, 5, 5), seed=23)
    out0 = dw.deconvolve_rl_hip(im, k, _opts(0))
    assert np.array_equal(out0, im)  # method_rl.c:120-123
    d = np.zeros((5, 5, 5), np.float32)
    d[2, 2, 2] = 1
    s = _opts(3)
    out = dw.deconvolve_rl_hip(im, d, s)
    np.testing.assert_allclose(out, np.maximum(im, s.bg), rtol=2e-6)


@pytest.mark.parametrize("bq", [0, 1])
def test_rl_border_quality_exact_sizes(bq):
    if bq == 1:
        shape, ks = (11, 26, 43), (7, 7, 9)
    else:
        shape, ks = (12, 27, 40), (5, 5, 5)
    im, k = case(shape, ks, seed=24)
    s = _opts(5, borderQuality=bq)
    out = dw.deconvolve_rl_hip(im, k, s)
    o32, e32, _, _ = po.rl(im, k, niter=5, border_quality=bq)
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_rl_flux_conservation_bq0_no_clamp():
    """bq 0, image filling the work volume, bg = 0: sum(x_n) = sum(g)."""
    im, k = case((12, 27, 40), (5, 5, 5), seed=25)
    s = _opts(6, borderQuality=0, bg_auto=0, bg=0.0)
    out = dw.deconvolve_rl_hip(im, k, s)
    assert abs(out.astype(np.float64).sum() / im.astype(np.float64).sum() - 1) < 1e-5


def test_rl_rel_stopping_and_mse():
    im, k = case((10, 24, 30), (5, 7, 7), seed=26)
    s = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=0.02, maxiter=40, method=dw.METHOD_RL)
    out = dw.deconvolve_rl_hip(im, k, s)
    o32, e32, _, n = po.rl(im, k, iter_type=po.ITER_REL, err_rel=0.02, maxiter=40, work=s.stats["work"])
    assert s.stats["iterations"] == n
    assert rel_l2(out, o32) <= 1e-5
    s = _opts(5, metric=dw.METRIC_MSE, start_condition=dw.START_IDENTITY)
    out = dw.deconvolve_rl_hip(im, k, s)
    o32, e32, _, _ = po.rl(im, k, niter=5, metric=po.METRIC_MSE, start=po.START_IDENTITY,
                           work=s.stats["work"])
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


@pytest.mark.parametrize("groups", [1, 2])
def test_rl_device_entry_and_engine_bit_identical(groups):
    """Host entry, device entry and the batched engine (RL and SHB jobs in
    one batch, in one group or in two groups on two streams) give the same
    bytes."""
    from deconwolf_amd.serve import Engine
    im, k = case((12, 30, 34), (5, 7, 7), seed=27)
    im2, k2 = case((9, 22, 26), (5, 5, 7), seed=28)
    a = dw.deconvolve_rl_hip(im, k, _opts(7))
    b = dw.deconvolve_rl_hip_device(torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), _opts(7))
    torch.cuda.synchronize()
    assert np.array_equal(a, b.cpu().numpy())
    shb = dw.deconvolve_shb_hip(im2, k2, dw.DwOpts(nIter=5, iter_type=dw.ITER_FIXED))
    eng = Engine(hbm_budget_gb=1.0, max_active=8, groups=groups)
    eng.submit(1, torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), _opts(7))
    eng.submit(2, torch.from_numpy(im2).to(DEV), torch.from_numpy(k2).to(DEV),
               dw.DwOpts(nIter=5, iter_type=dw.ITER_FIXED))
    done = []
    while len(done) < 2:
        done += eng.step()
    o1, st1 = eng.result(1)
    o2, st2 = eng.result(2)
    eng.close()
    assert st1["iterations"] == 7 and st2["iterations"] == 5
    assert np.array_equal(o1.cpu().numpy(), a)
    assert np.array_equal(o2.cpu().numpy(), shb)
