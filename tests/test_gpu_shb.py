"""GPU parity of the HIP SHB path against the CPU oracle (oracle/).

Every test calls libdwhip through its C-ABI (deconwolf_amd wraps it with
ctypes).  Tolerances are written in each test:

* transforms: max error relative to the largest magnitude <= 2e-6;
* SHB loop at the same transform size as the oracle's float32 run:
  relative L2 <= 1e-5 after up to 20 iterations, and <= 1e-4 against the
  float64 oracle run (the bar set for deconwolf's float SHB);
* error (I-divergence) trace: relative 1e-4 per iteration.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import synth
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def case(shape, psfshape, seed=1, npts=30, bg=50.0, psf="gauss"):
    P, N, M = shape
    rng = np.random.default_rng(seed)
    pP, pN, pM = psfshape
    if psf == "gauss":
        k = synth.gaussian_psf(pM, pN, pP, 1.3, 1.6, 1.9)
    else:
        k = psf
    V = synth.point_sources(shape, npts, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), bg, rng=rng)
    im += np.float32(5.0)  # dw_run adds the offset before the method (dw.c:2835-2838)
    return im, k


# (nz, ny, nx): cover every radix of the planner (2..9, 15, 16, 20, 25),
# one to four stages, n = 1, and the split step for odd and even h = nx/2
FFT_SHAPES = [(10, 12, 16), (15, 21, 70), (32, 64, 128), (14, 243, 250), (9, 5, 2), (16, 256, 1250),
              (140, 36, 40), (162, 30, 48), (96, 1215, 36), (4, 2401, 6), (1, 1, 8)]


@pytest.mark.parametrize("shape", FFT_SHAPES)
def test_fft3_forward_inverse(shape):
    print("fft3", shape, flush=True)
    rng = np.random.default_rng(7)
    vol = rng.standard_normal(shape).astype(np.float32)
    t = torch.from_numpy(vol).to(DEV)
    spec = dw.fft3_device(t).cpu().numpy()
    ref = np.fft.rfftn(vol.astype(np.float64))
    err = np.abs(spec - ref).max() / np.abs(ref).max()
    assert err < 2e-6, err
    back = torch.empty_like(t)
    dw.fft3_device(back, inverse=True, spec=torch.from_numpy(spec.astype(np.complex64)).to(DEV))
    back = back.cpu().numpy() / vol.size
    err2 = np.abs(back - vol).max() / np.abs(vol).max()
    assert err2 < 2e-6, err2


def test_fft3_matches_oracle_fft():
    rng = np.random.default_rng(3)
    vol = rng.standard_normal((7, 9, 10)).astype(np.float32)
    spec = dw.fft3_device(torch.from_numpy(vol).to(DEV)).cpu().numpy()
    ref = po.rfft3(vol)
    assert np.abs(spec - ref).max() / np.abs(ref).max() < 2e-6


def _opts(n, **kw):
    s = dw.DwOpts(nIter=n, iter_type=dw.ITER_FIXED)
    for k, v in kw.items():
        setattr(s, k, v)
    return s


@pytest.mark.parametrize("niter", [1, 2, 3, 10, 20])
def test_shb_matches_oracle(niter):
    print("shb small niter", niter, flush=True)
    im, k = case((16, 36, 40), (7, 9, 9))
    s = _opts(niter)
    out = dw.deconvolve_shb_hip(im, k, s)
    work = s.stats["work"]
    o32, e32, bg32, n32 = po.shb(im, k, niter=niter, work=work)
    o64, e64, _, _ = po.shb(im, k, niter=niter, work=work, dtype=np.float64)
    assert s.stats["iterations"] == niter == n32
    assert abs(s.bg - bg32) == 0
    assert rel_l2(out, o32) <= 1e-5
    assert rel_l2(out, o64) <= 1e-4
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_shb_reference_work_size_vs_padded():
    """The GPU pads the work volume to 2,3,5,7-smooth sizes; the oracle at
    deconwolf's exact work size gives the same image region."""
    im, k = case((13, 31, 35), (5, 7, 7), seed=5)   # reference work (41, 37, 17)
    s = _opts(8)
    out = dw.deconvolve_shb_hip(im, k, s)
    assert s.stats["work"] != s.stats["ref_work"]
    oref, _, _, _ = po.shb(im, k, niter=8, dtype=np.float64)
    assert rel_l2(out, oref) <= 1e-4


def test_shb_iter0_and_iter1():
    im, k = case((8, 20, 24), (5, 5, 5), seed=2)
    out0 = dw.deconvolve_shb_hip(im, k, _opts(0))
    assert np.array_equal(out0, im)  # method_shb.c:101-105
    s = _opts(1)
    out1 = dw.deconvolve_shb_hip(im, k, s)
    wM, wN, wP = s.stats["ref_work"]
    c = np.float32(np.float32(im.astype(np.float64).sum()) / np.float32(wM * wN * wP))
    assert np.all(out1 == c)


def test_shb_delta_psf_recovers_image():
    im, _ = case((8, 20, 24), (5, 5, 5), seed=3)
    d = np.zeros((5, 5, 5), np.float32)
    d[2, 2, 2] = 1
    out = dw.deconvolve_shb_hip(im, d, _opts(2))
    np.testing.assert_allclose(out, im, rtol=2e-6)


@pytest.mark.parametrize("bq", [0, 1])
def test_shb_border_quality_exact_sizes(bq):
    # sizes chosen so that the reference work size is 2,3,5,7-smooth with even x
    if bq == 1:
        shape, ks = (11, 26, 43), (7, 7, 9)   # w = (43+5, 26+4, 11+4) = (48, 30, 15)
    else:
        shape, ks = (12, 27, 40), (5, 5, 5)   # w = max(image, psf) = (40, 27, 12)
    im, k = case(shape, ks, seed=4)
    w, r, rc = dw.work_size(im.shape, k.shape, bq)
    assert rc == 0 and w == r, (w, r)
    s = _opts(6, borderQuality=bq)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, e32, _, _ = po.shb(im, k, niter=6, border_quality=bq)
    assert rel_l2(out, o32) <= 1e-5


def test_shb_mse_metric_and_no_positivity():
    im, k = case((10, 24, 30), (5, 7, 7), seed=6)
    s = _opts(6, metric=dw.METRIC_MSE, positivity=0, alphamax=0.5)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, e32, _, _ = po.shb(im, k, niter=6, metric=po.METRIC_MSE, positivity=0, alphamax=0.5,
                            work=s.stats["work"])
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_shb_identity_start_and_fixed_bg():
    im, k = case((10, 24, 30), (5, 7, 7), seed=8)
    s = _opts(5, start_condition=dw.START_IDENTITY, bg_auto=0, bg=20.0)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, _, bg, _ = po.shb(im, k, niter=5, start=po.START_IDENTITY, bg_auto=0, bg=20.0,
                           work=s.stats["work"])
    assert s.bg == 20.0 and bg == np.float32(20.0)
    assert rel_l2(out, o32) <= 1e-5


def test_shb_rel_stopping_matches_oracle():
    im, k = case((10, 24, 30), (5, 7, 7), seed=9)
    s = dw.DwOpts(iter_type=dw.ITER_REL, err_rel=0.05, maxiter=40)
    out = dw.deconvolve_shb_hip(im, k, s)
    o32, e32, _, n = po.shb(im, k, iter_type=po.ITER_REL, err_rel=0.05, maxiter=40,
                            work=s.stats["work"])
    assert s.stats["iterations"] == n
    assert rel_l2(out, o32) <= 1e-5


def test_device_entry_matches_host_entry():
    im, k = case((12, 30, 34), (5, 7, 7), seed=10)
    s1 = _opts(7)
    a = dw.deconvolve_shb_hip(im, k, s1)
    s2 = _opts(7)
    b = dw.deconvolve_shb_hip_device(torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV), s2)
    torch.cuda.synchronize()
    assert np.array_equal(a, b.cpu().numpy())


def test_shb_lp_start_matches_oracle():
    """DW_START_LP: the image smoothed with sigma 8 (fim_gsmooth) inserted as
    the start (method_shb.c:278-287)."""
    from deconwolf_amd import filters
    im, k = case((10, 24, 30), (5, 7, 7), seed=11)
    s = _opts(6, start_condition=dw.START_LP)
    out = dw.deconvolve_shb_hip(im, k, s)
    lp = filters.gsmooth(im, 8.0)
    o32, _, _, _ = po.shb(im, k, niter=6, start=po.START_LP, start_image=lp, work=s.stats["work"])
    assert rel_l2(out, o32) <= 1e-5


@pytest.mark.parametrize("shape,psfshape", [((1, 40, 50), (1, 9, 9)),     # 2-D image: z transforms of length 1
                                            ((6, 20, 24), (5, 31, 31)),   # PSF wider than the image
                                            ((9, 1, 64), (5, 1, 7)),      # one row per plane
                                            ((3, 17, 1), (3, 5, 1))])     # one column: x transform of length 2
def test_shb_degenerate_shapes(shape, psfshape):
    print("degenerate", shape, psfshape, flush=True)
    rng = np.random.default_rng(31)
    P, N, M = shape
    im = (rng.random(shape) * 200 + 20).astype(np.float32)
    k = rng.random(psfshape).astype(np.float32) + 0.05
    k[tuple(d // 2 for d in psfshape)] += 2
    k = (k / k.sum(dtype=np.float64)).astype(np.float32)
    for method in (dw.METHOD_SHB, dw.METHOD_RL):
        s = _opts(6, method=method)
        run = dw.deconvolve_rl_hip if method == dw.METHOD_RL else dw.deconvolve_shb_hip
        out = run(im, k, s)
        ref = po.rl if method == dw.METHOD_RL else po.shb
        o32, e32, _, _ = ref(im, k, niter=6, work=s.stats["work"])
        assert rel_l2(out, o32) <= 1e-5, (method, rel_l2(out, o32))
        np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)
