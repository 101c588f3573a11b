"""Parity at the benchmark's full sizes.

* c1 (128x128x32, Born-Wolf PSF 127x127x63, transform 256x256x96): all 20
  iterations of the configuration against the float32 oracle with its own
  transforms, relative L2 <= 1e-5, error trace to 1e-4.
* c2 (1024x1024x60, PSF 181x181x79, transform 1250x1215x140): 3 iterations
  against the float32 oracle whose transforms go through MKL's FFTW3
  interface (the own-FFT oracle is too slow at 200 M voxels), same bars;
  the MKL-backed oracle is first checked against the own-FFT oracle on a
  small case.
* c5 (one 1024x1024x81 channel, PSF 181x181x81 at dz 300 nm, transform
  1250x1215x162: a z plan of radices 6, 3, 3, 3): 2 iterations against the
  same MKL-backed float32 oracle, same bars.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import workloads
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _gpu(im, psf, niter):
    s = dw.DwOpts(nIter=niter, iter_type=dw.ITER_FIXED)
    out = dw.deconvolve_shb_hip_device(torch.from_numpy(im).to(DEV), torch.from_numpy(psf).to(DEV), s)
    torch.cuda.synchronize()
    return out.cpu().numpy(), s


def test_c1_full_size():
    print("c1 full size", flush=True)
    im, psf, s0, _ = workloads.make("c1")
    out, s = _gpu(im, psf, s0.nIter)
    assert s.stats["work"] == (256, 256, 96)
    o32, e32, bg, n = po.shb(im, psf, niter=s0.nIter, work=s.stats["work"])
    assert n == s0.nIter and s.bg == bg
    assert rel_l2(out, o32) <= 1e-5
    np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)


def test_c2_full_size():
    print("c2 full size", flush=True)
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    if not po.set_fft_backend(True, 16):
        pytest.skip("MKL's FFTW3 interface is not available on this host")
    try:
        # the external transforms agree with the oracle's own on a small case
        rng = np.random.default_rng(3)
        im_s = (rng.random((10, 24, 30)) * 100 + 10).astype(np.float32)
        k_s = np.ones((3, 5, 5), np.float32) / 75
        ext, _, _, _ = po.shb(im_s, k_s, niter=4)
        po.set_fft_backend(False)
        own, _, _, _ = po.shb(im_s, k_s, niter=4)
        assert rel_l2(ext, own) <= 1e-5
        po.set_fft_backend(True, 16)

        im, psf, _, _ = workloads.make("c2")
        out, s = _gpu(im, psf, 3)
        assert s.stats["work"] == (1250, 1215, 140)
        o32, e32, bg, n = po.shb(im, psf, niter=3, work=s.stats["work"], nthreads=16)
        assert n == 3 and s.bg == bg
        assert rel_l2(out, o32) <= 1e-5
        np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)
    finally:
        po.set_fft_backend(False)


def test_c5_channel_full_size():
    print("c5 full size, channel 690 nm", flush=True)
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    if not po.set_fft_backend(True, 16):
        pytest.skip("MKL's FFTW3 interface is not available on this host")
    try:
        job = workloads.make_c5(0, channels=(3,), device=DEV, iters=2)[0]
        im = job["im"].cpu().numpy()
        psf = job["psf"].cpu().numpy()
        out, s = _gpu(im, psf, 2)
        assert s.stats["work"][2] == 162 and s.stats["work"][2] >= im.shape[0] + psf.shape[0] - 1
        o32, e32, bg, n = po.shb(im, psf, niter=2, work=s.stats["work"], nthreads=16)
        assert n == 2 and s.bg == bg
        assert rel_l2(out, o32) <= 1e-5
        np.testing.assert_allclose(s.err_trace, e32, rtol=1e-4)
    finally:
        po.set_fft_backend(False)
