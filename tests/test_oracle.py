"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

deconwolf ships no known-answer vectors for the SHB loop (its own tests
cover the FFT round trip, src/fft.c:708-780, and tiling only), and the
reference cannot be built here (src/dw.h:49 and src/fft.c:22 include
<fftw3.h>, absent from this image).  The oracle is pinned by:
  * the DFT definition (brute force) and numpy's pocketfft;
  * closed forms of the loop: --iter 0 returns the input, --iter 1 the flat
    start, a delta PSF recovers the image, an identity start with a delta
    PSF is a fixed point;
  * a spatial-domain restatement on tiny volumes: circular convolution and
    correlation by direct sums (not FFT), which fixes the OTF orientation,
    the conjugate in the adjoint and the 1/wMNP scaling;
  * invariances: padding the work volume beyond deconwolf's size leaves
    the image region unchanged (border quality 2), float32 vs float64.
"""
import numpy as np
import pytest

from oracle import pyoracle as po


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 7, 8, 9, 12, 16, 30, 49, 61, 67, 127, 254, 1250])
def test_cfft_vs_definition(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if n <= 64:
        k = np.arange(n)
        ref = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    else:
        ref = np.fft.fft(x)
    assert np.abs(po.cfft(x) - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(po.cfft(po.cfft(x), +1) / n - x).max() <= 1e-12 * max(1.0, np.abs(x).max())


@pytest.mark.parametrize("shape", [(3, 4, 5), (4, 6, 9), (5, 8, 10), (2, 3, 127)])
def test_rfft3_vs_numpy(shape):
    rng = np.random.default_rng(1)
    v = rng.standard_normal(shape)
    S = po.rfft3(v)
    assert np.abs(S - np.fft.rfftn(v)).max() <= 1e-10
    assert np.abs(po.irfft3(S, shape[2]) / v.size - v).max() <= 1e-12


def _case(shape=(6, 10, 12), pshape=(3, 5, 5), seed=1):
    rng = np.random.default_rng(seed)
    im = (rng.random(shape) * 100 + 10).astype(np.float32)
    psf = rng.random(pshape).astype(np.float32)
    psf[tuple(s // 2 for s in pshape)] += 3
    psf = (psf / psf.sum(dtype=np.float64)).astype(np.float32)
    return im, psf


def test_iter0_returns_input_and_iter1_flat_start():
    im, psf = _case()
    out, err, bg, n = po.shb(im, psf, niter=0)
    assert n == 0 and np.array_equal(out, im)
    out, err, bg, n = po.shb(im, psf, niter=1)
    wM, wN, wP = po.work_size(2, im.shape, psf.shape)
    c = np.float32(np.float32(im.astype(np.float64).sum()) / np.float32(wM * wN * wP))
    assert n == 1 and np.all(out == c)
    assert bg == np.float32(max(im.min(), np.float32(1e-2)))


def test_delta_psf_recovers_image():
    im, _ = _case()
    d = np.zeros((3, 5, 5), np.float32)
    d[1, 2, 2] = 1
    out, _, _, _ = po.shb(im, d, niter=2)
    np.testing.assert_allclose(out, im, rtol=1e-6)
    out, err, _, _ = po.shb(im, d, niter=6, start=po.START_IDENTITY)
    np.testing.assert_allclose(out, im, rtol=1e-6)
    # y == g everywhere, so the I-divergence is 0 up to rounding
    assert np.all(np.abs(err) < 1e-5)


def _circ_conv(x, k):
    """sum_t k[t] x[s - t] (circular) by direct sums."""
    out = np.zeros_like(x)
    for idx in np.ndindex(k.shape):
        if k[idx] != 0:
            out += k[idx] * np.roll(x, idx, axis=(0, 1, 2))
    return out


def _circ_corr(x, k):
    """sum_t k[t] x[s + t] (circular) by direct sums: the adjoint."""
    out = np.zeros_like(x)
    for idx in np.ndindex(k.shape):
        if k[idx] != 0:
            out += k[idx] * np.roll(x, tuple(-i for i in idx), axis=(0, 1, 2))
    return out


def _shb_spatial(im, psf, niter):
    """deconvolve_shb (method_shb.c:77-424) in float64 with the convolutions
    done as direct circular sums over the work volume."""
    P, N, M = im.shape
    pP, pN, pM = psf.shape
    wP, wN, wM = P + pP - 1, N + pN - 1, M + pM - 1
    # OTF kernel: PSF inserted at the corner, circularly shifted so that the
    # argmax (first occurrence) sits at index 0
    Z = np.zeros((wP, wN, wM))
    Z[:pP, :pN, :pM] = psf
    mid = np.unravel_index(np.argmax(Z), Z.shape)
    K = np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2))
    box = np.zeros((wP, wN, wM))
    box[:P, :N, :M] = 1
    P1 = _circ_corr(box, K)
    W = np.where(P1 > 0.01, 1.0 / np.where(P1 > 0.01, P1, 1), 0.0)
    bg = max(float(im.min()), 1e-2)
    g = np.zeros((wP, wN, wM))
    g[:P, :N, :M] = im
    x = np.full((wP, wN, wM), im.astype(np.float64).sum() / (wM * wN * wP))
    xp = x.copy()
    errs = []
    for it in range(niter):
        a = min(max((it - 1.0) / (it + 2.0), 0.0), 1.0)
        p = np.maximum(x + a * (x - xp), bg)
        y = _circ_conv(p, K)
        yi = y[:P, :N, :M]
        msk = (yi > 0) & (im > 0)
        errs.append(np.sum(im[msk] * np.log(im[msk] / yi[msk]) - (im[msk] - yi[msk])) / im.size)
        r = np.zeros_like(y)
        ys = np.where(np.abs(yi) < 1e-6, np.copysign(1e-6, yi), yi)
        r[:P, :N, :M] = im / ys
        xn = _circ_corr(r, K) * p * W
        xp = x
        x = np.maximum(xn, bg)
    return xp[:P, :N, :M], np.array(errs)


@pytest.mark.parametrize("niter", [1, 2, 4])
def test_oracle_vs_spatial_restatement(niter):
    im, psf = _case(shape=(4, 6, 7), pshape=(3, 3, 5), seed=3)
    # an asymmetric PSF makes a transposed or unconjugated OTF visible
    psf = psf * np.linspace(0.5, 1.5, psf.size).reshape(psf.shape).astype(np.float32)
    psf = (psf / psf.sum(dtype=np.float64)).astype(np.float32)
    ref, rerr = _shb_spatial(im.astype(np.float64), psf.astype(np.float64), niter)
    out, err, _, n = po.shb(im, psf, niter=niter, dtype=np.float64)
    assert n == niter
    assert rel_l2(out, ref) < 1e-12
    np.testing.assert_allclose(err, rerr, rtol=1e-6)
    out32, err32, _, _ = po.shb(im, psf, niter=niter)
    assert rel_l2(out32, ref) < 1e-5


def test_padding_invariance_bq2():
    im, psf = _case(shape=(6, 10, 12), pshape=(3, 5, 5), seed=4)
    ref, e1, _, _ = po.shb(im, psf, niter=9, dtype=np.float64)
    pad, e2, _, _ = po.shb(im, psf, niter=9, dtype=np.float64, work=(20, 16, 10))
    assert rel_l2(pad, ref) < 1e-12
    np.testing.assert_allclose(e1, e2, rtol=1e-12)


def test_float32_tracks_float64():
    im, psf = _case(shape=(8, 14, 16), pshape=(5, 5, 5), seed=5)
    o32, _, _, _ = po.shb(im, psf, niter=15)
    o64, _, _, _ = po.shb(im, psf, niter=15, dtype=np.float64)
    assert rel_l2(o32, o64) < 1e-5


def test_rel_stopping_iterator():
    """dw_iterator_next REL rule (dw.c:235-268): stop before iteration k when
    |e_{k-1} - e_{k-2}| / e_{k-1} < err_rel, with iteration 1 compared
    against 2*e_0*err_rel."""
    im, psf = _case(shape=(8, 14, 16), pshape=(5, 5, 5), seed=6)
    _, err, _, n = po.shb(im, psf, iter_type=po.ITER_REL, err_rel=0.05, maxiter=50)
    assert 2 <= n < 50
    e = err.astype(np.float32)
    for k in range(2, n):
        assert abs(e[k - 1] - e[k - 2]) / e[k - 1] >= np.float32(0.05)
    assert abs(e[n - 1] - e[n - 2]) / e[n - 1] < np.float32(0.05)


def test_setup_weights_closed_form():
    """W * P1 = 1 where P1 > 0.01, 0 elsewhere (method_shb.c:216-253), and
    a delta PSF gives W = 1 on the image box, 0 outside."""
    im, psf = _case(shape=(6, 8, 10), pshape=(3, 5, 5), seed=7)
    cK, W, x0, bg, has_w = po.shb_setup(im, psf, dtype=np.float64)
    P, N, M = im.shape
    wP, wN, wM = W.shape
    box = np.zeros(W.shape)
    box[:P, :N, :M] = 1
    Z = np.zeros(W.shape)
    Z[:3, :5, :5] = psf
    mid = np.unravel_index(np.argmax(Z), Z.shape)
    K = np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2))
    P1 = _circ_corr(box, K)
    np.testing.assert_allclose(W[P1 > 0.01] * P1[P1 > 0.01], 1.0, rtol=1e-12)
    assert np.all(W[P1 <= 0.01] == 0)
    np.testing.assert_allclose(cK, np.fft.rfftn(K), atol=1e-12)
    d = np.zeros((3, 5, 5))
    d[1, 2, 2] = 1
    _, Wd, _, _, _ = po.shb_setup(im, d, dtype=np.float64)
    np.testing.assert_allclose(Wd, box, atol=1e-12)
