"""Pins the Richardson-Lucy restatement of the oracle (deconvolve_rl,
src/method_rl.c:28-400) before it checks the device RL path.

deconwolf has no RL vectors either; the restatement is pinned by closed
forms (--iter 0 copy, a delta PSF returns the image after one update),
RL's flux conservation (with a normalised PSF, periodic boundaries and no
clamp, sum(x_k) = sum(g) after every update: a transposed or unconjugated
adjoint breaks it for an asymmetric PSF), a constant image as a fixed
point, a spatial-domain restatement by direct sums, and the padding
invariance of border quality 2.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from test_oracle import _case, _circ_conv, _circ_corr, rel_l2


def _asym(psf):
    psf = psf * np.linspace(0.5, 1.5, psf.size).reshape(psf.shape).astype(np.float32)
    return (psf / psf.sum(dtype=np.float64)).astype(np.float32)


def test_rl_iter0_copy_and_delta_psf():
    im, psf = _case()
    out, err, bg, n = po.rl(im, psf, niter=0)
    assert n == 0 and np.array_equal(out, im)
    d = np.zeros((3, 5, 5), np.float32)
    d[1, 2, 2] = 1
    # x_1 = H^T(g / H x0) * x0 * W = g on the image box (W = 1 there)
    out, err, bg, n = po.rl(im, d, niter=1)
    assert n == 1
    np.testing.assert_allclose(out, np.maximum(im, bg), rtol=1e-6)
    out, err, _, _ = po.rl(im, d, niter=5)
    np.testing.assert_allclose(out, np.maximum(im, bg), rtol=1e-6)
    assert np.all(np.abs(err[1:]) < 1e-5)


def test_rl_flux_conservation_periodic():
    """bq 0 with the image filling the work volume, bg = 0 (no clamp):
    sum_i x_{k+1,i} = sum_j g_j / y_j sum_i h(j - i) x_{k,i} = sum_j g_j."""
    im, psf = _case(shape=(6, 10, 12), pshape=(3, 5, 5), seed=11)
    psf = _asym(psf)
    for niter in (1, 2, 5):
        out, _, _, _ = po.rl(im, psf, niter=niter, dtype=np.float64, border_quality=0,
                             bg_auto=0, bg=0.0)
        np.testing.assert_allclose(out.sum(), im.astype(np.float64).sum(), rtol=1e-12)
    # a non-conjugated adjoint (H in place of H^T) would not conserve it
    f = np.asarray(np.random.default_rng(3).random(im.shape) + 1.0)
    Z = np.zeros(im.shape)
    Z[:3, :5, :5] = psf
    mid = np.unravel_index(np.argmax(Z), Z.shape)
    K = np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2))
    r = im / _circ_conv(f, K)
    assert abs((_circ_corr(r, K) * f).sum() / im.astype(np.float64).sum() - 1) < 1e-12
    assert abs((_circ_conv(r, K) * f).sum() / im.astype(np.float64).sum() - 1) > 1e-4


def test_rl_constant_image_fixed_point():
    im = np.full((6, 10, 12), 37.5, np.float32)
    _, psf = _case(shape=(6, 10, 12), pshape=(3, 5, 5), seed=12)
    out, err, _, _ = po.rl(im, _asym(psf), niter=4, dtype=np.float64, border_quality=0)
    np.testing.assert_allclose(out, 37.5, rtol=1e-12)
    assert np.all(np.abs(err) < 1e-9)


def _rl_spatial(im, psf, niter, bq2=True):
    """deconvolve_rl in float64 with direct circular sums over the work
    volume (border quality 2)."""
    P, N, M = im.shape
    pP, pN, pM = psf.shape
    wP, wN, wM = P + pP - 1, N + pN - 1, M + pM - 1
    Z = np.zeros((wP, wN, wM))
    Z[:pP, :pN, :pM] = psf
    mid = np.unravel_index(np.argmax(Z), Z.shape)
    K = np.roll(Z, tuple(-m for m in mid), axis=(0, 1, 2))
    box = np.zeros((wP, wN, wM))
    box[:P, :N, :M] = 1
    P1 = _circ_corr(box, K)
    W = np.where(P1 > 0.01, 1.0 / np.where(P1 > 0.01, P1, 1), 0.0)
    bg = max(float(im.min()), 1e-2)
    f = np.full((wP, wN, wM), im.astype(np.float64).sum() / (wM * wN * wP))
    errs = []
    for _ in range(niter):
        y = _circ_conv(f, K)
        yi = y[:P, :N, :M]
        msk = (yi > 0) & (im > 0)
        errs.append(np.sum(im[msk] * np.log(im[msk] / yi[msk]) - (im[msk] - yi[msk])) / im.size)
        r = np.full_like(y, float(np.float32(1e-6)))
        r[:P, :N, :M] = np.where(yi > 0, im / np.where(yi > 0, yi, 1), bg)
        f = np.maximum(_circ_corr(r, K) * f * W, bg)
    return f[:P, :N, :M], np.array(errs)


@pytest.mark.parametrize("niter", [1, 3])
def test_rl_vs_spatial_restatement(niter):
    im, psf = _case(shape=(4, 6, 7), pshape=(3, 3, 5), seed=13)
    psf = _asym(psf)
    ref, rerr = _rl_spatial(im.astype(np.float64), psf.astype(np.float64), niter)
    out, err, _, n = po.rl(im, psf, niter=niter, dtype=np.float64)
    assert n == niter
    assert rel_l2(out, ref) < 1e-12
    np.testing.assert_allclose(err, rerr, rtol=1e-6)
    out32, _, _, _ = po.rl(im, psf, niter=niter)
    assert rel_l2(out32, ref) < 1e-5


def test_rl_padding_invariance_bq2():
    im, psf = _case(shape=(6, 10, 12), pshape=(3, 5, 5), seed=14)
    psf = _asym(psf)
    ref, e1, _, _ = po.rl(im, psf, niter=7, dtype=np.float64)
    pad, e2, _, _ = po.rl(im, psf, niter=7, dtype=np.float64, work=(20, 16, 10))
    assert rel_l2(pad, ref) < 1e-12
    np.testing.assert_allclose(e1, e2, rtol=1e-12)
