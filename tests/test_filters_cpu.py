"""LP start smoothing (deconwolf_amd.filters, restating fim_gsmooth) and
the Born-Wolf PSF generator (deconwolf_amd.psf_bw) against closed forms."""
import math

import numpy as np
import pytest
from scipy import special

from deconwolf_amd import filters, psf_bw


def brute_conv1(v, K):
    """fim_conv1_vector, normalized, written as plain loops (one line)."""
    n, nk = len(v), len(K)
    k2 = (nk - 1) // 2
    out = np.empty(n, np.float32)
    for i in range(n):
        acc, kacc, full = 0.0, 0.0, True
        for t in range(nk):
            j = i - k2 + t
            if 0 <= j < n:
                acc += float(np.float32(K[t] * v[j]))
                kacc += float(K[t])
            else:
                full = False
        out[i] = np.float32(acc if (full and nk <= n) else acc / kacc)
    return out


def test_gaussian_kernel():
    K = filters.gaussian_kernel(8.0)
    assert K.shape == (49,)
    assert abs(K.astype(np.float64).sum() - 1) < 1e-6
    assert np.allclose(K, K[::-1])
    assert np.argmax(K) == 24


@pytest.mark.parametrize("n", [5, 20, 60])
def test_conv1_matches_loops(n):
    rng = np.random.default_rng(n)
    v = rng.random(n).astype(np.float32) * 100
    K = filters.gaussian_kernel(3.0)
    got = filters.conv1_normalized(v[None, None, :], K, 2)[0, 0]
    np.testing.assert_array_equal(got, brute_conv1(v, K))


def test_gsmooth_constant_and_trim():
    V = np.full((3, 20, 30), 7.0, np.float32)
    out = filters.gsmooth(V, 8.0)
    np.testing.assert_allclose(out, 7.0, rtol=1e-6)   # normalised edges keep constants
    # handed to the library as an x-fastest buffer
    assert out.flags["C_CONTIGUOUS"] and out.dtype == np.float32


def test_bw_in_focus_is_airy():
    """At z = 0 the Born-Wolf integral is int_0^1 J0(v rho) rho drho = J1(v)/v."""
    lam, NA = 520.0, 1.4
    r = np.linspace(0, 3000, 301)
    prof = psf_bw.radial_profile(r, 0.0, lam, NA, 1.515)
    v = 2 * math.pi / lam * NA * r
    airy = np.where(v > 0, special.j1(v) / np.where(v > 0, v, 1), 0.5) ** 2
    np.testing.assert_allclose(prof, airy, rtol=1e-9, atol=1e-12)


def test_bw_psf_symmetries():
    V = psf_bw.bw_psf(31, 31, 15, 520.0, dz_nm=200.0)
    assert abs(V.astype(np.float64).sum() - 1) < 1e-5
    np.testing.assert_allclose(V, V[::-1], rtol=1e-6)          # z
    np.testing.assert_allclose(V, V[:, ::-1], rtol=1e-6)       # y
    np.testing.assert_allclose(V, V[:, :, ::-1], rtol=1e-6)    # x
    np.testing.assert_allclose(V, np.transpose(V, (0, 2, 1)), rtol=1e-6)  # x <-> y
    assert np.unravel_index(np.argmax(V), V.shape) == (7, 15, 15)


def test_lanczos5_interpolation():
    """lanczos.c:102-160: samples are returned at integer positions (and
    within 1e-9 of them), indices below 0 mirror, a smooth profile is
    reproduced between samples."""
    v = np.cos(0.05 * np.arange(60)) + 2.0
    idx, w = psf_bw.lanczos5_weights(np.array([3.0, 7.0 + 5e-10, 12.0 - 5e-10]))
    got = (v[idx] * w).sum(-1)
    np.testing.assert_array_equal(got, [v[3], v[7], v[12]])
    x = np.linspace(0.3, 40.7, 97)
    idx, w = psf_bw.lanczos5_weights(x)
    assert idx.min() >= 0 and np.all(idx[:, :5] == np.abs(np.floor(x)[:, None].astype(int) - np.arange(5)))
    np.testing.assert_allclose((v[idx] * w).sum(-1), np.cos(0.05 * x) + 2.0, rtol=2e-3)


def test_gauss_legendre_rule():
    x, w = psf_bw.gauss_legendre01(12)
    for k in range(24):   # exact for polynomials up to degree 23
        assert abs((w * x ** k).sum() - 1.0 / (k + 1)) < 1e-13
