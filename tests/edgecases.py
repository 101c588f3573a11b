"""Inputs and closed forms shared by the oracle pins (tests/test_oracle_pins.py)
and the GPU edge-case tests (tests/test_gpu_edges.py).

Each case is built so that one step of deconvolve_shb (src/method_shb.c)
has a value fixed by the mathematics alone:

* delta PSF, flat start below bg: p_0 = bg everywhere, so y_0 = H p_0 = bg
  and the iteration-0 error is the metric of (g, bg) — MSE (dw.c:1461-1490)
  and I-divergence (dw.c:1492-1524) in closed form;
* tiny image, bg 0: the flat start c is below 1e-6, so y_0 = c and every
  image voxel takes the |y| < 1e-6 branch of the ratio (method_shb.c:41-42):
  x_1 = (g / 1e-6) * c on the image box (delta PSF, W = 1 there);
* a PSF with two equal maxima: the OTF shift uses the lowest linear index
  (fim_argmax_max, fim.c:307-319).
"""
import numpy as np


def delta_psf(n=9, nz=9):
    d = np.zeros((nz, n, n), np.float32)
    d[nz // 2, n // 2, n // 2] = 1
    return d


def metric_case(seed=0, shape=(10, 12, 14)):
    """Image in [100, 110] with a 9x9x9 delta PSF: the flat start
    sum(g)/wMNP is far below bg = min(g)."""
    rng = np.random.default_rng(seed)
    im = (100 + 10 * rng.random(shape)).astype(np.float32)
    return im, delta_psf()


def metric_closed_forms(im):
    """(MSE_0, IDIV_0) for y_0 == bg = min(im) (bg_auto), in float64."""
    g = im.astype(np.float64)
    bg = float(max(im.min(), np.float32(1e-2)))
    mse = np.mean((bg - g) ** 2)
    pos = g > 0
    idiv = np.sum(g[pos] * np.log(g[pos] / bg) - (g[pos] - bg)) / g.size
    return mse, idiv, bg


def flat_constant(im, work):
    """Flat start of method_shb.c:255-268 for deconwolf's work size."""
    wM, wN, wP = work
    return np.float32(np.float32(im.astype(np.float64).sum()) / np.float32(wM * wN * wP))


def eps_case(seed=0, shape=(8, 16, 16)):
    """Values in [1e-8, 5e-6] with zero patches; 9x9x9 delta PSF.  With bg 0
    the flat start is about 5.6e-7 < 1e-6."""
    rng = np.random.default_rng(seed)
    im = (1e-8 + 5e-6 * rng.random(shape)).astype(np.float32)
    im[:, 2:6, 3:9] = 0
    im[5, :, :4] = 0
    return im, delta_psf()


def eps_closed_form(im, c):
    """x_1 on the image box: IFFT(conj(K) FFT(ratio)) * p_0 * W with K = 1,
    W = 1, p_0 = c, ratio = g / (float)copysign(1e-6, y_0 = c)."""
    return (im / np.float32(1e-6)).astype(np.float32) * np.float32(c)


def tie_psf(seed=0, shape=(5, 7, 7)):
    """Random PSF (sum 1) with its maximum at two voxels; returns the PSF
    and the two linear indices (x fastest), lower first."""
    rng = np.random.default_rng(seed)
    k = (0.1 + rng.random(shape)).astype(np.float32)
    a = (1, 2, 2)   # (z, y, x)
    b = (3, 4, 5)
    k[a] = 4.0
    k[b] = 4.0
    k = (k / k.sum(dtype=np.float64)).astype(np.float32)
    assert k[a] == k[b] == k.max()
    ia = np.ravel_multi_index(a, shape)
    ib = np.ravel_multi_index(b, shape)
    return k, (min(ia, ib), max(ia, ib)), (a, b)
