"""psf_autocrop (src/dw.c:1764-1997) on the host side of the method.

The crop index is integer output of float arithmetic, so it must match
deconwolf's bit for bit.  The brute force below is a direct loop over the
reference's steps with every sum and product in float32; the threshold
case is built so that float32 and float64 plane sums choose different
crops.
"""
import numpy as np
import pytest

from deconwolf_amd import method_shb as ms


def brute_autocrop(psf, imshape, bq=2, xycropfactor=0.001):
    """(m0, m1, n0, n1, p0, p1) of the byImage crop and `first` of the XY
    crop, by plain loops over the reference's arithmetic."""
    P, N, M = imshape
    p, n, m = psf.shape
    mopt, nopt, popt = (M - 1) * 2 + 1, (N - 1) * 2 + 1, (P - 1) * 2 + 1
    if bq == 0:
        mopt, nopt, popt = M, N, P
    box = (0, m - 1, 0, n - 1, 0, p - 1)
    if not p < popt and (m > mopt or n > nopt or p > popt):
        m0, m1, n0, n1, p0, p1 = box
        if m > mopt:
            m0, m1 = (m - mopt) // 2, m1 - (m - mopt) // 2
        if n > nopt:
            n0, n1 = (n - nopt) // 2, n1 - (n - nopt) // 2
        if p > popt:
            p0, p1 = (p - popt) // 2, p1 - (p - popt) // 2
        box = (m0, m1, n0, n1, p0, p1)
        psf = psf[p0:p1 + 1, n0:n1 + 1, m0:m1 + 1]
    first = -1
    if bq > 0 and xycropfactor > 0:
        pP, pN, pM = psf.shape
        maxsum = np.float32(0)
        for xx in range(pM):
            s = np.float32(0)
            for yy in range(pN):
                for zz in range(pP):
                    s = np.float32(s + psf[zz, yy, xx])
            if s > maxsum:
                maxsum = s
        thr = np.float32(np.float32(xycropfactor) * maxsum)
        s = np.float32(0)
        while s < thr:
            first += 1
            s = np.float32(0)
            for yy in range(pN):
                for zz in range(pP):
                    s = np.float32(s + psf[zz, yy, first])
    return box, first


def crop_with(psf, box, first):
    m0, m1, n0, n1, p0, p1 = box
    c = psf[p0:p1 + 1, n0:n1 + 1, m0:m1 + 1]
    if first >= 1:
        pP, pN, pM = c.shape
        c = c[:, first:pN - first, first:pM - first]
    return c


def gauss_psf(size, sz, sxy, sigz, rng):
    z, y, x = np.meshgrid(np.arange(sz) - sz // 2, np.arange(size) - size // 2, np.arange(size) - size // 2,
                          indexing="ij")
    g = np.exp(-(x ** 2 + y ** 2) / (2 * sxy ** 2) - z ** 2 / (2 * sigz ** 2))
    g *= 1 + 0.01 * rng.random(g.shape)
    return (g / g.sum()).astype(np.float32)


@pytest.mark.parametrize("seed,size,sz,sxy,imshape,xyf", [
    (0, 31, 15, 2.0, (20, 64, 64), 0.001),
    (1, 41, 21, 4.5, (12, 64, 64), 0.001),
    (2, 25, 9, 1.2, (40, 64, 64), 0.01),
    (3, 33, 33, 3.0, (8, 12, 14), 0.001),   # byImage crop first (33 > 2*12-1 in x and y, 33 > 15 in z)
    (4, 21, 11, 2.5, (10, 30, 30), 0.2),
])
def test_autocrop_matches_brute_force(seed, size, sz, sxy, imshape, xyf):
    rng = np.random.default_rng(seed)
    psf = gauss_psf(size, sz, sxy, sxy * 1.5, rng)
    box, first = brute_autocrop(psf, imshape, 2, xyf)
    s = ms.DwOpts(xycropfactor=xyf)
    got = ms.psf_autocrop(psf, imshape, s)
    assert np.array_equal(got, crop_with(psf, box, first)), (got.shape, box, first)


def test_autocrop_bq0_skips_xy_crop():
    rng = np.random.default_rng(5)
    psf = gauss_psf(31, 15, 2.0, 3.0, rng)
    s = ms.DwOpts(xycropfactor=0.001, borderQuality=0)
    got = ms.psf_autocrop(psf, (20, 64, 64), s)
    assert got.shape == psf.shape


def test_autocrop_threshold_float_order():
    """Plane x = 0 holds 1.0 first (y = z = 0) followed by 120 values of
    2^-25: in float32, accumulated in deconwolf's order, each tiny value is
    below half an ulp of 1.0 and is lost (sum 1.0); in float64 the sum is
    1 + 120 * 2^-25.  With the threshold between the two, deconwolf's float
    arithmetic crops one plane on each side and float64 sums would not."""
    pP, pN, pM = 11, 11, 11
    psf = np.zeros((pP, pN, pM), np.float32)
    psf[:, :, 0] = np.float32(2.0 ** -25)
    psf[0, 0, 0] = 1.0
    psf[:, :, pM - 1] = psf[:, :, 0]
    psf[:, :, 1:pM - 1] = np.float32(10.0 / (pP * pN))      # plane sums about 10
    psf[pP // 2, pN // 2, pM // 2] += np.float32(1000.0)   # the largest plane
    sums = ms.xplane_sums_f32(psf)
    assert sums[0] == np.float32(1.0)
    assert psf[:, :, 0].astype(np.float64).sum() == 1.0 + 120 * 2.0 ** -25
    maxsum = sums.max()
    xyf = np.float32(1.0 + 1.5e-6) / maxsum
    thr = np.float32(np.float32(xyf) * maxsum)
    assert np.float32(1.0) < thr < 1.0 + 120 * 2.0 ** -25
    box, first = brute_autocrop(psf, (40, 64, 64), 2, float(xyf))
    assert first == 1
    got = ms.psf_autocrop(psf, (40, 64, 64), ms.DwOpts(xycropfactor=float(xyf)))
    assert got.shape == (pP, pN - 2, pM - 2)
    assert np.array_equal(got, psf[:, 1:pN - 1, 1:pM - 1])


def test_autocrop_factor_above_one_is_an_error():
    rng = np.random.default_rng(6)
    psf = gauss_psf(15, 7, 2.0, 3.0, rng)
    with pytest.raises(ValueError):
        ms.psf_autocrop(psf, (20, 64, 64), ms.DwOpts(xycropfactor=1.5))
