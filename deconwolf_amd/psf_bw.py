"""Born-Wolf widefield PSF, the model of deconwolf's `dw_bw`
(src/dw_bwpsf.c:538-775, src/bw_gsl.c:79-157, src/lanczos.c:102-160).

Per plane at defocus z the radial amplitude is

    A(r, z) = int_0^1 J0(k0 NA r rho) exp(-i k0 NA^2 z rho^2 / (2 ni)) rho drho

(bw_gsl.c:79-111, k0 = 2 pi / lambda) and the intensity |A|^2 is sampled
at 17 radii per pixel (dw_bwpsf.c:181, 660-670).  Each pixel holds the mean
over the pixel square of the Lanczos-5 interpolation of that profile
(dw_bwpsf.c:563-634, lanczos.c:102-160), zero for pixels whose centre lies
beyond radmax; one octant of each plane is computed and mirrored
(dw_bwpsf.c:726-766), and planes are mirrored in z (:553-558).

dw_bw integrates adaptively with GSL (QAG, tolerance 1e-6); here the rho
integral is Gauss-Legendre with enough nodes for the oscillations, and the
pixel mean an 8 x 8 Gauss-Legendre product rule.  GSL is not available, so
agreement with dw_bw itself is not checked; the closed form at focus
(A = J1(v)/v, v = k0 NA r), the interpolation and the symmetries are.
The same algorithm runs on the GPU (`bw_psf_device`, libdwhip's
dwhip_bw_psf_device).
"""
import ctypes

import numpy as np
from scipy import special

OVERSAMPLING = 17   # radial samples per pixel (dw_bwpsf.c:181)
PIXEL_NODES = 8     # Gauss-Legendre nodes per axis of the pixel mean


def gauss_legendre01(n):
    """Nodes and weights of the n-point Gauss-Legendre rule on [0, 1]."""
    x, w = np.polynomial.legendre.leggauss(n)
    return 0.5 * (x + 1.0), 0.5 * w


def n_nodes(rmax_nm, z_nm, lambda_nm, NA, ni):
    """Gauss-Legendre nodes for the rho integral: enough for the
    oscillations of J0 (argument up to k0 NA rmax) and of the defocus phase."""
    k0 = 2.0 * np.pi / lambda_nm
    vmax = k0 * NA * rmax_nm
    wmax = abs(k0 * NA * NA * z_nm / (2.0 * ni))
    return int(64 + 2.0 * (vmax + wmax) / np.pi)


def radial_profile(r_nm, z_nm, lambda_nm, NA, ni, nquad=None):
    """|A(r, z)|^2 for an array of radii (nm) at one defocus z (nm)."""
    r = np.asarray(r_nm, dtype=np.float64)
    k0 = 2.0 * np.pi / lambda_nm
    if nquad is None:
        nquad = n_nodes(float(r.max()) if r.size else 0.0, z_nm, lambda_nm, NA, ni)
    rho, w = gauss_legendre01(nquad)
    phase = -k0 * NA * NA * z_nm * rho * rho / (2.0 * ni)
    out = np.empty(r.shape, dtype=np.float64)
    flat = r.ravel()
    o = out.ravel()
    step = 512
    for i in range(0, flat.size, step):
        rr = flat[i:i + step, None]
        b = special.j0(k0 * NA * rr * rho[None, :]) * rho[None, :] * w[None, :]
        re = (b * np.cos(phase)[None, :]).sum(axis=1)
        im = (b * np.sin(phase)[None, :]).sum(axis=1)
        o[i:i + step] = re * re + im * im
    return out


def _sinc(t):
    x = np.pi * t
    small = np.abs(x) <= 1e-6
    xs = np.where(small, 1.0, x)
    return np.where(small, 1 - x * x / 6 + x ** 4 / 120, np.sin(xs) / xs)


def lanczos5_weights(x):
    """Indices (..., 10) and weights (..., 10) of the Lanczos-5 interpolation
    at positions x >= 0 of a profile symmetric about 0 (lanczos.c:102-160):
    five samples before (mirrored at 0) and five after floor(x); positions
    within 1e-9 of a sample take that sample."""
    x = np.asarray(x, dtype=np.float64)
    n = np.floor(x).astype(np.int64)
    d = x - n
    kk = np.arange(5)
    idx = np.concatenate([np.abs(n[..., None] - kk), n[..., None] + 1 + kk], axis=-1)
    tb = d[..., None] + kk
    ta = (kk + 1) - d[..., None]
    t = np.concatenate([tb, ta], axis=-1)
    wts = _sinc(t) * _sinc(t / 5.0)
    lo = d < 1e-9
    hi = d > 1 - 1e-9
    wts = np.where((lo | hi)[..., None], 0.0, wts)
    wts[..., 0] = np.where(lo, 1.0, wts[..., 0])
    wts[..., 5] = np.where(hi, 1.0, wts[..., 5])
    return idx, wts


def _octant(N):
    """Pixels (x, y) with x <= y of the first quadrant (dw_bwpsf.c:729-730)."""
    h = N // 2
    xs, ys = [], []
    for x in range(h + 1):
        for y in range(x, h + 1):
            xs.append(x)
            ys.append(y)
    return np.array(xs), np.array(ys)


def _mirror(plane_vals, xs, ys, N):
    """Write the octant values into an N x N plane (dw_bwpsf.c:750-765)."""
    V = np.zeros((N, N), np.float64)
    xf = N - xs - 1
    yf = N - ys - 1
    for a, b in ((xs, ys), (xs, yf), (ys, xs), (ys, xf), (xf, ys), (xf, yf), (yf, xs), (yf, xf)):
        V[b, a] = plane_vals   # V[x + N*y] with x = a, y = b
    return V


def bw_psf(M, N, P, lambda_nm, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=200.0, normalize=True):
    """Born-Wolf PSF volume shaped (P, N, M) with M == N (dw_bw writes
    N x N x P), centred at ((N-1)/2, (N-1)/2, (P-1)/2), float32; optionally
    normalised to sum 1 (dw_run normalises the PSF it reads)."""
    if M != N:
        raise ValueError("dw_bw makes square planes (M == N)")
    x0 = (N - 1) / 2.0
    radmax = int(round(np.sqrt((N - x0) ** 2 + (N - x0) ** 2))) + 1
    nr = (radmax + 1) * OVERSAMPLING + 2
    rs = np.arange(nr, dtype=np.float64) / OVERSAMPLING
    xs, ys = _octant(N)
    gx, gw = gauss_legendre01(PIXEL_NODES)
    off = gx - 0.5
    X = (xs - x0)[:, None, None] + off[None, :, None]
    Y = (ys - x0)[:, None, None] + off[None, None, :]
    r2 = X * X + Y * Y
    r = np.where(np.abs(r2) > 1e-10, np.sqrt(np.abs(r2)), 0.0)
    idx, wts = lanczos5_weights(r * OVERSAMPLING)
    pw = (gw[:, None] * gw[None, :])[None, :, :]
    inside = np.sqrt((xs - x0) ** 2 + (ys - x0) ** 2) < radmax
    V = np.zeros((P, N, N), dtype=np.float64)
    for z in range((P + 1) // 2):
        defocus = float(np.float32(dz_nm * (z - (P - 1.0) / 2.0)))
        prof = radial_profile(rs * dx_nm, defocus, lambda_nm, NA, ni,
                              nquad=n_nodes(rs[-1] * dx_nm, defocus, lambda_nm, NA, ni))
        val = (prof[idx] * wts).sum(axis=-1)
        pix = np.where(inside, (val * pw).sum(axis=(1, 2)), 0.0)
        V[z] = _mirror(pix, xs, ys, N)
        V[P - 1 - z] = V[z]
    V = V.astype(np.float32)
    if normalize:
        V = (V.astype(np.float64) / V.astype(np.float64).sum()).astype(np.float32)
    return V


def bw_psf_device(N, P, lambda_nm, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=200.0, out=None):
    """The same PSF (not normalised, as dw_bw writes it) computed on the GPU
    by dwhip_bw_psf_device; returns a float32 (P, N, N) torch tensor."""
    import torch
    from ._lib import check, lib
    if out is None:
        out = torch.empty((P, N, N), dtype=torch.float32, device="cuda")
    check(lib().dwhip_bw_psf_device(out.data_ptr(), int(N), int(P), float(lambda_nm), float(NA), float(ni),
                                    float(dx_nm), float(dz_nm), torch.cuda.current_stream().cuda_stream),
          "dwhip_bw_psf_device")
    return out
