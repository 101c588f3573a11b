"""Born-Wolf widefield PSF, the model of deconwolf's `dw_bw`
(src/dw_bwpsf.c:538-775, src/bw_gsl.c:79-157), used to make the inputs of
the benchmark and of the parity cases.

Per plane at defocus z the radial amplitude is

    A(r, z) = int_0^1 J0(k0 NA r rho) exp(-i k0 NA^2 z rho^2 / (2 ni)) rho drho

(bw_gsl.c:79-111, k0 = 2 pi / lambda) and the intensity is |A|^2.  Each
pixel holds the mean of |A|^2 over the pixel square (dw_bwpsf.c:600-637),
the plane is filled using the 8-fold symmetry of the square grid
(:753-766) and the volume using the symmetry in z (:547-558).

dw_bw integrates with GSL's adaptive QAG and interpolates the radial
profile with Lanczos-5; here the rho integral is Gauss-Legendre and the
pixel mean is a Gauss-Legendre product rule over a linearly interpolated,
finely sampled profile.  GSL is not available, so agreement with dw_bw
itself is not checked; the closed form at focus (A = J1(v)/v,
v = k0 NA r) and the symmetries are.
"""
import numpy as np
from scipy import special


def radial_profile(r_nm, z_nm, lambda_nm, NA, ni, nquad=None):
    """|A(r, z)|^2 for an array of radii (nm) at one defocus z (nm)."""
    r = np.asarray(r_nm, dtype=np.float64)
    k0 = 2.0 * np.pi / lambda_nm
    vmax = k0 * NA * (float(r.max()) if r.size else 0.0)
    wmax = abs(k0 * NA * NA * z_nm / (2.0 * ni))
    if nquad is None:
        # enough nodes for the oscillations of J0 and of the defocus phase
        nquad = int(64 + 2.0 * (vmax + wmax) / np.pi)
    x, w = np.polynomial.legendre.leggauss(nquad)
    rho = 0.5 * (x + 1.0)
    w = 0.5 * w
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


def bw_psf(M, N, P, lambda_nm, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=200.0,
           oversampling=8, pixel_nodes=4, normalize=True):
    """Born-Wolf PSF volume shaped (P, N, M), centred at ((M-1)/2, (N-1)/2,
    (P-1)/2), float32, optionally normalised to sum 1."""
    x0 = (M - 1) / 2.0
    y0 = (N - 1) / 2.0
    radmax = int(round(np.sqrt((M - x0) ** 2 + (N - y0) ** 2))) + 1
    nr = (radmax + 1) * oversampling + 2
    rs = np.arange(nr, dtype=np.float64) / oversampling          # pixels
    gx, gw = np.polynomial.legendre.leggauss(pixel_nodes)
    sub = 0.5 * gx                                                # offsets in a pixel
    sw = 0.5 * gw
    xs = np.arange(M, dtype=np.float64) - x0
    ys = np.arange(N, dtype=np.float64) - y0
    # sub-pixel radii for every pixel: (N, M, q, q)
    X = xs[None, :, None, None] + sub[None, None, :, None]
    Y = ys[:, None, None, None] + sub[None, None, None, :]
    R = np.sqrt(X * X + Y * Y) * oversampling
    W2 = (sw[:, None] * sw[None, :])[None, None, :, :]
    V = np.zeros((P, N, M), dtype=np.float64)
    for z in range((P + 1) // 2):
        defocus = dz_nm * (z - (P - 1.0) / 2.0)
        prof = radial_profile(rs * dx_nm, defocus, lambda_nm, NA, ni)
        i0 = np.minimum(np.floor(R).astype(np.int64), nr - 2)
        f = R - i0
        val = prof[i0] * (1.0 - f) + prof[i0 + 1] * f
        V[z] = (val * W2).sum(axis=(2, 3))
        V[P - 1 - z] = V[z]
    if normalize:
        V /= V.sum()
    return V.astype(np.float32)
