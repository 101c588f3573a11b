"""Seeded synthetic microscopy volumes: point sources or Gaussian blobs on a
constant background, blurred by a PSF, Poisson noise.  Volumes are shaped
(P, N, M), x fastest."""
import numpy as np


def gaussian_psf(M, N, P, sx=1.5, sy=1.5, sz=2.0):
    """Separable Gaussian PSF centred at the middle voxel, sum 1."""
    z = np.arange(P) - (P - 1) / 2.0
    y = np.arange(N) - (N - 1) / 2.0
    x = np.arange(M) - (M - 1) / 2.0
    g = (np.exp(-0.5 * (z / sz) ** 2)[:, None, None] * np.exp(-0.5 * (y / sy) ** 2)[None, :, None]
         * np.exp(-0.5 * (x / sx) ** 2)[None, None, :])
    return (g / g.sum()).astype(np.float32)


def point_sources(shape, n, amp=(500.0, 5000.0), rng=None):
    """n deltas at uniform integer voxels, amplitudes U(amp)."""
    rng = rng or np.random.default_rng(0)
    P, N, M = shape
    V = np.zeros(shape, dtype=np.float64)
    pz = rng.integers(0, P, n)
    py = rng.integers(0, N, n)
    px = rng.integers(0, M, n)
    a = rng.uniform(amp[0], amp[1], n)
    np.add.at(V, (pz, py, px), a)
    return V


def gaussian_blobs(shape, n, sigma=(1.0, 3.0), amp=(200.0, 4000.0), rng=None):
    """n isotropic Gaussian blobs, sigma ~ U(sigma) px, peak ~ U(amp)."""
    rng = rng or np.random.default_rng(0)
    P, N, M = shape
    V = np.zeros(shape, dtype=np.float64)
    cz = rng.uniform(0, P, n)
    cy = rng.uniform(0, N, n)
    cx = rng.uniform(0, M, n)
    s = rng.uniform(sigma[0], sigma[1], n)
    a = rng.uniform(amp[0], amp[1], n)
    for i in range(n):
        r = int(np.ceil(4 * s[i]))
        z0, z1 = max(0, int(cz[i]) - r), min(P, int(cz[i]) + r + 1)
        y0, y1 = max(0, int(cy[i]) - r), min(N, int(cy[i]) + r + 1)
        x0, x1 = max(0, int(cx[i]) - r), min(M, int(cx[i]) + r + 1)
        gz = np.exp(-0.5 * ((np.arange(z0, z1) - cz[i]) / s[i]) ** 2)
        gy = np.exp(-0.5 * ((np.arange(y0, y1) - cy[i]) / s[i]) ** 2)
        gx = np.exp(-0.5 * ((np.arange(x0, x1) - cx[i]) / s[i]) ** 2)
        V[z0:z1, y0:y1, x0:x1] += a[i] * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
    return V


def blur(V, psf, device=None):
    """Linear ('same') convolution of V with a centred PSF, zero padded.
    float64 numpy by default; device='cuda' uses torch in float32 (large
    volumes)."""
    P, N, M = V.shape
    pP, pN, pM = psf.shape
    full = (P + pP - 1, N + pN - 1, M + pM - 1)
    if device is None:
        ax = (0, 1, 2)
        F = np.fft.rfftn(V, full, axes=ax) * np.fft.rfftn(psf.astype(np.float64), full, axes=ax)
        out = np.fft.irfftn(F, full, axes=ax)
    else:
        import torch
        a = torch.as_tensor(np.asarray(V, dtype=np.float32), device=device)
        b = torch.as_tensor(np.asarray(psf, dtype=np.float32), device=device)
        out = torch.fft.irfftn(torch.fft.rfftn(a, full) * torch.fft.rfftn(b, full), full)
        out = out.cpu().numpy().astype(np.float64)
    cz, cy, cx = (pP - 1) // 2, (pN - 1) // 2, (pM - 1) // 2
    return out[cz:cz + P, cy:cy + N, cx:cx + M]


def noisy_image(V, background, rng=None):
    """Poisson(V + background) as float32."""
    rng = rng or np.random.default_rng(0)
    lam = np.clip(V, 0, None) + background
    return rng.poisson(lam).astype(np.float32)
