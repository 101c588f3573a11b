"""Gaussian smoothing used by the LP start condition of deconvolve_shb
(method_shb.c:278-287: fim_gsmooth(image, 8)).  Host-side set-up only; the
smoothed image is handed to the GPU as the start volume.

Restates fim_gsmooth / fim_gsmooth_aniso (src/fim.c:2240-2278),
gaussian_kernel (fim.c:2077-2105), fim_convn1 (fim.c:3348-3400) and the
normalised edge handling of fim_conv1_vector (fim.c:1944-2056): products
of float32 values, sums in float64, one float32 result per sample.
"""
import math

import numpy as np


def gaussian_kernel(sigma):
    """fim.c:2077-2105: 2*ceil(3 sigma)+1 taps, normalised to sum 1."""
    sigma = np.float32(sigma)
    ln = int(math.ceil(3.0 * float(sigma)))
    ln = max(ln, 1)
    n = 2 * ln + 1
    mid = np.float32((n - 1) // 2)
    s2 = np.float32(float(sigma) ** 2)
    k0 = np.float32(1.0 / float(sigma) / math.sqrt(2.0 * math.pi))
    K = np.empty(n, np.float32)
    acc = 0.0
    for kk in range(n):
        x = np.float32(np.float32(kk) - mid)
        K[kk] = np.float32(float(k0) * math.exp(-0.5 * float(x) ** 2 / float(s2)))
        acc += float(K[kk])
    return (K.astype(np.float64) / acc).astype(np.float32)


def conv1_normalized(V, K, axis):
    """fim_conv1_vector with normalized=1 along `axis` of V (float32)."""
    V = np.moveaxis(np.asarray(V, np.float32), axis, -1)
    nV = V.shape[-1]
    nK = K.shape[0]
    k2 = (nK - 1) // 2
    out = np.empty(V.shape, np.float32)
    Kd = K.astype(np.float64)
    for vv in range(nV):
        lo = max(0, vv - k2)
        hi = min(nV, vv + k2 + 1)
        ks = K[lo - vv + k2:hi - vv + k2]
        prod = (V[..., lo:hi] * ks).astype(np.float64)   # float products, double sums
        acc = np.zeros(V.shape[:-1], np.float64)
        for t in range(prod.shape[-1]):
            acc = acc + prod[..., t]
        full = (hi - lo) == nK and nK <= nV
        if full:
            out[..., vv] = acc.astype(np.float32)
        else:
            kacc = 0.0
            for t in range(ks.shape[0]):
                kacc += float(Kd[lo - vv + k2 + t])
            out[..., vv] = (acc / kacc).astype(np.float32)
    return np.moveaxis(out, -1, axis)


def gsmooth(V, sigma):
    """fim_gsmooth(V, M, N, P, sigma) on a (P, N, M) array; returns a new
    array.  x and y with the full kernel, z with the kernel trimmed to at
    most 2P-1 taps (fim.c:2255-2266)."""
    V = np.asarray(V, np.float32)
    K = gaussian_kernel(sigma)
    out = conv1_normalized(V, K, 2)
    out = conv1_normalized(out, K, 1)
    P = V.shape[0]
    if P > 1:
        Ka = K
        while Ka.shape[0] > 2 * P - 1:
            Ka = Ka[1:-1]
        out = conv1_normalized(out, Ka, 0)
    return np.ascontiguousarray(out)
