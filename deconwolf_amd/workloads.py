"""Synthetic benchmark workloads (seeded), prepared the way deconwolf's
dw_run prepares a method call (src/dw.c:2799-2852): PSF normalised to sum
1, autocropped, normalised again; the offset added to the image.

    c1  128x128x32, 40 point sources U(500,5000), background 100,
        Born-Wolf PSF 127x127x63 (NA 1.40, ni 1.515, 460 nm, 65/200 nm),
        Poisson noise, 20 iterations
    c2  1024x1024x60, 2000 Gaussian blobs (sigma U(1,3) px, amplitude
        U(200,4000)), background 120, Born-Wolf PSF 181x181x79 (520 nm),
        Poisson noise, 50 iterations

Only numpy and scipy are used (no GPU), so a CPU-only process can make
the same inputs.
"""
import os

import numpy as np

from . import psf_bw, synth
from .method_shb import DwOpts, fim_normalize_sum1, psf_autocrop

CONFIGS = {
    "c1": dict(shape=(32, 128, 128), kind="points", n=40, amp=(500.0, 5000.0), bg=100.0,
               psf=(127, 127, 63), lambda_nm=460.0, dz=200.0, iters=20, seed=1),
    "c2": dict(shape=(60, 1024, 1024), kind="blobs", n=2000, sigma=(1.0, 3.0), amp=(200.0, 4000.0),
               bg=120.0, psf=(181, 181, 79), lambda_nm=520.0, dz=200.0, iters=50, seed=2),
}


def describe(name):
    c = CONFIGS[name]
    P, N, M = c["shape"]
    pM, pN, pP = c["psf"]
    obj = ("%d point sources" % c["n"]) if c["kind"] == "points" else ("%d Gaussian blobs" % c["n"])
    return ("%dx%dx%d float32, %s, background %g, Born-Wolf PSF %dx%dx%d (NA 1.40, ni 1.515, %g nm, "
            "dx 65 nm, dz %g nm), Poisson noise, %d SHB iterations"
            % (M, N, P, obj, c["bg"], pM, pN, pP, c["lambda_nm"], c["dz"], c["iters"]))


def _blur_scipy(V, psf):
    """'same' linear convolution in float32 through scipy.fft (threads)."""
    from scipy import fft as sfft
    P, N, M = V.shape
    pP, pN, pM = psf.shape
    full = (P + pP - 1, N + pN - 1, M + pM - 1)
    w = os.cpu_count() or 1
    a = sfft.rfftn(V.astype(np.float32), full, workers=w)
    a *= sfft.rfftn(psf.astype(np.float32), full, workers=w)
    out = sfft.irfftn(a, full, workers=w)
    cz, cy, cx = (pP - 1) // 2, (pN - 1) // 2, (pM - 1) // 2
    return out[cz:cz + P, cy:cy + N, cx:cx + M]


def make(name):
    """Returns (im, psf, opts, raw_psf_shape): im with the offset added,
    psf normalised and autocropped, opts a DwOpts with the iteration count."""
    c = CONFIGS[name]
    rng = np.random.default_rng(c["seed"])
    pM, pN, pP = c["psf"]
    psf = psf_bw.bw_psf(pM, pN, pP, c["lambda_nm"], NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=c["dz"])
    if c["kind"] == "points":
        V = synth.point_sources(c["shape"], c["n"], amp=c["amp"], rng=rng)
    else:
        V = synth.gaussian_blobs(c["shape"], c["n"], sigma=c["sigma"], amp=c["amp"], rng=rng)
    B = _blur_scipy(V, psf)
    del V
    im = synth.noisy_image(B, c["bg"], rng=rng)
    del B
    s = DwOpts(nIter=c["iters"], iter_type=2)
    ps = fim_normalize_sum1(psf)
    ps = psf_autocrop(ps, im.shape, s)
    ps = fim_normalize_sum1(ps)
    im += np.float32(s.offset)
    return np.ascontiguousarray(im), np.ascontiguousarray(ps), s, psf.shape


# ---------------------------------------------------------------------------
# c4: mixed-size request stream
# ---------------------------------------------------------------------------
C4 = dict(xy=(256, 384, 512, 768, 1024), z=(20, 40, 60, 80), lambdas=(460.0, 520.0, 600.0, 690.0),
          iters=(20, 50, 100), psf=(181, 181, 79), bg=120.0, blob_density=2000.0 / (1024 * 1024 * 60))


def describe_stream(n, rate, xy=None):
    xy = xy or C4["xy"]
    return ("c4 stream: %d jobs, Poisson arrivals at %g jobs/s, XY uniform over {%s}^2, "
            "Z uniform over {20,40,60,80}, Gaussian blobs at the c2 density, background 120, Born-Wolf PSF "
            "181x181x79 at 460/520/600/690 nm (NA 1.40), iterations uniform over {20,50,100}"
            % (n, rate, ",".join(str(v) for v in xy)))


def c4_psf(lam, device="cuda:0"):
    """Born-Wolf PSF of the c4 stream at wavelength lam (nm), from the device
    generator, normalised like bw_psf."""
    pM, pN, pP = C4["psf"]
    raw = psf_bw.bw_psf_device(pN, pP, lam, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=200.0).cpu().numpy()
    return fim_normalize_sum1(raw)


def make_c4_job(M, P, lam, iters, rng, gen_seed, device="cuda:0", psf=None):
    """One c4 job of XY size M x M and depth P: Gaussian blobs at the c2
    density drawn from rng, blurred on the GPU (torch.fft, input generation
    only), background 120, Poisson noise seeded with gen_seed; the PSF
    prepared like dw_run (normalise, autocrop, normalise), the offset
    added to the image."""
    import torch
    N = M
    pM, pN, pP = C4["psf"]
    psf = c4_psf(lam, device) if psf is None else psf
    nb = max(1, int(round(C4["blob_density"] * M * N * P)))
    V = synth.gaussian_blobs((P, N, M), nb, sigma=(1.0, 3.0), amp=(200.0, 4000.0), rng=rng)
    full = (P + pP - 1, N + pN - 1, M + pM - 1)
    a = torch.as_tensor(V.astype(np.float32), device=device)
    b = torch.as_tensor(psf, device=device)
    B = torch.fft.irfftn(torch.fft.rfftn(a, full) * torch.fft.rfftn(b, full), full)
    cz, cy, cx = (pP - 1) // 2, (pN - 1) // 2, (pM - 1) // 2
    B = B[cz:cz + P, cy:cy + N, cx:cx + M].clamp_min(0) + C4["bg"]
    g = torch.Generator(device=device)
    g.manual_seed(int(gen_seed))
    im = torch.poisson(B, generator=g).float().contiguous()
    del a, B
    s = DwOpts(nIter=iters, iter_type=2)
    ps = fim_normalize_sum1(psf)
    ps = psf_autocrop(ps, (P, N, M), s)
    ps = fim_normalize_sum1(ps)
    im += float(s.offset)
    return dict(im=im, psf=torch.as_tensor(ps, device=device).contiguous(), opts=s, shape=(M, N, P), iters=iters,
                lam=lam)


def make_stream(n, rate, seed=4, device="cuda:0", xy=None, reuse=False):
    """n jobs of the c4 stream: list of dicts with im, psf (torch tensors on
    device, prepared like dw_run), opts, arrival (s).  The PSFs come from
    the device Born-Wolf generator; blurring runs on the GPU through
    torch.fft (input generation only).  reuse: one generated image per
    (XY, Z, wavelength) class, shared by the jobs of that class (long
    streams; the jobs' shapes, iteration counts and arrivals are drawn as
    without it)."""
    rng = np.random.default_rng(seed)
    xy = xy or C4["xy"]
    psfs = {}
    made = {}
    jobs = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        M = int(rng.choice(xy))
        P = int(rng.choice(C4["z"]))
        lam = float(rng.choice(C4["lambdas"]))
        iters = int(rng.choice(C4["iters"]))
        if lam not in psfs:
            psfs[lam] = c4_psf(lam, device)
        key = (M, P, lam)
        if reuse and key in made:
            j = dict(made[key])
            j["opts"] = DwOpts(nIter=iters, iter_type=2)
            j["iters"] = iters
        else:
            j = make_c4_job(M, P, lam, iters, rng, seed * 100003 + i, device=device, psf=psfs[lam])
            made[key] = j
        j["arrival"] = t
        jobs.append(j)
    return jobs


# ---------------------------------------------------------------------------
# c3: one large image in overlapping tiles
# ---------------------------------------------------------------------------
C3 = dict(shape=(60, 2048, 2048), n=8000, sigma=(1.0, 3.0), amp=(200.0, 4000.0), bg=120.0, psf=(181, 181, 79),
          lambda_nm=520.0, iters=50, tile=512, overlap=48, seed=3)


def describe_c3():
    c = C3
    P, N, M = c["shape"]
    return ("c3: %dx%dx%d float32, %d Gaussian blobs, background %g, Born-Wolf PSF 181x181x79 (%g nm), Poisson "
            "noise, %d SHB iterations, tiles of %d px with %d px overlap (tiling.c)"
            % (M, N, P, c["n"], c["bg"], c["lambda_nm"], c["iters"], c["tile"], c["overlap"]))


def make_c3(device="cuda:0"):
    """(im GPU tensor without offset, psf numpy normalised and autocropped for
    the whole image, DwOpts).  PSF from the device generator, blur on the
    GPU with torch.fft (input generation only)."""
    import torch
    c = C3
    rng = np.random.default_rng(c["seed"])
    pM, pN, pP = c["psf"]
    raw = psf_bw.bw_psf_device(pN, pP, c["lambda_nm"], NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=200.0).cpu().numpy()
    psf = fim_normalize_sum1(raw)
    V = synth.gaussian_blobs(c["shape"], c["n"], sigma=c["sigma"], amp=c["amp"], rng=rng)
    P, N, M = c["shape"]
    full = (P + pP - 1, N + pN - 1, M + pM - 1)
    a = torch.as_tensor(V.astype(np.float32), device=device)
    del V
    b = torch.as_tensor(psf, device=device)
    B = torch.fft.irfftn(torch.fft.rfftn(a, full) * torch.fft.rfftn(b, full), full)
    del a
    cz, cy, cx = (pP - 1) // 2, (pN - 1) // 2, (pM - 1) // 2
    B = B[cz:cz + P, cy:cy + N, cx:cx + M].clamp_min(0) + c["bg"]
    g = torch.Generator(device=device)
    g.manual_seed(c["seed"])
    im = torch.poisson(B, generator=g).float().contiguous()
    del B
    s = DwOpts(nIter=c["iters"], iter_type=2)
    ps = psf_autocrop(psf, (P, N, M), s)
    ps = fim_normalize_sum1(ps)
    return im, ps, s


# ---------------------------------------------------------------------------
# c5: multi-channel fields
# ---------------------------------------------------------------------------
C5 = dict(shape=(81, 1024, 1024), lambdas=(460.0, 520.0, 600.0, 690.0), psf=(181, 181, 81), dz=300.0,
          density=(500, 5000), bg=(80.0, 200.0), sigma=(1.0, 3.0), amp=(200.0, 4000.0), iters=100, fields=64)


def describe_c5(fields):
    return ("c5: %d of %d fields x 4 channels (460/520/600/690 nm), each 1024x1024x81 float32, Gaussian blobs "
            "with a per-channel count U(500,5000) and background U(80,200), Born-Wolf PSF 181x181x81 (NA 1.40, "
            "dx 65 nm, dz 300 nm), Poisson noise, 100 SHB iterations" % (fields, C5["fields"]))


def make_c5(field, channels=(0, 1, 2, 3), device="cuda:0", iters=None):
    """Channels of field `field` of c5: list of dicts with im (torch, offset
    added), psf (torch, prepared like dw_run), opts, lam.  PSFs from the
    device generator, blur with torch.fft on the GPU (input generation
    only); seeds follow (field, channel)."""
    import torch
    c = C5
    P, N, M = c["shape"]
    pM, pN, pP = c["psf"]
    out = []
    for ch in channels:
        rng = np.random.default_rng(5000 + 4 * field + ch)
        lam = c["lambdas"][ch]
        raw = psf_bw.bw_psf_device(pN, pP, lam, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=c["dz"]).cpu().numpy()
        psf = fim_normalize_sum1(raw)
        nb = int(rng.integers(c["density"][0], c["density"][1] + 1))
        bg = float(rng.uniform(*c["bg"]))
        V = synth.gaussian_blobs(c["shape"], nb, sigma=c["sigma"], amp=c["amp"], rng=rng)
        full = (P + pP - 1, N + pN - 1, M + pM - 1)
        a = torch.as_tensor(V.astype(np.float32), device=device)
        del V
        b = torch.as_tensor(psf, device=device)
        B = torch.fft.irfftn(torch.fft.rfftn(a, full) * torch.fft.rfftn(b, full), full)
        del a
        cz, cy, cx = (pP - 1) // 2, (pN - 1) // 2, (pM - 1) // 2
        B = B[cz:cz + P, cy:cy + N, cx:cx + M].clamp_min(0) + bg
        g = torch.Generator(device=device)
        g.manual_seed(5000 + 4 * field + ch)
        im = torch.poisson(B, generator=g).float().contiguous()
        del B
        s = DwOpts(nIter=iters or c["iters"], iter_type=2)
        ps = psf_autocrop(psf, (P, N, M), s)
        ps = fim_normalize_sum1(ps)
        im += float(s.offset)
        out.append(dict(im=im, psf=torch.as_tensor(ps, device=device).contiguous(), opts=s, lam=lam, bg=bg,
                        blobs=nb))
    return out
