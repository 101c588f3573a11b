"""Python host side of the HIP SHB method.

Mirrors deconwolf's method interface for the scaled-heavy-ball loop:

* :class:`DwOpts` holds the `dw_opts` fields that `deconvolve_shb` reads
  (src/dw.h:104-211) with the defaults of `dw_opts_new` (src/dw.c:340-413);
* :func:`deconvolve_shb_hip` has the meaning of the `dw_function`
  `deconvolve_shb(im, M, N, P, psf, pM, pN, pP, s)` (src/dw.h:89-91,
  src/method_shb.c:77-424): same inputs (image with the offset already
  added, PSF normalised and autocropped), same output (the x_{n-1} iterate
  cropped to M x N x P), and it writes the resolved background back into
  ``s.bg`` as the reference does (method_shb.c:108-117);
* :func:`deconvolve` mirrors the in-memory `dw_run` path used by the Python
  wrapper (src/pywrapper/src/pydeconwolf.c:284-395, src/dw.c:2799-2852):
  normalise, autocrop, add the offset, run the method.

Volumes are numpy arrays shaped (P, N, M), x fastest, like pydeconwolf.
All arithmetic runs in libdwhip on the GPU; this module only marshals.
"""
import ctypes
from dataclasses import dataclass, field
from typing import Optional, Tuple

import numpy as np

from . import _lib
from ._lib import (DwhipParams, DwhipStats, check, lib, ITER_ABS, ITER_REL, ITER_FIXED,
                   METRIC_MSE, METRIC_IDIV, START_FLAT, START_LP, START_IDENTITY,
                   METHOD_RL, METHOD_ID, METHOD_SHB, DWHIP_METHOD_SHB, DWHIP_METHOD_RL,
                   FFTW_ESTIMATE, FFTW_MEASURE, PLAN_MEASURE, PLAN_RULE)


@dataclass
class DwOpts:
    """The dw_opts fields on the SHB and RL paths (src/dw.h:104-211)."""

    method: int = METHOD_SHB          # dw.c:376 (dw_method values, dw.h:52-60)
    nIter: int = 1                    # dw.c:349
    iter_type: int = ITER_REL         # dw.c:361
    maxiter: int = 250                # dw.c:350
    err_rel: float = 0.02             # dw.c:351
    err_abs: float = 1.0              # dw.c:352
    bg: float = 1e-2                  # dw.c:389
    bg_auto: int = 1                  # dw.c:390
    borderQuality: int = 2            # dw.c:382
    positivity: int = 1               # dw.c:388
    alphamax: float = 1.0             # dw.c:401
    metric: int = METRIC_IDIV         # dw.c:398
    start_condition: int = START_FLAT  # dw.c:410
    offset: float = 5.0               # dw.c:394
    xycropfactor: float = 0.001       # dw.c:377
    outFormat: int = 16               # dw.c:383 (16: uint16 with scaling, 32: float32)
    scaling: float = -1.0             # dw.c:384 (<= 0: automatic)
    verbosity: int = 0
    # FFTW planner flag (dw.c:400): FFTW_ESTIMATE plans the transforms by
    # rule, anything else measures a new transform size's plans once
    fftw3_planning: int = FFTW_MEASURE
    iterdump: int = 0                 # dw.c:378: the guess every iterdump iterations
    fulldump: int = 0                 # dw.c:386: the final guess and shifted PSF at transform size
    # per-iteration callable(iter, error, x) for iterdump/fulldump/tsv-like
    # uses: x the (P,N,M) guess x_k every `iterdump` iterations (else None);
    # with fulldump, iter -1 gives the result and -2 the shifted PSF at the
    # transform size (nz, ny, nx).  Set: every pass runs directly, the host
    # waits for each iteration.
    iter_hook: Optional[object] = field(default=None, repr=False)
    # transform size; None = automatic (smallest 2,3,5,7-smooth size >= the
    # reference work size for borderQuality 2)
    work: Optional[Tuple[int, int, int]] = None
    # filled by the method: error per iteration, stats
    err_trace: Optional[np.ndarray] = field(default=None, repr=False)
    stats: Optional[dict] = field(default=None, repr=False)


def _params(s: DwOpts) -> DwhipParams:
    p = DwhipParams()
    lib().dwhip_params_default(ctypes.byref(p), int(s.nIter))
    p.niter = int(s.nIter)
    p.iter_type = int(s.iter_type)
    p.maxiter = int(s.maxiter)
    p.err_rel = float(s.err_rel)
    p.err_abs = float(s.err_abs)
    p.bg = float(s.bg)
    p.bg_auto = int(s.bg_auto)
    p.border_quality = int(s.borderQuality)
    p.positivity = int(s.positivity)
    p.alphamax = float(s.alphamax)
    p.metric = int(s.metric)
    p.start = int(s.start_condition)
    p.start_image = None
    w = s.work or (0, 0, 0)
    for i in range(3):
        p.work[i] = int(w[i])
    p.verbosity = int(s.verbosity)
    # dwhip_params.method takes the dw_method values themselves
    if int(s.method) not in (DWHIP_METHOD_SHB, DWHIP_METHOD_RL):
        raise ValueError("method %r has no GPU path (SHB and RL do)" % (s.method,))
    p.method = int(s.method)
    p.planning = PLAN_RULE if int(s.fftw3_planning) == FFTW_ESTIMATE else PLAN_MEASURE
    if s.iter_hook is not None:
        user = s.iter_hook

        def _hook(_u, it, err, x, nx, ny, nz):
            arr = None
            if x:
                arr = np.ctypeslib.as_array(x, shape=(int(nz), int(ny), int(nx))).copy()
            user(int(it), float(err), arr)

        p.hook = _lib.HOOK(_hook)
        p.hook_every = int(s.iterdump)
        p.hook_full = 1 if s.fulldump else 0
        p._keep = p.hook  # the callback object lives as long as the params
    return p


def plan_tune(work, retune=False):
    """Measures the fused passes' plans for a transform size (as
    :func:`work_size` returns it) and caches the choice, like FFTW's
    planner with FFTW_MEASURE; returns (tuned_ms, rule_ms), the summed pass
    times of the chosen plan and of the rule planner's."""
    w = (ctypes.c_int64 * 3)(*[int(v) for v in work])
    tb, tr = ctypes.c_double(), ctypes.c_double()
    check(lib().dwhip_plan_tune(w, 1 if retune else 0, ctypes.byref(tb), ctypes.byref(tr)), "dwhip_plan_tune")
    return tb.value, tr.value


def plan_get(work):
    """(measured, radices, tiles) of the plan a job of transform size work
    uses: radices per axis (x rows of length nx/2, y, z), tiles (x rows, y
    and z columns per workgroup)."""
    w = (ctypes.c_int64 * 3)(*[int(v) for v in work])
    r = (ctypes.c_int * 12)()
    t = (ctypes.c_int * 3)()
    rc = lib().dwhip_plan_get(w, r, t)
    if rc < 0:
        check(rc, "dwhip_plan_get")
    rad = [[v for v in r[4 * a:4 * a + 4] if v] for a in range(3)]
    return bool(rc), rad, tuple(t)


def plan_cache_path(path):
    """Plan cache file ("" for none, None for the default); loads it."""
    check(lib().dwhip_plan_cache_path(None if path is None else path.encode()), "dwhip_plan_cache_path")


def _stats_dict(st: DwhipStats) -> dict:
    return {
        "bg": st.bg, "iterations": st.iterations,
        "work": tuple(st.work), "ref_work": tuple(st.ref_work),
        "setup_ms": st.setup_ms, "iter_ms": st.iter_ms,
    }


def work_size(imshape, psfshape, borderQuality=2):
    """(work, ref_work) as (wM, wN, wP) tuples for (P,N,M) / (pP,pN,pM) shapes."""
    P, N, M = imshape
    pP, pN, pM = psfshape
    w = (ctypes.c_int64 * 3)()
    r = (ctypes.c_int64 * 3)()
    rc = lib().dwhip_work_size(M, N, P, pM, pN, pP, borderQuality, w, r)
    return tuple(w), tuple(r), rc


def _finish(s, st, err):
    n = st.iterations
    s.err_trace = err[:n].copy()
    s.stats = _stats_dict(st)
    # the background is resolved at set-up (method_shb.c:108-117), before the
    # loop, so it is written back even when no iteration ran; with nIter 0
    # the method returns first and st.bg is the caller's own value
    if s.bg_auto:
        s.bg = float(st.bg)


def _lp_start(im, s):
    if s.start_condition == START_LP:
        from .filters import gsmooth
        return gsmooth(np.array(im, dtype=np.float32, copy=True), 8.0)
    return None


def deconvolve_shb_hip(im: np.ndarray, psf: np.ndarray, s: DwOpts) -> np.ndarray:
    """SHB deconvolution on the GPU, host arrays in and out.

    im: (P, N, M) float32 image, offset already added (borrowed).
    psf: (pP, pN, pM) float32 PSF, sum 1, autocropped.
    Returns a new (P, N, M) float32 array: x_{n-1} after n iterations
    (method_shb.c:385-413)."""
    return _run_host("dwhip_shb_run", im, psf, s)


def deconvolve_rl_hip(im: np.ndarray, psf: np.ndarray, s: DwOpts) -> np.ndarray:
    """Richardson-Lucy on the GPU with the meaning of deconvolve_rl
    (src/method_rl.c:105-400): same inputs as :func:`deconvolve_shb_hip`;
    returns the last iterate x_n cropped to (P, N, M)."""
    return _run_host("dwhip_rl_run", im, psf, s)


def _run_host(entry, im, psf, s):
    im = np.ascontiguousarray(im, dtype=np.float32)
    psf = np.ascontiguousarray(psf, dtype=np.float32)
    if im.ndim != 3 or psf.ndim != 3:
        raise ValueError("im and psf must be 3-D (P, N, M) arrays")
    P, N, M = im.shape
    pP, pN, pM = psf.shape
    p = _params(s)
    start = _lp_start(im, s)
    if start is not None:
        start = np.ascontiguousarray(start, dtype=np.float32)  # the library reads it x-fastest
        p.start_image = start.ctypes.data
    out = np.empty_like(im)
    nerr = max(int(s.nIter), int(s.maxiter), 1)
    err = np.zeros(nerr, dtype=np.float32)
    st = DwhipStats()
    st.err_trace = err.ctypes.data
    st.err_trace_len = nerr
    check(getattr(lib(), entry)(im.ctypes.data, M, N, P, psf.ctypes.data, pM, pN, pP,
                                ctypes.byref(p), out.ctypes.data, ctypes.byref(st)), entry)
    _finish(s, st, err)
    return out


def deconvolve_shb_hip_device(im, psf, s: DwOpts, out=None, start_image=None):
    """Same as :func:`deconvolve_shb_hip` for torch tensors already on the
    GPU (float32, contiguous, shapes (P,N,M) / (pP,pN,pM)).  Work is queued
    on torch's current stream.  Returns `out` (allocated if None)."""
    return _run_device("dwhip_shb_run_device", im, psf, s, out, start_image)


def deconvolve_shb_hip_device_noalloc(im, psf, s: DwOpts, out, start_image=None):
    """:func:`deconvolve_shb_hip_device` through the entry point that
    allocates no device memory (the job's state in the library's static
    block of 256 MiB per GPU; larger jobs raise DWHIP_ERR_SCRATCH).  `out`
    must be given.  Calls that overlap on one GPU race on that block."""
    if out is None:
        raise ValueError("the no-allocation entry point needs `out`")
    return _run_device("dwhip_shb_run_device_noalloc", im, psf, s, out, start_image)


def deconvolve_rl_hip_device(im, psf, s: DwOpts, out=None, start_image=None):
    """:func:`deconvolve_rl_hip` for GPU tensors, as
    :func:`deconvolve_shb_hip_device`."""
    return _run_device("dwhip_rl_run_device", im, psf, s, out, start_image)


def _run_device(entry, im, psf, s, out, start_image):
    import torch
    if not (im.is_cuda and psf.is_cuda):
        raise ValueError("im and psf must be GPU tensors")
    im = im.contiguous()
    psf = psf.contiguous()
    if im.dtype != torch.float32 or psf.dtype != torch.float32:
        raise ValueError("im and psf must be float32")
    P, N, M = im.shape
    pP, pN, pM = psf.shape
    if out is None:
        out = torch.empty_like(im)
    p = _params(s)
    if start_image is not None:
        start_image = start_image.contiguous()
        p.start_image = start_image.data_ptr()
    elif s.start_condition == START_LP:
        raise ValueError("LP start on the device path needs start_image (the smoothed image)")
    nerr = max(int(s.nIter), int(s.maxiter), 1)
    err = np.zeros(nerr, dtype=np.float32)
    st = DwhipStats()
    st.err_trace = err.ctypes.data
    st.err_trace_len = nerr
    stream = torch.cuda.current_stream().cuda_stream
    check(getattr(lib(), entry)(im.data_ptr(), M, N, P, psf.data_ptr(), pM, pN, pP,
                                ctypes.byref(p), out.data_ptr(), ctypes.byref(st), stream), entry)
    _finish(s, st, err)
    return out


def fft3_device(vol, inverse=False, spec=None):
    """Unnormalised 3-D R2C (or C2R with inverse=True) through the engine's
    transforms.  vol: float32 (nz, ny, nx) GPU tensor, nx even; spec:
    complex64 (nz, ny, nx//2+1) GPU tensor.  Returns spec (R2C) or vol (C2R)."""
    import torch
    nz, ny, nx = vol.shape
    if spec is None:
        spec = torch.empty((nz, ny, nx // 2 + 1), dtype=torch.complex64, device=vol.device)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().dwhip_fft3_device(vol.data_ptr(), spec.data_ptr(), nx, ny, nz, 1 if inverse else 0, stream),
          "dwhip_fft3_device")
    return vol if inverse else spec


# ---------------------------------------------------------------------------
# caller-side steps of dw_run (src/dw.c:2799-2852) for the in-memory API
# ---------------------------------------------------------------------------
def fim_normalize_sum1(psf):
    """psf / sum(psf), sum in double (src/fim.c:900-912)."""
    psf = np.asarray(psf, dtype=np.float32)
    s = float(np.sum(psf, dtype=np.float64))
    return (psf.astype(np.float64) / s).astype(np.float32)


def psf_autocrop(psf, imshape, s: DwOpts):
    """psf_autocrop (src/dw.c:1973-1997): crop to at most 2*dim-1 per axis
    (psf_autocrop_byImage, dw.c:1764-1855), then drop outer x/y planes
    whose plane sum is below xycropfactor * the largest plane sum
    (psf_autocrop_XY, dw.c:1857-1971).

    The crop index is integer output of float arithmetic, so it follows the
    reference's: each x plane's sum accumulates in float, y outer and z
    inner (dw.c:1873-1882), the threshold is the float product
    xycropfactor * maxsum, and the first plane reaching it ends the scan.
    Where no plane reaches it (xycropfactor > 1), deconwolf reads past the
    PSF; here that is a ValueError."""
    P, N, M = imshape
    psf = np.asarray(psf, dtype=np.float32)
    p, n, m = psf.shape
    mopt, nopt, popt = (M - 1) * 2 + 1, (N - 1) * 2 + 1, (P - 1) * 2 + 1
    if s.borderQuality == 0:
        mopt, nopt, popt = M, N, P
    if not p < popt and (m > mopt or n > nopt or p > popt):
        m0, m1, n0, n1, p0, p1 = 0, m - 1, 0, n - 1, 0, p - 1
        if m > mopt:
            m0, m1 = (m - mopt) // 2, m1 - (m - mopt) // 2
        if n > nopt:
            n0, n1 = (n - nopt) // 2, n1 - (n - nopt) // 2
        if p > popt:
            p0, p1 = (p - popt) // 2, p1 - (p - popt) // 2
        psf = np.ascontiguousarray(psf[p0:p1 + 1, n0:n1 + 1, m0:m1 + 1])
    if s.borderQuality > 0 and s.xycropfactor > 0:
        sums = xplane_sums_f32(psf)
        maxsum = np.float32(0)
        for v in sums:  # sum > maxsum ? maxsum = sum : 0 (dw.c:1883)
            if v > maxsum:
                maxsum = v
        thr = np.float32(s.xycropfactor) * maxsum
        first = -1
        ssum = np.float32(0)
        while ssum < thr:
            first += 1
            if first >= sums.size:
                raise ValueError("psf_autocrop: no x plane reaches xycropfactor * the largest plane sum")
            ssum = sums[first]
        if first >= 1:
            pM = psf.shape[2]
            pN = psf.shape[1]
            psf = np.ascontiguousarray(psf[:, first:pN - first, first:pM - first])
    return psf


def xplane_sums_f32(psf):
    """Sum of each x plane of a (pP, pN, pM) PSF accumulated in float32 in
    deconwolf's order, y outer and z inner (psf_autocrop_XY, dw.c:1873-1882):
    one sequential float32 accumulation per plane."""
    a = np.ascontiguousarray(np.asarray(psf, dtype=np.float32).transpose(2, 1, 0))  # [x][y][z]
    m = a.shape[0]
    if a.size == 0:
        return np.zeros(m, np.float32)
    return np.cumsum(a.reshape(m, -1), axis=1, dtype=np.float32)[:, -1].copy()


def deconvolve(image, psf, iterations, bq=2, verbose=0, opts: Optional[DwOpts] = None):
    """In-memory deconvolution like pydeconwolf.deconvolve
    (src/pywrapper/src/pydeconwolf.c:284-395): FIXED iteration count, PSF
    normalised and autocropped, offset added before the method; the result
    keeps the offset, as dw_run's non-tiled path returns it (dw.c:2845-2849)."""
    s = opts or DwOpts()
    s.nIter = int(iterations)
    s.iter_type = ITER_FIXED
    s.borderQuality = int(bq)
    s.verbosity = int(verbose)
    im = np.array(image, dtype=np.float32, copy=True)
    if im.min() < 0:
        raise ValueError("the image has negative values (dw.c:2703-2710)")
    if im.max() < 1:
        raise ValueError("the image maximum is below 1 (dw.c:2711-2721)")
    ps = fim_normalize_sum1(psf)
    ps = psf_autocrop(ps, im.shape, s)
    ps = fim_normalize_sum1(ps)
    if s.offset > 0:
        im += np.float32(s.offset)
    if int(s.method) == METHOD_RL:
        return deconvolve_rl_hip(im, ps, s)
    return deconvolve_shb_hip(im, ps, s)
