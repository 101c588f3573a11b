"""ctypes front-end of the CPU SHB oracle (liboracle_shb.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py, always as the checker.  The product
(deconwolf_amd) never imports this module.

The oracle restates deconwolf's deconvolve_shb (src/method_shb.c:77-424);
see shb_core.inc for the per-step citations.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_shb.so")
_lib = None

ITER_ABS, ITER_REL, ITER_FIXED = 0, 1, 2
METRIC_MSE, METRIC_IDIV = 0, 1
START_FLAT, START_LP, START_IDENTITY = 0, 1, 2


class ShbOpts(ctypes.Structure):
    """Mirror of oracle_shb_opts (shb_oracle.h)."""

    _fields_ = [
        ("niter", ctypes.c_int),
        ("iter_type", ctypes.c_int),
        ("maxiter", ctypes.c_int),
        ("err_rel", ctypes.c_float),
        ("err_abs", ctypes.c_float),
        ("bg", ctypes.c_float),
        ("bg_auto", ctypes.c_int),
        ("border_quality", ctypes.c_int),
        ("positivity", ctypes.c_int),
        ("alphamax", ctypes.c_float),
        ("metric", ctypes.c_int),
        ("start", ctypes.c_int),
        ("start_image", ctypes.c_void_p),
        ("work", ctypes.c_int64 * 3),
        ("nthreads", ctypes.c_int),
    ]


def build():
    """Compile liboracle_shb.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.oracle_cfft.argtypes = [dp, ctypes.c_int, ctypes.c_int]
        L.oracle_rfft3.argtypes = [dp, i64, i64, i64, dp]
        L.oracle_irfft3.argtypes = [dp, i64, i64, i64, dp]
        for name in ("oracle_shb_f32", "oracle_shb_f64", "oracle_rl_f32", "oracle_rl_f64"):
            f = getattr(L, name)
            f.argtypes = [dp, i64, i64, i64, dp, i64, i64, i64,
                          ctypes.POINTER(ShbOpts), dp, dp, dp]
            f.restype = ctypes.c_int
        for name in ("oracle_shb_setup_f32", "oracle_shb_setup_f64"):
            f = getattr(L, name)
            f.argtypes = [dp, i64, i64, i64, dp, i64, i64, i64,
                          ctypes.POINTER(ShbOpts), dp, dp, dp, dp]
            f.restype = ctypes.c_int
        L.oracle_num_threads.restype = ctypes.c_int
        L.oracle_set_fft_backend.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_set_fft_backend.restype = ctypes.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def set_fft_backend(external, nthreads=0, double=False):
    """external=True: float transforms through MKL's FFTW3 interface when
    libmkl_rt.so loads (needs MKL_THREADING_LAYER=GNU in the environment
    before the process starts); double=True: the float64 oracle's too.
    Returns whether what was asked for is in use."""
    level = (2 if double else 1) if external else 0
    got = lib().oracle_set_fft_backend(level, int(nthreads))
    return got == level if external else True


def cfft(x, sign=-1):
    """1-D complex FFT (unnormalised), sign -1 forward / +1 inverse."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().oracle_cfft(_ptr(a), a.size, sign)
    return a


def rfft3(vol):
    """vol: numpy array shaped (nz, ny, nx) (x fastest).  Returns
    (nz, ny, nx//2+1) complex128, FFTW r2c layout."""
    v = np.ascontiguousarray(vol, dtype=np.float64)
    nz, ny, nx = v.shape
    out = np.empty((nz, ny, nx // 2 + 1), dtype=np.complex128)
    lib().oracle_rfft3(_ptr(v), nx, ny, nz, _ptr(out))
    return out


def irfft3(spec, nx):
    """Unnormalised C2R of an (nz, ny, nx//2+1) spectrum."""
    s = np.ascontiguousarray(spec, dtype=np.complex128)
    nz, ny, _ = s.shape
    out = np.empty((nz, ny, nx), dtype=np.float64)
    lib().oracle_irfft3(_ptr(s), nx, ny, nz, _ptr(out))
    return out


def make_opts(niter=10, iter_type=ITER_FIXED, maxiter=250, err_rel=0.02, err_abs=1.0,
              bg=1e-2, bg_auto=1, border_quality=2, positivity=1, alphamax=1.0,
              metric=METRIC_IDIV, start=START_FLAT, work=None, nthreads=0):
    """Defaults are those of dw_opts_new (src/dw.c:340-413)."""
    o = ShbOpts()
    o.niter, o.iter_type, o.maxiter = niter, iter_type, maxiter
    o.err_rel, o.err_abs = err_rel, err_abs
    o.bg, o.bg_auto, o.border_quality = bg, bg_auto, border_quality
    o.positivity, o.alphamax, o.metric, o.start = positivity, alphamax, metric, start
    o.start_image = None
    w = work or (0, 0, 0)
    for i in range(3):
        o.work[i] = int(w[i])
    o.nthreads = nthreads
    return o


def shb(im, psf, opts=None, dtype=np.float32, start_image=None, **kw):
    """Run the oracle.  im: (P, N, M) array, psf: (pP, pN, pM) array, both
    x fastest.  Returns (out, err_trace, bg, iterations_run)."""
    return _run("shb", im, psf, opts, dtype, start_image, kw)


def rl(im, psf, opts=None, dtype=np.float32, start_image=None, **kw):
    """Richardson-Lucy (method_rl.c) on the same options as shb(); same
    return value."""
    return _run("rl", im, psf, opts, dtype, start_image, kw)


def _run(method, im, psf, opts, dtype, start_image, kw):
    o = opts if opts is not None else make_opts(**kw)
    im_a = np.ascontiguousarray(im, dtype=dtype)
    psf_a = np.ascontiguousarray(psf, dtype=dtype)
    keep = None
    if start_image is not None:
        keep = np.ascontiguousarray(start_image, dtype=dtype)
        o.start_image = keep.ctypes.data
    P, N, M = im_a.shape
    pP, pN, pM = psf_a.shape
    out = np.empty_like(im_a)
    nmax = max(o.niter, o.maxiter, 1)
    err = np.zeros(nmax + 1, dtype=np.float32)
    stats = np.zeros(4, dtype=np.float64)
    f = getattr(lib(), "oracle_%s_%s" % (method, "f32" if dtype == np.float32 else "f64"))
    rc = f(_ptr(im_a), M, N, P, _ptr(psf_a), pM, pN, pP, ctypes.byref(o), _ptr(out),
           _ptr(err), _ptr(stats))
    del keep
    if rc != 0:
        raise RuntimeError("oracle_%s failed with code %d" % (method, rc))
    n = int(stats[1])
    global last_times
    last_times = (float(stats[2]), float(stats[3]))
    return out, err[:n].copy(), float(stats[0]), n


# (setup seconds, iteration seconds) of the last shb() call
last_times = (0.0, 0.0)


def work_size(bq, imshape, psfshape):
    """Reference work size (method_shb.c:137-153) as (wM, wN, wP)."""
    P, N, M = imshape
    pP, pN, pM = psfshape
    if bq == 2:
        return (M + pM - 1, N + pN - 1, P + pP - 1)
    if bq == 1:
        return (M + (pM + 1) // 2, N + (pN + 1) // 2, P + (pP + 1) // 2)
    return (max(M, pM), max(N, pN), max(P, pP))


def shb_setup(im, psf, opts=None, dtype=np.float32, **kw):
    """Setup products at transform size: (otf complex (wP,wN,wM/2+1),
    W (wP,wN,wM), x0 (wP,wN,wM), bg, has_W)."""
    o = opts if opts is not None else make_opts(**kw)
    im_a = np.ascontiguousarray(im, dtype=dtype)
    psf_a = np.ascontiguousarray(psf, dtype=dtype)
    P, N, M = im_a.shape
    pP, pN, pM = psf_a.shape
    wref = work_size(o.border_quality, im_a.shape, psf_a.shape)
    w = [o.work[i] if o.work[i] > 0 else wref[i] for i in range(3)]
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    cK = np.empty((w[2], w[1], w[0] // 2 + 1), dtype=cdt)
    W = np.empty((w[2], w[1], w[0]), dtype=dtype)
    x0 = np.empty_like(W)
    stats = np.zeros(2, dtype=np.float64)
    f = lib().oracle_shb_setup_f32 if dtype == np.float32 else lib().oracle_shb_setup_f64
    rc = f(_ptr(im_a), M, N, P, _ptr(psf_a), pM, pN, pP, ctypes.byref(o), _ptr(cK), _ptr(W),
           _ptr(x0), _ptr(stats))
    if rc != 0:
        raise RuntimeError("oracle_shb_setup failed with code %d" % rc)
    return cK, W, x0, float(stats[0]), bool(stats[1])
