"""CPU-side checks of libdwhip and the host code around it (no GPU needed):
the C-ABI library loads and exports every function include/dwhip.h
declares; the work-size rules follow deconwolf's (method_shb.c:137-153);
defaults follow dw_opts_new (dw.c:340-413); the dw_opts adapter maps
fields one to one; without a GPU the entry points fail with
DWHIP_ERR_NODEVICE instead of running anything on the host."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import deconwolf_amd as dw
from deconwolf_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "dwhip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(dwhip_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_function():
    L = _lib.lib()
    names = header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding covers the same set
    bound = {s[0] for s in _lib.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def _ws(M, N, P, pM, pN, pP, bq):
    w = (ctypes.c_int64 * 3)()
    r = (ctypes.c_int64 * 3)()
    rc = _lib.lib().dwhip_work_size(M, N, P, pM, pN, pP, bq, w, r)
    return rc, tuple(w), tuple(r)


def smooth(n):
    for f in (2, 3, 5, 7):
        while n % f == 0:
            n //= f
    return n == 1


@pytest.mark.parametrize("shape", [(128, 128, 32, 127, 127, 63), (1024, 1024, 60, 181, 181, 79),
                                   (1024, 1024, 81, 181, 181, 81), (560, 608, 60, 181, 181, 79),
                                   (40, 36, 16, 9, 9, 7), (1, 1, 1, 1, 1, 1), (101, 201, 40, 181, 181, 79)])
def test_work_size_bq2(shape):
    M, N, P, pM, pN, pP = shape
    rc, w, r = _ws(M, N, P, pM, pN, pP, 2)
    assert r == (M + pM - 1, N + pN - 1, P + pP - 1)
    assert rc == 0
    assert w[0] % 2 == 0 and smooth(w[0] // 2) and smooth(w[1]) and smooth(w[2])
    assert all(a >= b for a, b in zip(w, r))
    # padding stays small
    assert all(a <= 1.1 * b + 8 for a, b in zip(w, r))


def test_work_size_bq1_bq0_exact():
    """bq 0 and 1 run at deconwolf's exact work size whatever its factors:
    odd x extents and prime factors go to the any-size path (Bluestein);
    only extents whose transform does not fit a workgroup's LDS (Bluestein
    length m >= 2n - 1 with 16 m bytes above 150 KiB) are refused."""
    rc, w, r = _ws(43, 26, 11, 9, 7, 7, 1)
    assert r == (43 + 5, 26 + 4, 11 + 4) and w == r and rc == 0
    rc, w, r = _ws(40, 27, 12, 5, 5, 5, 0)
    assert r == (40, 27, 12) and w == r and rc == 0
    rc, w, r = _ws(41, 27, 12, 5, 5, 5, 0)   # odd x extent
    assert r == (41, 27, 12) and w == r and rc == 0
    rc, w, r = _ws(40, 11, 12, 5, 5, 5, 0)   # 11 is prime
    assert w == r == (40, 11, 12) and rc == 0
    rc, w, r = _ws(1024, 1024, 60, 181, 181, 79, 1)   # 1115 = 5 * 223
    assert w == r == (1115, 1115, 100) and rc == 0
    rc, w, r = _ws(9973, 16, 8, 3, 3, 3, 0)   # prime beyond the LDS limit
    assert rc == -2
    assert _ws(0, 1, 1, 1, 1, 1, 2)[0] == -1
    assert _ws(1, 1, 1, 1, 1, 1, 3)[0] == -1


def test_params_default_match_dw_opts_new():
    p = _lib.DwhipParams()
    _lib.lib().dwhip_params_default(ctypes.byref(p), 17)
    assert p.niter == 17 and p.iter_type == _lib.ITER_FIXED and p.maxiter == 250
    assert abs(p.err_rel - 0.02) < 1e-9 and p.err_abs == 1.0
    assert abs(p.bg - 1e-2) < 1e-9 and p.bg_auto == 1 and p.border_quality == 2
    assert p.positivity == 1 and p.alphamax == 1.0 and p.metric == _lib.METRIC_IDIV
    assert p.start == _lib.START_FLAT
    s = dw.DwOpts()
    assert (s.maxiter, s.borderQuality, s.offset, s.metric) == (250, 2, 5.0, _lib.METRIC_IDIV)


def test_strerror_and_job_bytes():
    L = _lib.lib()
    for c in (0, -1, -2, -3, -4, -5, -6, -7, -8):
        assert len(L.dwhip_strerror(c)) > 0
    w = (ctypes.c_int64 * 3)(1250, 1215, 140)
    b = L.dwhip_job_bytes(w, 1024, 1024, 60, 181, 181, 79, 2, 50)
    V = 1250 * 1215 * 140
    assert 4.9 * V * 4 < b < 5.2 * V * 4   # C, K (V/2+ complex each) + U0, U1, W (V floats each)


def test_no_gpu_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    im = np.ones((4, 6, 8), np.float32) * 10
    k = np.zeros((3, 3, 3), np.float32)
    k[1, 1, 1] = 1
    with pytest.raises(dw.DwhipError) as e:
        dw.deconvolve_shb_hip(im, k, dw.DwOpts(nIter=3, iter_type=dw.ITER_FIXED))
    assert "NODEVICE" in str(e.value)
    # nIter == 0 returns a copy without touching a device (method_shb.c:101-105)
    out = dw.deconvolve_shb_hip(im, k, dw.DwOpts(nIter=0, iter_type=dw.ITER_FIXED))
    assert np.array_equal(out, im)


def test_engine_config_validated_before_device_use():
    """dwhip_engine_create checks its configuration before touching a
    device: bad values give DWHIP_ERR_ARG on any host; the defaults are
    FIFO admission, no voxel cap, one job group, 64 active jobs, up to 8
    rounds per step, no aging."""
    L = _lib.lib()
    cfg = _lib.EngineConfig()
    L.dwhip_engine_config_default(ctypes.byref(cfg))
    assert (cfg.max_active, cfg.order, cfg.active_mvox, cfg.groups) == (64, _lib.ORDER_FIFO, 0, 1)
    assert (cfg.rounds_per_step, cfg.aging_ms) == (8, 0)
    assert ctypes.sizeof(_lib.EngineConfig) == 40   # hbm_budget + 7 ints, 8-byte aligned, as in dwhip.h
    h = ctypes.c_void_p()
    for field, bad in (("groups", 3), ("groups", -1), ("order", 7), ("active_mvox", -5), ("max_active", 0),
                       ("rounds_per_step", -1), ("aging_ms", -3)):
        c = _lib.EngineConfig()
        L.dwhip_engine_config_default(ctypes.byref(c))
        setattr(c, field, bad)
        assert _lib.ERRORS.get(L.dwhip_engine_create(ctypes.byref(c), None, ctypes.byref(h))) == "DWHIP_ERR_ARG", field


class View(ctypes.Structure):
    _fields_ = [(n, t) for n, t in [
        ("nIter", ctypes.c_int), ("iter_type", ctypes.c_int), ("maxiter", ctypes.c_int),
        ("err_rel", ctypes.c_float), ("err_abs", ctypes.c_float), ("bg", ctypes.c_float),
        ("bg_auto", ctypes.c_int), ("borderQuality", ctypes.c_int), ("positivity", ctypes.c_int),
        ("alphamax", ctypes.c_float), ("metric", ctypes.c_int), ("start_condition", ctypes.c_int),
        ("verbosity", ctypes.c_int), ("fftw3_planning", ctypes.c_int)]]


@pytest.fixture(scope="module")
def adapter():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "deconwolf_amd", "adapter")], check=True)
    L = ctypes.CDLL(os.path.join(ROOT, "deconwolf_amd", "adapter", "libdwhip_adapter.so"))
    L.dwhip_adapter_params.argtypes = [ctypes.POINTER(View), ctypes.POINTER(_lib.DwhipParams)]
    L.dwhip_adapter_trace_len.argtypes = [ctypes.POINTER(View)]
    L.dwhip_adapter_log_iterations.argtypes = [ctypes.c_void_p, ctypes.POINTER(View), ctypes.c_void_p,
                                               ctypes.c_int]
    return L


def test_adapter_maps_every_field(adapter):
    v = View(33, 1, 77, 0.05, 2.5, 12.0, 0, 1, 0, 0.5, 0, 2, 0, _lib.FFTW_ESTIMATE)
    p = _lib.DwhipParams()
    adapter.dwhip_adapter_params(ctypes.byref(v), ctypes.byref(p))
    assert p.planning == _lib.PLAN_RULE        # --noplan (dw.c:1068)
    v.fftw3_planning = _lib.FFTW_MEASURE       # dw_opts_new's default (dw.c:400)
    adapter.dwhip_adapter_params(ctypes.byref(v), ctypes.byref(p))
    assert p.planning == _lib.PLAN_MEASURE
    assert (p.niter, p.iter_type, p.maxiter) == (33, 1, 77)
    assert abs(p.err_rel - 0.05) < 1e-7 and p.err_abs == 2.5 and p.bg == 12.0
    assert (p.bg_auto, p.border_quality, p.positivity) == (0, 1, 0)
    assert p.alphamax == 0.5 and p.metric == 0 and p.start == 2
    assert tuple(p.work) == (0, 0, 0) and p.start_image is None
    assert adapter.dwhip_adapter_trace_len(ctypes.byref(v)) == 77
    v.iter_type = 2
    assert adapter.dwhip_adapter_trace_len(ctypes.byref(v)) == 33


def test_adapter_log_lines(adapter, tmp_path):
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    path = str(tmp_path / "log.txt")
    f = libc.fopen(path.encode(), b"w")
    v = View(3, 2, 250, 0.02, 1.0, 0.01, 1, 2, 1, 1.0, 1, 0, 0)
    err = np.array([1.5, 0.25, 0.125], np.float32)
    adapter.dwhip_adapter_log_iterations(f, ctypes.byref(v), err.ctypes.data, 3)
    libc.fclose(f)
    lines = open(path).read().splitlines()
    # dw_iterator_show's log format (src/dw.c:327-332)
    assert lines == ["Iteration   1/  3, Idiv=1.500000e+00", "Iteration   2/  3, Idiv=2.500000e-01",
                     "Iteration   3/  3, Idiv=1.250000e-01"]


def test_plan_cache_file_and_rule_plan(tmp_path):
    """dwhip_plan_get without a measured plan gives the rule planner's:
    fewest stages, then the smallest largest radix (c2's transform:
    625 = 25*25, 1215 = 15*9*9, 140 = 7*5*4) and the largest tiles; a plan
    line in the cache file (the format dwhip_plan_tune writes) replaces it,
    and lines whose radices are not a plan the tuner could choose (each
    radix multiset in one of its two monotone orders) are ignored."""
    w = (1250, 1215, 140)
    dw.plan_cache_path("")
    measured, rad, tiles = dw.plan_get(w)
    assert not measured and rad == [[25, 25], [15, 9, 9], [7, 5, 4]] and tiles == (10, 16, 32)
    f = tmp_path / "plans.txt"
    f.write_text("gfx950 1250 1215 140 x=5,5,25 xrows=5 y=9,9,15 ycols=8 z=4,5,7 zcols=16\n"
                 "gfx950 256 256 96 x=11,11 xrows=4 y=16,16 ycols=8 z=6,4,4 zcols=16\n"     # 11 not a radix
                 "gfx942 256 256 96 x=16,8 xrows=4 y=16,16 ycols=8 z=6,4,4 zcols=16\n")     # other device
    try:
        dw.plan_cache_path(str(f))
        measured, rad, tiles = dw.plan_get(w)
        assert measured and rad == [[5, 5, 25], [9, 9, 15], [4, 5, 7]] and tiles == (5, 8, 16)
        measured, rad, _ = dw.plan_get((256, 256, 96))
        assert not measured and rad[0] == [16, 8]
        f.write_text("gfx950 1250 1215 140 x=5,25,5 xrows=5 y=9,9,15 ycols=8 z=4,5,7 zcols=16\n")
        dw.plan_cache_path(str(f))
        assert not dw.plan_get(w)[0]
    finally:
        dw.plan_cache_path("")
