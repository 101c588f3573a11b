"""ctypes binding of libdwhip.so (include/dwhip.h).

The library is built in-tree by deconwolf_amd.build.  There is no fallback:
if the library is missing or cannot be loaded, every entry point raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdwhip.so")

DWHIP_OK = 0
ERRORS = {
    -1: "DWHIP_ERR_ARG", -2: "DWHIP_ERR_SIZE", -3: "DWHIP_ERR_HIP", -4: "DWHIP_ERR_NOMEM",
    -5: "DWHIP_ERR_PSF", -6: "DWHIP_ERR_NODEVICE", -7: "DWHIP_ERR_STATE", -8: "DWHIP_ERR_SCRATCH",
}
# size of the static device block of dwhip_shb_run_device_noalloc
SCRATCH_BYTES = 256 << 20

ITER_ABS, ITER_REL, ITER_FIXED = 0, 1, 2
METRIC_MSE, METRIC_IDIV = 0, 1
START_FLAT, START_LP, START_IDENTITY = 0, 1, 2
# dwhip_to_u16_device modes
U16_ROUND, U16_TRUNC, U16_RINT = 0, 1, 2
# dwhip_engine_config.order
ORDER_FIFO, ORDER_SHORTEST = 0, 1
# dwhip_params.method: the dw_method values (src/dw.h:52-60)
DWHIP_METHOD_RL, DWHIP_METHOD_SHB = 0, 2
# dw_method (src/dw.h:52-60), the values of DwOpts.method
METHOD_RL, METHOD_ID, METHOD_SHB = 0, 1, 2
# dwhip_params.planning
PLAN_RULE, PLAN_MEASURE = 0, 1
# FFTW planner flags, the values of DwOpts.fftw3_planning (src/dw.c:400,1068)
FFTW_MEASURE, FFTW_ESTIMATE = 0, 1 << 6


# dwhip_params.hook: void (*)(void *user, int iter, float error, const float *x, int64 nx, ny, nz)
HOOK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64)


class DwhipParams(ctypes.Structure):
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
        ("verbosity", ctypes.c_int),
        ("method", ctypes.c_int),
        ("planning", ctypes.c_int),
        ("hook", HOOK),
        ("hook_user", ctypes.c_void_p),
        ("hook_every", ctypes.c_int),
        ("hook_full", ctypes.c_int),
    ]


class DwhipStats(ctypes.Structure):
    _fields_ = [
        ("bg", ctypes.c_float),
        ("iterations", ctypes.c_int),
        ("work", ctypes.c_int64 * 3),
        ("ref_work", ctypes.c_int64 * 3),
        ("setup_ms", ctypes.c_double),
        ("iter_ms", ctypes.c_double),
        ("err_trace", ctypes.c_void_p),
        ("err_trace_len", ctypes.c_int),
    ]


class EngineConfig(ctypes.Structure):
    _fields_ = [
        ("hbm_budget", ctypes.c_size_t),
        ("max_active", ctypes.c_int),
        ("order", ctypes.c_int),
        ("active_mvox", ctypes.c_int),
        ("groups", ctypes.c_int),
        ("rounds_per_step", ctypes.c_int),
        ("aging_ms", ctypes.c_int),
        ("reserved", ctypes.c_int * 1),
    ]


class DwhipTile(ctypes.Structure):
    _fields_ = [("pos", ctypes.c_int64 * 6), ("xpos", ctypes.c_int64 * 6), ("xsize", ctypes.c_int64 * 3)]


class DwhipError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = ERRORS.get(code, str(code))
        try:
            msg += ": " + lib().dwhip_strerror(code).decode()
        except Exception:  # pragma: no cover - message decoration only
            pass
        super().__init__((what + ": " if what else "") + msg)


# (name, restype, argtypes) of every symbol include/dwhip.h declares
_vp, _i64, _int = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
_P64 = ctypes.POINTER(ctypes.c_int64)
SIGNATURES = [
    ("dwhip_params_default", None, [ctypes.POINTER(DwhipParams), _int]),
    ("dwhip_strerror", ctypes.c_char_p, [_int]),
    ("dwhip_init", _int, []),
    ("dwhip_work_size", _int, [_i64, _i64, _i64, _i64, _i64, _i64, _int, _P64, _P64]),
    ("dwhip_job_bytes", ctypes.c_size_t, [_P64, _i64, _i64, _i64, _i64, _i64, _i64, _int, _int]),
    ("dwhip_plan_tune", _int, [_P64, _int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    ("dwhip_plan_cache_path", _int, [ctypes.c_char_p]),
    ("dwhip_plan_get", _int, [_P64, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    ("dwhip_shb_run", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                             ctypes.POINTER(DwhipParams), _vp, ctypes.POINTER(DwhipStats)]),
    ("dwhip_shb_run_device", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                                    ctypes.POINTER(DwhipParams), _vp, ctypes.POINTER(DwhipStats), _vp]),
    ("dwhip_shb_run_device_noalloc", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                                            ctypes.POINTER(DwhipParams), _vp, ctypes.POINTER(DwhipStats), _vp]),
    ("dwhip_rl_run", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                            ctypes.POINTER(DwhipParams), _vp, ctypes.POINTER(DwhipStats)]),
    ("dwhip_rl_run_device", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                                   ctypes.POINTER(DwhipParams), _vp, ctypes.POINTER(DwhipStats), _vp]),
    ("dwhip_output_stats_device", _int, [_vp, _i64, ctypes.c_float, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_int64), _vp]),
    ("dwhip_to_u16_device", _int, [_vp, _i64, ctypes.c_float, _int, _vp, _vp]),
    ("dwhip_bw_psf_device", _int, [_vp, _i64, _i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, ctypes.c_double, _vp]),
    ("dwhip_fft3_device", _int, [_vp, _vp, _i64, _i64, _i64, _int, _vp]),
    ("dwhip_engine_config_default", None, [ctypes.POINTER(EngineConfig)]),
    ("dwhip_engine_create", _int, [ctypes.POINTER(EngineConfig), _vp, ctypes.POINTER(_vp)]),
    ("dwhip_engine_destroy", None, [_vp]),
    ("dwhip_engine_submit", _int, [_vp, _vp, _i64, _i64, _i64, _vp, _i64, _i64, _i64,
                                   ctypes.POINTER(DwhipParams), _vp, _i64]),
    ("dwhip_engine_step", _int, [_vp, _P64, _int, ctypes.POINTER(_int)]),
    ("dwhip_engine_status", _int, [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int), ctypes.POINTER(ctypes.c_size_t)]),
    ("dwhip_engine_result", _int, [_vp, _i64, ctypes.POINTER(DwhipStats), _int]),
    ("dwhip_tiling_create", _int, [_i64, _i64, _i64, _i64, _i64, ctypes.POINTER(DwhipTile), _int, ctypes.POINTER(_int)]),
    ("dwhip_tile_extract_device", _int, [_vp, _i64, _i64, _i64, ctypes.POINTER(DwhipTile), ctypes.c_float, _vp, _vp]),
    ("dwhip_tiles_blend_device", _int, [ctypes.POINTER(DwhipTile), ctypes.POINTER(_vp), _int, _i64, _i64, _i64,
                                        ctypes.c_float, _vp, _vp]),
    ("dwhip_profile_enable", _int, [_int]),
    ("dwhip_profile_get", _int, [_int, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
]

N_PROFILE_KERNELS = 16


def profile_enable(on=True):
    check(lib().dwhip_profile_enable(1 if on else 0), "dwhip_profile_enable")


def profile_read():
    """{kernel name: (launches, total_ms, algorithmic bytes)} for kernels that ran."""
    out = {}
    for k in range(N_PROFILE_KERNELS):
        name = ctypes.c_char_p()
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        by = ctypes.c_double()
        check(lib().dwhip_profile_get(k, ctypes.byref(name), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(by)))
        if n.value:
            out[name.value.decode()] = (int(n.value), float(ms.value), float(by.value))
    return out

_lib = None


def lib():
    """Loads libdwhip.so (or the build named by $DWHIP_LIBRARY, e.g. one
    made with other tile parameters); raises if it has not been built."""
    global _lib
    if _lib is None:
        path = os.environ.get("DWHIP_LIBRARY") or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(
                "libdwhip.so is not built (expected at %s); run `python -m deconwolf_amd.build`" % path)
        L = ctypes.CDLL(path)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code, what=""):
    if code != DWHIP_OK:
        raise DwhipError(code, what)
    return code
