"""MI355X-native scaled-heavy-ball (SHB) Richardson-Lucy deconvolution.

A drop-in for deconwolf's `deconvolve_shb` method (src/method_shb.c), and
for its plain Richardson-Lucy method `deconvolve_rl` (src/method_rl.c),
whose loops run as hand-written HIP kernels for gfx950 in libdwhip.so.
"""
from .method_shb import (DwOpts, deconvolve, deconvolve_rl_hip, deconvolve_rl_hip_device,
                         deconvolve_shb_hip, deconvolve_shb_hip_device, deconvolve_shb_hip_device_noalloc,
                         fft3_device, fim_normalize_sum1, plan_cache_path, plan_get, plan_tune,
                         psf_autocrop, work_size)
from ._lib import (ITER_ABS, ITER_REL, ITER_FIXED, METRIC_MSE, METRIC_IDIV, START_FLAT, START_LP,
                   START_IDENTITY, METHOD_RL, METHOD_SHB, DwhipError)

__all__ = [
    "DwOpts", "deconvolve", "deconvolve_shb_hip", "deconvolve_shb_hip_device",
    "deconvolve_shb_hip_device_noalloc", "deconvolve_rl_hip",
    "deconvolve_rl_hip_device", "fft3_device", "fim_normalize_sum1", "psf_autocrop", "work_size",
    "DwhipError", "plan_tune", "plan_get", "plan_cache_path", "ITER_ABS", "ITER_REL", "ITER_FIXED", "METRIC_MSE", "METRIC_IDIV",
    "START_FLAT", "START_LP", "START_IDENTITY", "METHOD_RL", "METHOD_SHB",
]
