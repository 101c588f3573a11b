"""The entry point that allocates no device memory,
dwhip_shb_run_device_noalloc, against dwhip_shb_run_device through the
C-ABI: two calls in a row (the second on a smaller input, so the static
block holds the first call's leftovers), each bit-identical to the
allocating entry point's result, error trace and bg; a job larger than
the block is refused with DWHIP_ERR_SCRATCH and nothing written.
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import _lib, synth

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


def _case(seed, shape, psf_shape):
    rng = np.random.default_rng(seed)
    pz, py, px = psf_shape
    k = synth.gaussian_psf(px, py, pz, 1.5, 1.8, 2.1)
    V = synth.point_sources(shape, 40, rng=rng)
    im = synth.noisy_image(synth.blur(V, k), 60.0, rng=rng) + np.float32(5.0)
    return torch.from_numpy(im).to(DEV), torch.from_numpy(k).to(DEV)


@pytest.mark.parametrize("iter_type", [dw.ITER_FIXED, dw.ITER_REL])
def test_noalloc_equals_allocating_entry(iter_type):
    print("no-alloc entry point, iter_type", iter_type, flush=True)
    cases = [_case(1, (24, 96, 110), (15, 15, 11)), _case(2, (12, 40, 52), (9, 9, 7))]
    for im, k in cases:
        s_ref = dw.DwOpts(nIter=12, iter_type=iter_type, err_rel=0.01, maxiter=30)
        s_na = dw.DwOpts(nIter=12, iter_type=iter_type, err_rel=0.01, maxiter=30)
        ref = dw.deconvolve_shb_hip_device(im, k, s_ref)
        out = torch.full_like(im, float("nan"))
        dw.deconvolve_shb_hip_device_noalloc(im, k, s_na, out)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        assert s_na.stats["iterations"] == s_ref.stats["iterations"] > 0
        assert np.array_equal(s_na.err_trace, s_ref.err_trace)
        assert s_na.bg == s_ref.bg


def test_noalloc_refuses_jobs_beyond_the_block():
    print("no-alloc entry point, oversized job", flush=True)
    im = torch.full((60, 400, 400), 50.0, device=DEV)
    k = torch.full((31, 31, 31), 1.0 / 31 ** 3, device=DEV)
    w, _, _ = dw.work_size(tuple(im.shape), tuple(k.shape))
    L = _lib.lib()
    need = L.dwhip_job_bytes((ctypes.c_int64 * 3)(*w), 400, 400, 60, 31, 31, 31, 2, 3)
    assert need > _lib.SCRATCH_BYTES
    out = torch.zeros_like(im)
    with pytest.raises(dw.DwhipError) as e:
        dw.deconvolve_shb_hip_device_noalloc(im, k, dw.DwOpts(nIter=3, iter_type=dw.ITER_FIXED), out)
    assert e.value.code == -8
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(out)) == 0
