"""The device Born-Wolf generator (dwhip_bw_psf_device) against the host
restatement of the same algorithm (deconwolf_amd.psf_bw.bw_psf, pinned by
closed forms in test_filters_cpu.py): same values to float rounding, exact
octant and z symmetry, maximum at the centre."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from deconwolf_amd import psf_bw

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,P,lam,dz", [(31, 15, 520.0, 200.0), (181, 79, 600.0, 200.0), (64, 20, 460.0, 300.0)])
def test_device_psf_matches_host(N, P, lam, dz):
    print("bw psf", N, P, lam, flush=True)
    dev = psf_bw.bw_psf_device(N, P, lam, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=dz).cpu().numpy()
    host = psf_bw.bw_psf(N, N, P, lam, NA=1.40, ni=1.515, dx_nm=65.0, dz_nm=dz, normalize=False)
    assert dev.shape == (P, N, N)
    assert np.abs(dev - host).max() <= 1e-6 * host.max()
    assert np.array_equal(dev, dev[::-1])
    assert np.array_equal(dev, dev[:, ::-1, :]) and np.array_equal(dev, dev[:, :, ::-1])
    assert np.array_equal(dev, np.transpose(dev, (0, 2, 1)))
    if N % 2 and P % 2:
        assert np.unravel_index(np.argmax(dev), dev.shape) == (P // 2, N // 2, N // 2)
