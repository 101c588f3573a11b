"""GPU side of the output path: the device statistics and uint16
conversions (dwhip_output_stats_device, dwhip_to_u16_device) against the
C semantics of deconwolf's writers, restated with numpy, bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import deconwolf_amd as dw
from deconwolf_amd import output

pytestmark = pytest.mark.gpu

DEV = "cuda:0"

# half-way cases, rounding edges, clamps and non-finite values
EDGES = np.array([0.0, 0.5, 1.5, 2.5, 3.49999, 65534.5, 65535.0, 65535.4, 65535.6, 70000.0, -0.4, -3.0,
                  np.inf, -np.inf, np.nan, 1e-30], np.float32)


def ref_u16(v, scaling, mode):
    x = (v.astype(np.float32) * np.float32(scaling)).astype(np.float32)
    if mode == output.U16_ROUND:          # fim_tiff.c:1142-1149
        x = np.where(x > 65535, np.float32(65535), x)
        x = np.where(np.isfinite(x), x, np.float32(0))
        x64 = x.astype(np.float64)       # |x| + 0.5 is exact in double
        r = np.sign(x64) * np.floor(np.abs(x64) + 0.5)
    elif mode == output.U16_TRUNC:        # (uint16_t)(v*s)
        r = np.trunc(x)
    else:                                 # nearbyintf: half to even
        r = np.rint(x)
    r = np.where(np.isnan(r), 0, r)
    return np.clip(r, 0, 65535).astype(np.uint16)


@pytest.mark.parametrize("mode", [output.U16_ROUND, output.U16_TRUNC, output.U16_RINT])
def test_to_u16_modes(mode):
    rng = np.random.default_rng(mode)
    v = np.concatenate([EDGES, (rng.random(100000) * 70000).astype(np.float32),
                        np.arange(0, 200, 0.5, dtype=np.float32)])
    for scaling in (1.0, 0.731):
        got = output.to_u16(torch.from_numpy(v).to(DEV), scaling, mode)
        want = ref_u16(v, scaling, mode)
        assert np.array_equal(got, want), (mode, scaling, v[got != want][:8], got[got != want][:8],
                                           want[got != want][:8])


def test_output_stats():
    rng = np.random.default_rng(5)
    v = (rng.random((7, 33, 45)) * 900 + 12.5).astype(np.float32)
    v[rng.random(v.shape) < 0.3] = np.float32(12.5)
    v[0, 0, 0] = np.nan
    vmax, nbg = output.stats(torch.from_numpy(v).to(DEV), 12.5)
    assert vmax == np.nanmax(v)
    assert nbg == int(np.sum(v == np.float32(12.5)))
    w = -np.abs(v)
    w[0, 0, 0] = -1e30
    vmax, _ = output.stats(torch.from_numpy(w).to(DEV), 0.0)
    assert vmax == np.nanmax(w)


@pytest.mark.parametrize("name,tiled", [("o.tif", False), ("o.npy", False), ("o.tif", True), ("o.npy", True)])
def test_imwrite_matches_reference_rules(tmp_path, name, tiled):
    rng = np.random.default_rng(9)
    out = (rng.random((4, 20, 24)) * 3000 + 7).astype(np.float32)
    s = dw.DwOpts(bg=7.0)
    path = str(tmp_path / name)
    sc = output.imwrite(path, torch.from_numpy(out).to(DEV), s, tiled=tiled)
    vmax = out.max()
    if tiled:
        assert sc == float(np.float32(65535) / np.float32(vmax))
        mode = output.U16_RINT if name.endswith(".npy") else output.U16_TRUNC
    else:
        assert sc == float(np.float32(65535.0 / float(vmax)))
        mode = output.U16_TRUNC if name.endswith(".npy") else output.U16_ROUND
    want = ref_u16(out, sc, mode)
    if name.endswith(".npy"):
        got = np.load(path)
    else:
        from PIL import Image
        im = Image.open(path)
        got = []
        for k in range(im.n_frames):
            im.seek(k)
            got.append(np.array(im).astype(np.uint16))
        got = np.stack(got)
    assert got.dtype == np.uint16 and np.array_equal(got, want)
