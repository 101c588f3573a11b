"""Host side of the output path (deconwolf_amd/output.py): the 16-bit
scaling rules of dw_run and deconvolve_tiles and the TIFF/NPY writers,
read back with Pillow and numpy (independent readers)."""
import numpy as np
import pytest

from deconwolf_amd import output


def test_scaling_rules():
    # scaling_for_u16: (2^16-1)/max in double, rounded to float (dw.c:98-103)
    assert output.scaling_for_u16(1234.5) == float(np.float32(65535.0 / 1234.5))
    # tiled: 65535/max in float when no scaling is given (dw.c:2166-2172)
    assert output.tiled_scaling(1234.5, -1) == float(np.float32(65535) / np.float32(1234.5))
    assert output.tiled_scaling(1234.5, 3.0) == 3.0
    assert output.tiled_scaling(0.0, -1) == -1.0


def _pil_pages(path):
    from PIL import Image
    im = Image.open(path)
    pages = []
    for k in range(im.n_frames):
        im.seek(k)
        pages.append(np.array(im))
    return pages, im


@pytest.mark.parametrize("dtype", [np.uint16, np.float32])
@pytest.mark.parametrize("big", [False, True])
def test_tiff_roundtrip(tmp_path, dtype, big):
    rng = np.random.default_rng(1)
    P, N, M = 3, 7, 11          # odd sizes: width != height, odd plane bytes
    if dtype == np.uint16:
        vol = rng.integers(0, 65536, size=(P, N, M)).astype(np.uint16)
    else:
        vol = rng.standard_normal((P, N, M)).astype(np.float32)
    path = str(tmp_path / "v.tif")
    output.write_tiff(path, vol, description=output.imagej_description(P, 200.0), software="dwhip",
                      resolution=1 / 65.0, bigtiff=big)
    try:
        pages, im = _pil_pages(path)
    except Exception as e:  # pragma: no cover - Pillow without BigTIFF support
        if big:
            pytest.skip("Pillow cannot read BigTIFF: %s" % e)
        raise
    assert len(pages) == P
    for z in range(P):
        assert pages[z].shape == (N, M)
        assert np.array_equal(pages[z].astype(dtype), vol[z])
    im.seek(0)
    assert im.tag_v2[270].startswith("ImageJ=1.52r\nimages=3\nslices=3\nunit=nm\nspacing=200.0")
    assert im.tag_v2[305] == "dwhip"


def test_npy_roundtrip(tmp_path):
    vol = np.arange(2 * 3 * 4, dtype=np.uint16).reshape(2, 3, 4)
    path = str(tmp_path / "v.npy")
    output.write_npy(path, vol)
    back = np.load(path)
    assert back.dtype == np.uint16 and back.shape == (2, 3, 4) and np.array_equal(back, vol)
