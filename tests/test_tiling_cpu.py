"""Tiling: the oracle restatement of tiling.c (oracle/tiling_oracle.py)
pinned by the reference's own tiling tests (src/tiling_ut.c:8-152) and
hand-checked divisions, and libdwhip's host planner against it."""
import numpy as np
import pytest

from deconwolf_amd import tiling
from oracle import tiling_oracle as to


def test_weight1d_reference_case():
    # tiling_ut.c:8-26: a=1, b=6, c=8, d=12; 1 on [b, c], below 1 elsewhere
    x = np.arange(14)
    w = to.weight1d(1, 6, 8, 12, x)
    assert np.all(w[(x >= 6) & (x <= 8)] == 1)
    assert np.all(w[(x < 6) | (x > 8)] < 1)
    assert w[0] == 0 and w[13] == 0 and w[1] == np.float32(1 / 6)


def test_division_2048_512_48():
    tiles = to.create(2048, 2048, 60, 512, 48)
    assert len(tiles) == 16
    xs = sorted({t["xsize"][0] for t in tiles})
    assert xs == [560, 608]
    assert tiles[1]["pos"][:4] == [0, 511, 512, 1023]   # N-divisions inner
    assert tiles[4]["xpos"][:2] == [464, 1071]


@pytest.mark.parametrize("args", [(64, 48, 3, 20, 4), (100, 37, 2, 30, 7), (2048, 2048, 1, 512, 48),
                                  (53, 53, 1, 17, 0)])
def test_weights_partition_of_unity(args):
    M, N, P, m, ov = args
    tiles = to.create(M, N, P, m, ov)
    yy, xx = np.meshgrid(np.arange(N), np.arange(M), indexing="ij")
    ws = to.weight_sum(tiles, xx, yy)
    assert np.all(ws > 0)                      # tiling_ut.c test_weights
    tot = np.zeros_like(ws)
    for t in tiles:
        tot += to.tile_weight(t, xx, yy) / ws
    np.testing.assert_allclose(tot, 1.0, rtol=1e-6)


@pytest.mark.parametrize("args", [(64, 48, 3, 20, 4), (100, 37, 2, 30, 7)])
def test_copy_paste_identity(args):
    # tiling_ut.c test_copy_paste: a constant image survives extract + put
    M, N, P, m, ov = args
    tiles = to.create(M, N, P, m, ov)
    V = np.full((P, N, M), 7, np.float32)
    outs = [to.get_tile(t, V) for t in tiles]
    assert all(np.all(o == 7) for o in outs)
    T = to.put_tiles(tiles, outs, V.shape)
    assert np.abs(T - V).max() < 1e-4


@pytest.mark.parametrize("args", [(2048, 2048, 60, 512, 48), (64, 48, 3, 20, 4), (100, 37, 2, 30, 7),
                                  (1000, 700, 5, 333, 20), (10, 10, 1, 10, 3)])
def test_host_planner_matches_oracle(args):
    M, N, P, m, ov = args
    mine = [tiling.tile_dict(t) for t in tiling.plan(M, N, P, m, ov)]
    ref = to.create(M, N, P, m, ov)
    assert mine == ref
