"""Edge cases on the GPU: empty and single-point inputs, masks without a valid
cell, a 1x1 raster, points far outside the raster, a raster narrower than
one 64-cell block -- each against the oracle or the closed form."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu


def test_no_points_is_rejected(ctx):
    # the reference refuses an empty point set at input (ReadSamplePoints,
    # tiled:1625-1649); every compute call here reports MKDE_E_STATE
    e = np.zeros(0, np.float32)
    ctx.points_upload(e, e, None)
    g = M.grid(70, 50, 0.0, 0.0, 1.0)
    mask = cases.disk_mask(70, 50, 20.0)
    for call in (lambda: ctx.set_bandwidth(2.0), lambda: ctx.KernelDensityEstimation_tile(g, mask),
                 lambda: ctx.BandWidth_rule_of_thumb()):
        with pytest.raises(M.MkdeError) as ei:
            call()
        assert ei.value.code == 4


def test_single_point_closed_form(ctx, orc):
    x = np.array([10.25], np.float32)
    y = np.array([12.75], np.float32)
    ctx.points_upload(x, y, None)
    ctx.set_bandwidth(1.5)
    ctx.set_ec_weights(np.ones(1, np.float32))
    g = M.grid(24, 26, 0.0, 0.0, 1.0)
    den = ctx.KernelDensityEstimation_tile(g, None)
    _, o64, _ = orc.density_tile(x, y, np.ones(1, np.float32), np.ones(1, np.float32), np.full(1, 1.5, np.float32),
                                 orc.grid(24, 26, 0.0, 0.0, 1.0), None)
    pos = o64 > 0
    assert np.all(den[~pos] == 0.0)
    assert np.max(np.abs(den[pos] - o64[pos]) / o64[pos]) <= 1e-5
    # closed form at the nearest cell centre: K(d) = exp(-d^2 / (2 h^2)) / (2 pi h^2)
    r, c = 26 - 1 - 12, 10
    d2 = (c + 0.5 - 10.25) ** 2 + (12.5 - 12.75) ** 2
    assert abs(den[r, c] - np.exp(-d2 / 4.5) / (2 * np.pi * 2.25)) <= 1e-6


def test_mask_without_valid_cells(ctx):
    x, y = cases.uniform_points(500, 50.0, seed=3)
    ctx.points_upload(x, y, None)
    g = M.grid(50, 50, 0.0, 0.0, 1.0)
    mask = np.full((50, 50), -9999.0, np.float32)
    ctx.mask_upload(g, mask)
    flags, nb = ctx.MarkBoundary()
    assert nb == 0 and not flags.any()
    d = ctx.CalcDist2Boundary()
    assert np.all(d == np.finfo(np.float64).max)
    ctx.set_bandwidth(2.0)
    ew, n_eval = ctx.EdgeCorrectionWeightsExact()
    assert np.all(ew == 1.0) and n_eval == 0
    den = ctx.KernelDensityEstimation_tile(g, mask)
    assert np.all(den == -9999.0)


def test_one_cell_raster_and_far_points(ctx, orc):
    x = np.array([0.5, 1e6, -1e6, 0.4], np.float32)
    y = np.array([0.5, 0.5, 3.0, 0.6], np.float32)
    w = np.array([1.0, 2.0, 3.0, 0.5], np.float32)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(0.7)
    ctx.set_ec_weights(np.ones(4, np.float32))
    g = M.grid(1, 1, 0.0, 0.0, 1.0)
    den = ctx.KernelDensityEstimation_tile(g, None)
    _, o64, _ = orc.density_tile(x, y, w, np.ones(4, np.float32), np.full(4, 0.7, np.float32),
                                 orc.grid(1, 1, 0.0, 0.0, 1.0), None)
    assert den.shape == (1, 1) and abs(den[0, 0] - o64[0, 0]) / o64[0, 0] <= 1e-5


def test_narrow_raster_edge_blocks(ctx, orc):
    # 5 x 200 and 200 x 3 rasters: every 64x64 block is an edge block
    x, y, w = cases.thomas_points(3, 80, 4.0, 60.0, seed=9)
    ctx.points_upload(x, y, w)
    hs = np.linspace(1.0, 4.0, len(x)).astype(np.float32)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(np.ones(len(x), np.float32))
    mx, my = float(np.median(x)), float(np.median(y))
    for nc, nr in ((5, 200), (200, 3)):
        xll, yll = mx - nc * 0.15, my - nr * 0.15   # centred on the points
        g = M.grid(nc, nr, xll, yll, 0.3)
        den = ctx.KernelDensityEstimation_tile(g, None)
        _, o64, _ = orc.density_tile(x, y, w, np.ones(len(x), np.float32), hs, orc.grid(nc, nr, xll, yll, 0.3), None)
        pos = o64 > 0
        assert pos.any()
        assert np.all(den[~pos] == 0.0)
        assert np.max(np.abs(den[pos] - o64[pos]) / o64[pos]) <= 1e-5
