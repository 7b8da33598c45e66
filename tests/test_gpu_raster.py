"""A GeoTIFF-style raster of tiles in one pass (mkde_density_raster) against
the oracle run tile by tile with the origins the reference gives each tile
(AsciiRasterFromGeoTIFFTile, kde_cuda_kdtr_tiled.cu:2218-2220; tile
geotransform GeoTIFF.cpp:190-195), and against mkde_density_tile on each
tile (the same kernel on the same geometry: bit-identical)."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu
REL = 1e-5


def tile_grid(orc, gt, ncols, nrows, xoff, yoff, tc, tr, nodata):
    g0 = gt[0] + gt[1] * xoff
    g3 = gt[3] + gt[5] * yoff
    return orc.grid(tc, tr, float(np.float32(g0)), float(np.float32(g3 + tr * gt[5])), float(np.float32(gt[1])),
                    nodata)


def tiles(ncols, nrows, tw, th):
    for yoff in range(0, nrows, th):
        for xoff in range(0, ncols, tw):
            yield xoff, yoff, min(tw, ncols - xoff), min(th, nrows - yoff)


@pytest.mark.parametrize("with_mask", [True, False])
def test_raster_tiles_match_oracle_per_tile(orc, ctx, with_mask):
    ncols, nrows, tw, th = 520, 410, 200, 150     # ragged edge tiles, tiles not multiples of 64
    gt = [1234.56, 0.37, 0.0, 987.65, 0.0, -0.37]
    nodata = -9999.0
    r = np.random.default_rng(41)
    ext_x, ext_y = ncols * gt[1], nrows * -gt[5]
    n = 1500
    x = (gt[0] + r.uniform(-2, ext_x + 2, n)).astype(np.float32)
    y = (gt[3] - r.uniform(-2, ext_y + 2, n)).astype(np.float32)
    w = r.uniform(0.5, 2.0, n).astype(np.float32)
    hs = r.uniform(1.0, 6.0, n).astype(np.float32)
    ew = r.uniform(1.0, 1.5, n).astype(np.float32)
    mask = None
    if with_mask:
        mask = np.ones((nrows, ncols), np.float32)
        mask[r.random((nrows, ncols)) < 0.05] = nodata
        mask[100:180, 60:300] = nodata
    ctx.points_upload(x, y, w)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    den = ctx.KernelDensityEstimation_raster(ncols, nrows, gt, nodata, tw, th, mask)
    for xoff, yoff, tc, tr in tiles(ncols, nrows, tw, th):
        g = tile_grid(orc, gt, ncols, nrows, xoff, yoff, tc, tr, nodata)
        tm = None if mask is None else np.ascontiguousarray(mask[yoff:yoff + tr, xoff:xoff + tc])
        _, o64, _ = orc.density_tile(x, y, w, ew, hs, g, tm)
        got = den[yoff:yoff + tr, xoff:xoff + tc]
        valid = np.ones((tr, tc), bool) if tm is None else tm != nodata
        assert np.all(got[~valid] == nodata)
        pos = valid & (o64 > 0)
        assert np.all(got[valid & ~pos] == 0.0), (xoff, yoff)
        assert np.max(np.abs(got[pos] - o64[pos]) / o64[pos]) <= REL, (xoff, yoff)
        one = ctx.KernelDensityEstimation_tile(M.grid(g.ncols, g.nrows, g.xll, g.yll, g.cs, nodata), tm)
        assert np.array_equal(one, got), (xoff, yoff)


def test_raster_single_tile_equals_tile_call(ctx):
    # tile size >= raster: one tile whose origin is the raster's own
    x, y = cases.uniform_points(4000, 300.0, seed=7)
    w = None
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(4.0)
    ctx.set_ec_weights(np.ones(len(x), np.float32))
    gt = [0.0, 1.0, 0.0, 300.0, 0.0, -1.0]
    a = ctx.KernelDensityEstimation_raster(300, 300, gt, -9999.0, 4096, 4096, None)
    b = ctx.KernelDensityEstimation_tile(M.grid(300, 300, 0.0, 0.0, 1.0), None)
    assert np.array_equal(a, b)


def test_raster_keep_on_device_then_download(ctx):
    x, y = cases.uniform_points(3000, 256.0, seed=8)
    w = None
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(3.0)
    ctx.set_ec_weights(np.ones(len(x), np.float32))
    gt = [0.0, 1.0, 0.0, 256.0, 0.0, -1.0]
    a = ctx.KernelDensityEstimation_raster(256, 256, gt, -9999.0, 128, 64, None)
    assert ctx.KernelDensityEstimation_raster(256, 256, gt, -9999.0, 128, 64, None, keep_on_device=True) is None
    assert np.array_equal(ctx.tile_download(), a)


def test_raster_argument_errors(ctx):
    ctx.points_upload(*cases.uniform_points(10, 10.0, seed=1))
    ctx.set_bandwidth(1.0)
    with pytest.raises(M.MkdeError):
        ctx.KernelDensityEstimation_raster(10, 10, [0, 1, 0, 10, 0, -1], -9999.0, 0, 4, None)
    with pytest.raises(M.MkdeError):
        ctx.KernelDensityEstimation_raster(10, 10, [0, float("nan"), 0, 10, 0, -1], -9999.0, 4, 4, None)
    with pytest.raises(M.MkdeError):
        ctx.KernelDensityEstimation_raster(10, 10, [0, -1.0, 0, 10, 0, -1], -9999.0, 4, 4, None)


def test_split_blocks_dense_cluster(orc, ctx):
    # 24k points inside 40 x 40 cells: every block near them has more
    # candidates than the split threshold, so its work is cut into parts whose
    # FP64 partials are merged; the result stays within 1e-5 of the FP64
    # oracle, is deterministic, and the raster path agrees bit for bit
    r = np.random.default_rng(5)
    n = 24000
    x = (r.normal(60.0, 8.0, n)).astype(np.float32)
    y = (r.normal(70.0, 8.0, n)).astype(np.float32)
    w = r.uniform(0.5, 2.0, n).astype(np.float32)
    hs = r.uniform(2.0, 9.0, n).astype(np.float32)
    ew = np.ones(n, np.float32)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    g = M.grid(130, 140, 0.0, 0.0, 1.0)
    mask = np.ones((140, 130), np.float32)
    mask[60:64, :] = -9999.0
    a = ctx.KernelDensityEstimation_tile(g, mask)
    b = ctx.KernelDensityEstimation_tile(g, mask)
    assert np.array_equal(a, b)
    c = ctx.KernelDensityEstimation_raster(130, 140, [0.0, 1.0, 0.0, 140.0, 0.0, -1.0], -9999.0, 4096, 4096, mask)
    assert np.array_equal(a, c)
    og = orc.grid(130, 140, 0.0, 0.0, 1.0)
    _, o64, _ = orc.density_tile(x, y, w, ew, hs, og, mask)
    valid = mask != -9999.0
    assert np.all(a[~valid] == -9999.0)
    pos = valid & (o64 > 0)
    assert np.max(np.abs(a[pos] - o64[pos]) / o64[pos]) <= REL
    ctx.density_stats(True)
    ctx.KernelDensityEstimation_tile(g, mask)
    st = ctx.density_stats(False)
    # the same job without the work list (raster order, no split): every
    # candidate is classified exactly once either way
    import os
    os.environ["MKDE_DENSITY_ORDER"] = "0"
    try:
        c2 = M.Context(0)
    finally:
        del os.environ["MKDE_DENSITY_ORDER"]
    c2.points_upload(x, y, w)
    c2.set_bandwidths(hs)
    c2.set_ec_weights(ew)
    c2.density_stats(True)
    d = c2.KernelDensityEstimation_tile(g, mask)
    st2 = c2.density_stats(False)
    c2.close()
    assert st["candidates"] == st2["candidates"] and st["full_blocks"] == st2["full_blocks"]
    assert np.max(np.abs(d[pos] - a[pos]) / a[pos]) <= 1e-6
