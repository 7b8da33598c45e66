"""GPU parity: every kernel of the hot path against the oracle (the C
restatement of the reference CPU path) on the same seeded inputs.

Bars: bit-exact for flags, boundary distances, candidate sets and pair counts;
1e-5 relative against the FP64 oracle for densities, EC weights, bandwidths.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REL = 1e-5  # relative tolerance of densities / weights against the FP64 oracle


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.where(ref != 0, np.abs(ref), 1.0)
    return np.abs(got - ref) / den


def assert_density_close(got, ref64, tol=REL):
    zero = ref64 == 0
    assert np.all(got[zero] == 0), "cells outside every cut-off must be exactly 0"
    e = rel_err(got[~zero], ref64[~zero])
    assert e.max(initial=0.0) <= tol, "max rel err %.3g at %s" % (e.max(), np.argmax(e))


@pytest.fixture(scope="module")
def scene(orc):
    """Clustered weighted points over a 300x300 mask (disk with holes)."""
    x, y, w = cases.thomas_points(12, 250, 12.0, 300.0, seed=20260101)
    x, y, w = x[:3000], y[:3000], w[:3000]
    g = orc.grid(300, 300, 0.0, 0.0, 1.0)
    mask = cases.disk_mask(300, 300, 140, n_holes=4, hole_min=10, hole_max=40, seed=20260101)
    return dict(x=x, y=y, w=w, g=g, mask=mask)


def mgrid(M, g):
    return M.grid(g.ncols, g.nrows, g.xll, g.yll, g.cs, g.nodata)


def test_boundary_flags_and_distances(ctx, orc, scene):
    import mgpukde_tiled_amd as M

    s = scene
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    flags, nb = ctx.MarkBoundary()
    oflags, onb = orc.mark_boundary(s["g"], s["mask"])
    assert nb == onb
    assert np.array_equal(flags, oflags)
    d = ctx.CalcDist2Boundary()
    od = orc.dist_to_boundary(s["x"], s["y"], s["g"], oflags)
    assert np.array_equal(d, od), "boundary distances must be bit-exact"


@pytest.mark.parametrize("shape", ["edges_only", "scattered"])
def test_boundary_distances_large_raster(ctx, orc, shape):
    # ragged raster (odd leaf counts at every level), non-integer origin and
    # cellsize; points deep inside (edges only) and among scattered holes;
    # distances bit-exact against a brute force over all boundary cells
    import mgpukde_tiled_amd as M

    nc, nr, cs = 1037, 781, 0.37
    g = orc.grid(nc, nr, 12.5, -7.25, cs)
    mask = np.ones((nr, nc), np.float32)
    r = np.random.default_rng(11)
    if shape == "scattered":
        mask[r.random((nr, nc)) < 0.002] = -9999.0
        mask[300:340, 500:900] = -9999.0
    x = (12.5 + r.uniform(-5, nc * cs + 5, 4000)).astype(np.float32)
    y = (-7.25 + r.uniform(-5, nr * cs + 5, 4000)).astype(np.float32)
    ctx.points_upload(x, y, None)
    ctx.mask_upload(mgrid(M, g), mask)
    flags, nb = ctx.MarkBoundary()
    oflags, onb = orc.mark_boundary(g, mask)
    assert nb == onb and np.array_equal(flags, oflags)
    d = ctx.CalcDist2Boundary()
    rr, cc = np.nonzero(oflags.reshape(nr, nc))
    bx = np.array([orc.col_x(g, int(c)) for c in np.unique(cc)], np.float32)
    bxm = dict(zip(np.unique(cc), bx))
    by = np.array([orc.row_y(g, int(q)) for q in np.unique(rr)], np.float32)
    bym = dict(zip(np.unique(rr), by))
    cxs = np.array([bxm[c] for c in cc], np.float32)
    cys = np.array([bym[q] for q in rr], np.float32)
    ref = np.empty(len(x))
    for k in range(0, len(x), 256):
        dx = (cxs[None, :] - x[k:k + 256, None]).astype(np.float64)
        dy = (cys[None, :] - y[k:k + 256, None]).astype(np.float64)
        ref[k:k + 256] = (dx * dx + dy * dy).min(axis=1)
    assert np.array_equal(d, ref)


@pytest.mark.parametrize("h", [6.0, 17.5])
def test_ec_weights_fixed(ctx, orc, scene, h):
    import mgpukde_tiled_amd as M

    s = scene
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    ctx.set_bandwidth(h)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    oflags, _ = orc.mark_boundary(s["g"], s["mask"])
    od = orc.dist_to_boundary(s["x"], s["y"], s["g"], oflags)
    hs = np.full(len(s["x"]), h, np.float32)
    _, ew64, one = orc.ec_weights(s["x"], s["y"], od, hs, s["g"], s["mask"])
    assert ne == one, "EC pair count must equal the oracle's"
    skipped = ew64 == 1.0
    assert np.all(ew[skipped] == 1.0)
    assert rel_err(ew, ew64).max() <= REL


def test_ec_weights_adaptive(ctx, orc, scene):
    import mgpukde_tiled_amd as M

    s = scene
    r = np.random.default_rng(5)
    hs = r.uniform(3.0, 25.0, len(s["x"])).astype(np.float32)
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    ctx.set_bandwidths(hs)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    oflags, _ = orc.mark_boundary(s["g"], s["mask"])
    od = orc.dist_to_boundary(s["x"], s["y"], s["g"], oflags)
    _, ew64, one = orc.ec_weights(s["x"], s["y"], od, hs, s["g"], s["mask"])
    assert ne == one
    assert rel_err(ew, ew64).max() <= REL


@pytest.mark.parametrize("adaptive", [False, True])
def test_density_tile(ctx, orc, scene, adaptive):
    import mgpukde_tiled_amd as M

    s = scene
    n = len(s["x"])
    r = np.random.default_rng(7)
    hs = r.uniform(4.0, 30.0, n).astype(np.float32) if adaptive else np.full(n, 12.0, np.float32)
    ew = r.uniform(0.9, 1.6, n).astype(np.float32)
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    # a tile with its own origin, not aligned to 64-cell blocks
    tile = orc.grid(230, 170, 20.0, 60.0, 1.0)
    tmask = s["mask"][70:240, 20:250]
    got = ctx.KernelDensityEstimation_tile(mgrid(M, tile), tmask)
    o32, o64, ne = orc.density_tile(s["x"], s["y"], s["w"], ew, hs, tile, tmask)
    nod = tmask == cases.NODATA
    assert np.all(got[nod] == cases.NODATA)
    assert_density_close(got[~nod], o64[~nod])
    assert ctx.count_density_pairs(mgrid(M, tile), tmask) == ne


def test_density_tile_fine_grid_mode0(ctx, orc):
    """Mode-0 style: 100 uniform points, 1000x1000 raster at cs 0.1."""
    import mgpukde_tiled_amd as M

    x, y = cases.uniform_points(100, seed=20260001)
    w = np.ones(100, np.float32)
    h = cases.plain_small_bandwidth(x, y)
    ncols, nrows, xll, yll, cs = cases.mode0_density_tile(0.1)
    tile = orc.grid(ncols, nrows, xll, yll, cs)
    ew = np.random.default_rng(3).uniform(1.0, 2.0, 100).astype(np.float32)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(h)
    ctx.set_ec_weights(ew)
    got = ctx.KernelDensityEstimation_tile(mgrid(M, tile), None)
    _, o64, ne = orc.density_tile(x, y, w, ew, np.full(100, h, np.float32), tile, None)
    assert_density_close(got, o64)
    assert ctx.count_density_pairs(mgrid(M, tile), None) == ne


def test_cutoff_edge_is_exclusive(ctx, orc):
    """Integer points with h = 1: cells at distance exactly 3 are excluded
    (d2 = 9 = 9*h2 is not < 9*h2), as the reference's strict test."""
    import mgpukde_tiled_amd as M

    x = np.array([10.5, 10.5, 20.5, 33.5], np.float32)
    y = np.array([10.5, 10.5, 5.5, 33.5], np.float32)  # a duplicate point too
    w = np.ones(4, np.float32)
    tile = orc.grid(50, 50, 0.0, 0.0, 1.0)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(1.0)
    ctx.set_ec_weights(None)
    got = ctx.KernelDensityEstimation_tile(mgrid(M, tile), None)
    _, o64, ne = orc.density_tile(x, y, w, w, np.ones(4, np.float32), tile, None)
    assert_density_close(got, o64)
    # row of the point at y=10.5 is 39; the cell 3 columns right has d2 = 9: zero
    assert got[39, 13] == 0.0 and got[39, 12] > 0.0
    assert ctx.count_density_pairs(mgrid(M, tile), None) == ne


@pytest.mark.parametrize("kind", ["uniform", "integer"])
def test_range_candidates_exact(ctx, orc, kind):
    r = np.random.default_rng(11)
    n = 3000
    if kind == "uniform":
        x = r.uniform(0, 200, n).astype(np.float32)
        y = r.uniform(0, 200, n).astype(np.float32)
        h = r.uniform(1.0, 6.0, n).astype(np.float32)
    else:
        x = r.integers(0, 60, n).astype(np.float32)
        y = r.integers(0, 60, n).astype(np.float32)
        h = np.ones(n, np.float32)
    ctx.points_upload(x, y, None)
    ctx.set_bandwidths(h)
    offs, idx = ctx.SearchRange_all()
    o_offs, o_idx = orc.range_candidates(x, y, h)
    assert np.array_equal(offs, o_offs)
    assert np.array_equal(idx, o_idx)


@pytest.mark.parametrize("adaptive", [False, True])
def test_density_at_points(ctx, orc, scene, adaptive):
    s = scene
    n = len(s["x"])
    r = np.random.default_rng(13)
    hs = r.uniform(5.0, 20.0, n).astype(np.float32) if adaptive else np.full(n, 9.0, np.float32)
    ew = r.uniform(1.0, 1.5, n).astype(np.float32)
    ctx.points_upload(s["x"], s["y"], s["w"])
    if adaptive:
        ctx.set_bandwidths(hs)
    else:
        ctx.set_bandwidth(9.0)  # the single-bandwidth kernel
    ctx.set_ec_weights(ew)
    d0 = ctx.DensityAtPoints(loo=False)
    d1 = ctx.DensityAtPoints(loo=True)
    o = orc.loo_density(s["x"], s["y"], s["w"], ew, hs)
    assert rel_err(d0, o["den0_64"]).max() <= REL
    pos = o["den1_64"] > 0
    assert np.all(d1[~pos] == 0)
    assert rel_err(d1[pos], o["den1_64"][pos]).max() <= REL


def test_loo_loglik_batched(ctx, orc, scene):
    import mgpukde_tiled_amd as M

    s = scene
    n = len(s["x"])
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    hc = np.array([3.0, 5.5, 8.0, 12.0, 20.0], np.float32)
    got = ctx.LogLikelihood(hc, edge_correction=True)
    oflags, _ = orc.mark_boundary(s["g"], s["mask"])
    od = orc.dist_to_boundary(s["x"], s["y"], s["g"], oflags)
    for k, h in enumerate(hc):
        hs = np.full(n, h, np.float32)
        ew32, ew64, _ = orc.ec_weights(s["x"], s["y"], od, hs, s["g"], s["mask"])
        o = orc.loo_density(s["x"], s["y"], s["w"], ew32, hs)
        ref = orc.loglik_fp64(o["den1_64"])
        assert abs(got[k] - ref) <= 1e-5 * n, (k, got[k], ref)


def test_abramson(ctx, orc, scene):
    import mgpukde_tiled_amd as M

    s = scene
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    h0 = ctx.BandWidth_rule_of_thumb()
    assert h0 == orc.rule_of_thumb(s["x"], s["y"], s["w"])
    hs, d0, g = ctx.Abramson(h0, -0.5, edge_correction=True)
    n = len(s["x"])
    oflags, _ = orc.mark_boundary(s["g"], s["mask"])
    od = orc.dist_to_boundary(s["x"], s["y"], s["g"], oflags)
    ew32, ew64, _ = orc.ec_weights(s["x"], s["y"], od, np.full(n, h0, np.float32), s["g"], s["mask"])
    _, p64 = orc.pilot_density(s["x"], s["y"], s["w"], ew32, h0)
    assert rel_err(d0, p64).max() <= REL
    g64 = orc.compgml64(p64)
    assert abs(g - g64) <= 1e-6 * g64
    h64 = orc.abramson64(h0, -0.5, g64, p64)
    assert rel_err(hs, h64).max() <= REL


@pytest.mark.parametrize("adaptive", [False, True])
def test_density_at_points_cutoff_ties(ctx, orc, adaptive):
    """Integer coordinates with h = 1 (or h in {1, 2}): many pairs sit exactly
    on the cut-off d2 = 9 h2 and must be excluded, as in the oracle."""
    r = np.random.default_rng(17)
    n = 2000
    x = r.integers(0, 40, n).astype(np.float32)
    y = r.integers(0, 40, n).astype(np.float32)
    w = np.ones(n, np.float32)
    hs = (r.integers(1, 3, n) if adaptive else np.ones(n)).astype(np.float32)
    ctx.points_upload(x, y, w)
    if adaptive:
        ctx.set_bandwidths(hs)
    else:
        ctx.set_bandwidth(1.0)
    ctx.set_ec_weights(None)
    d0 = ctx.DensityAtPoints(loo=False)
    d1 = ctx.DensityAtPoints(loo=True)
    o = orc.loo_density(x, y, w, w, hs)
    assert rel_err(d0, o["den0_64"]).max() <= REL
    pos = o["den1_64"] > 0
    assert np.all(d1[~pos] == 0)
    assert rel_err(d1[pos], o["den1_64"][pos]).max() <= REL


@pytest.mark.parametrize("hs_range", [(180.0, 180.0), (450.0, 450.0), (40.0, 450.0)])
def test_ec_weights_wide_boxes(ctx, orc, hs_range):
    # boxes wider than one 1024-column chunk go to the wide EC kernel (whole
    # box in LDS); an adaptive mix sends points to both kernels in one pass
    import mgpukde_tiled_amd as M

    nc, nr = 1400, 900
    g = orc.grid(nc, nr, 100.0, 50.0, 1.0)
    r = np.random.default_rng(17)
    mask = np.ones((nr, nc), np.float32)
    mask[r.random((nr, nc)) < 0.01] = -9999.0
    mask[200:420, 300:700] = -9999.0
    mask[:, 1300:] = -9999.0
    n = 160
    x = (100.0 + r.uniform(0, nc, n)).astype(np.float32)
    y = (50.0 + r.uniform(0, nr, n)).astype(np.float32)
    w = r.uniform(0.5, 2.0, n).astype(np.float32)
    hs = r.uniform(hs_range[0], hs_range[1], n).astype(np.float32)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(mgrid(M, g), mask)
    ctx.set_bandwidths(hs)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    oflags, _ = orc.mark_boundary(g, mask)
    od = orc.dist_to_boundary(x, y, g, oflags)
    _, ew64, one = orc.ec_weights(x, y, od, hs, g, mask)
    assert ne == one
    assert rel_err(ew, ew64).max() <= REL


@pytest.mark.parametrize("origin_cs", [(12.5, -7.25, 0.37), (1000.25, 50.0, 0.5), (300000.1, 4000000.0, 2.0),
                                       (-55.0, 10.0, 1.0)])
def test_ec_weights_grid_geometries(ctx, orc, origin_cs):
    # the EC kernels form column centres in FP32 when the host finds every
    # column's centre exact in FP32 (integer-ish grids), else in FP64: both
    # must give the oracle's pair counts and weights (narrow and wide boxes)
    import mgpukde_tiled_amd as M

    xll, yll, cs = origin_cs
    nc, nr = 700, 500
    g = orc.grid(nc, nr, xll, yll, cs)
    r = np.random.default_rng(23)
    mask = np.ones((nr, nc), np.float32)
    mask[r.random((nr, nc)) < 0.01] = -9999.0
    mask[100:180, 200:420] = -9999.0
    n = 200
    x = (xll + r.uniform(0, nc * cs, n)).astype(np.float32)
    y = (yll + r.uniform(0, nr * cs, n)).astype(np.float32)
    hs = (cs * r.uniform(3.0, 260.0, n)).astype(np.float32)
    ctx.points_upload(x, y, None)
    ctx.mask_upload(mgrid(M, g), mask)
    ctx.set_bandwidths(hs)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    oflags, _ = orc.mark_boundary(g, mask)
    od = orc.dist_to_boundary(x, y, g, oflags)
    _, ew64, one = orc.ec_weights(x, y, od, hs, g, mask)
    assert ne == one
    assert rel_err(ew, ew64).max() <= REL
