"""The workloads as stated, end to end on the GPU against the oracle.

* plain_small: 100 uniform points, EC mask 100x100 at cs 1, density raster
  1000x1000 at cs 0.1 with an all-ones mask, the fixed bandwidth
  h = sqrt((sx^2+sy^2)/2) n^(-1/6) given explicitly (h_option -1), EC on.
* h_option 2 with Hooke-Jeeves: the selected (h0, alpha), the Abramson
  bandwidths, EC weights and the raster (CPU path tiled:550-578, GPU path
  tiled:1064-1128).
* mixture_search at an intermediate size: 20,000 points on a 1024^2 mask, 64
  log-spaced candidates 0.1x-10x the rule of thumb in batched passes.  Every
  candidate's EC weights are checked on a point sample, every candidate's
  leave-one-out log-likelihood against the FP64 oracle given those weights,
  and the top three candidates end to end (oracle EC for every point).
* lognormal_tiles at a reduced size: log-normal cluster intensities, fixed
  h = 16, EC on, the raster as 512^2 tiles; per-window relative error.
* a failed point upload leaves the context consistent.
"""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu
REL = 1e-5


def mgrid(g):
    return M.grid(g.ncols, g.nrows, g.xll, g.yll, g.cs, g.nodata)


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.abs(b)


def test_plain_small_as_stated(ctx, orc):
    x, y = cases.uniform_points(100, 100.0, seed=20260001)
    w = np.ones(100, np.float32)
    h = cases.plain_small_bandwidth(x, y)
    hs = np.full(100, h, np.float32)
    g_ec = orc.grid(100, 100, 0.0, 0.0, 1.0)
    mask_ec = cases.all_valid(100, 100, 0.0)
    ncols, nrows, xll, yll, cs = cases.mode0_density_tile(0.1)
    assert (ncols, nrows) == (1000, 1000)
    tile = orc.grid(ncols, nrows, xll, yll, cs)
    tmask = cases.all_valid(ncols, nrows, 1.0)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(mgrid(g_ec), mask_ec)
    ctx.tile_upload(mgrid(tile), tmask)
    ctx.set_bandwidths(hs)
    st = ctx.run_job(hoption=-1, edge_correction=True)
    den = ctx.tile_download()
    ew = ctx.get_ec_weights()
    assert np.array_equal(ctx.get_bandwidths(), hs)
    # the reference CPU path: boundary, distances, EC weights, the cell loop
    flags, _ = orc.mark_boundary(g_ec, mask_ec)
    d = orc.dist_to_boundary(x, y, g_ec, flags)
    ew32, ew64, ne_ec = orc.ec_weights(x, y, d, hs, g_ec, mask_ec)
    assert st["n_ec_pairs"] == ne_ec
    assert rel(ew, ew64).max() <= REL
    o32, o64, ne = orc.density_tile(x, y, w, ew64.astype(np.float32), hs, tile, tmask)
    r32, _, _ = orc.density_tile(x, y, w, ew32, hs, tile, tmask, want64=False)
    assert ctx.count_density_pairs(mgrid(tile), tmask) == ne
    pos = o64 > 0
    assert np.all(den[~pos] == 0.0)
    assert rel(den[pos], o64[pos]).max() <= REL
    # the reference's own float arithmetic (EC weights and cell sums in float)
    assert rel(den[pos], r32[pos]).max() <= REL
    # mass: EC on and every kernel disk inside the EC mask except near its
    # edge, so sum(den) * cs^2 is close to the number of points
    mass = float(den.astype(np.float64).sum()) * float(np.float32(cs)) ** 2
    assert abs(mass - float(o64.sum()) * float(np.float32(cs)) ** 2) <= 1e-5 * mass


@pytest.fixture(scope="module")
def clustered(orc):
    x, y, w = cases.thomas_points(6, 150, 7.0, 96.0, seed=20260311)
    g = orc.grid(96, 96, 0.0, 0.0, 1.0)
    mask = cases.disk_mask(96, 96, 44, n_holes=3, hole_min=5, hole_max=14, seed=20260311)
    flags, _ = orc.mark_boundary(g, mask)
    dist = orc.dist_to_boundary(x, y, g, flags)
    return dict(x=x, y=y, w=w, g=g, mask=mask, dist=dist)


def oracle_adaptive_h(orc, s, h, alpha):
    n = len(s["x"])
    h = np.float32(h)
    _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], np.full(n, h, np.float32), s["g"], s["mask"])
    _, p64 = orc.pilot_density(s["x"], s["y"], s["w"], ew64.astype(np.float32), h)
    g64 = orc.compgml64(p64)
    return orc.abramson64(float(h), float(alpha), g64, p64).astype(np.float32)


def oracle_loglik_adaptive(orc, s, h, alpha):
    hs = oracle_adaptive_h(orc, s, h, alpha)
    _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    o = orc.loo_density(s["x"], s["y"], s["w"], ew64.astype(np.float32), hs)
    return orc.loglik_fp64(o["den1_64"])


def test_adaptive_hooke_jeeves_job_end_to_end(ctx, orc, clustered):
    s = clustered
    g = mgrid(s["g"])
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(g, s["mask"])
    ctx.tile_upload(g, s["mask"])
    st = ctx.run_job(hoption=2, edge_correction=True, optimise=True)
    h0 = orc.rule_of_thumb(s["x"], s["y"], s["w"])
    assert st["h0"] == h0
    step_h = float(np.float32(h0) / np.float32(10))
    (oh, oa, _), _ = orc.hooke_jeeves(h0, -0.5, step_h, 0.1, lambda hh, aa: oracle_loglik_adaptive(orc, s, hh, aa))
    # the selected bandwidth pair is identical
    assert (st["h_opt"], st["alpha_opt"]) == (oh, oa)
    hs = ctx.get_bandwidths()
    href = oracle_adaptive_h(orc, s, oh, oa)
    assert rel(hs, href).max() <= REL
    ew = ctx.get_ec_weights()
    _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    assert rel(ew, ew64).max() <= REL
    den = ctx.tile_download()
    _, o64, _ = orc.density_tile(s["x"], s["y"], s["w"], ew64.astype(np.float32), href, s["g"], s["mask"])
    v = s["mask"] != cases.NODATA
    assert np.all(den[~v] == cases.NODATA)
    pos = v & (o64 > 0)
    assert rel(den[pos], o64[pos]).max() <= REL


def search_candidates(h_ref, K=64):
    return (np.float32(h_ref) * (10.0 ** (-1.0 + 2.0 * np.arange(K) / (K - 1)))).astype(np.float32)


def test_search_64_candidates_intermediate(ctx, orc):
    side, n = 1024, 20000
    x, y = cases.mixture_points(n, float(side), n_centres=64, seed=20260313)
    w = np.ones(n, np.float32)
    g = orc.grid(side, side, 0.0, 0.0, 1.0)
    mask = cases.all_valid(side, side)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(mgrid(g), mask)
    h_ref = ctx.BandWidth_rule_of_thumb()
    hc = search_candidates(h_ref)
    got = ctx.LogLikelihood(hc, edge_correction=True)
    flags, _ = orc.mark_boundary(g, mask)
    d = orc.dist_to_boundary(x, y, g, flags)
    assert np.array_equal(ctx.CalcDist2Boundary(), d)
    r = np.random.default_rng(20260313)
    ref = np.empty(len(hc))
    for k, hk in enumerate(hc):
        hs = np.full(n, hk, np.float32)
        ctx.set_bandwidth(float(hk))
        ew, _ = ctx.EdgeCorrectionWeightsExact()
        # EC weights of this candidate on a sample: half near the boundary
        near = np.flatnonzero(d < 9.0 * float(np.float32(hk * hk)))
        far = np.flatnonzero(d >= 9.0 * float(np.float32(hk * hk)))
        smp = np.concatenate([r.choice(near, min(20, len(near)), replace=False) if len(near) else near,
                              r.choice(far, min(5, len(far)), replace=False) if len(far) else far])
        _, ew64, _ = orc.ec_weights(x[smp], y[smp], d[smp], hs[smp], g, mask)
        assert rel(ew[smp], ew64).max() <= REL, k
        # leave-one-out densities and log-likelihood given those weights
        den64, _ = orc.point_density_binned(x, y, w, ew, hs, loo=True)
        ref[k] = orc.loglik_fp64(den64)
    # each GPU log-likelihood is a sum of n logs of FP32 densities with
    # relative error < 1e-6: |dL| < n * 1e-6
    assert np.max(np.abs(got - ref)) <= 1e-6 * n
    k_best = int(np.argmax(got))
    assert k_best == int(np.argmax(ref))
    srt = np.sort(ref)
    assert srt[-1] - srt[-2] > 2e-6 * n   # the argmax is decided well outside the error
    # the three best candidates end to end: oracle EC weights of every point
    for k in np.argsort(-got)[:3]:
        hs = np.full(n, hc[k], np.float32)
        _, ew64, _ = orc.ec_weights(x, y, d, hs, g, mask)
        den64, _ = orc.point_density_binned(x, y, w, ew64.astype(np.float32), hs, loo=True)
        assert abs(got[k] - orc.loglik_fp64(den64)) <= 1e-6 * n, k


def test_lognormal_tiles_h16_reduced(ctx, orc):
    side, tile, n, h = 2048, 512, 200_000, 16.0
    x, y = cases.lognormal_clusters(n, float(side), 256, 64.0, seed=20260314)
    w = np.ones(n, np.float32)
    gt = [0.0, 1.0, 0.0, float(side), 0.0, -1.0]
    g = orc.grid(side, side, 0.0, 0.0, 1.0)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidth(h)
    ctx.mask_upload(mgrid(g), cases.all_valid(side, side, 1.0))
    ew, _ = ctx.EdgeCorrectionWeightsExact()
    den = ctx.KernelDensityEstimation_raster(side, side, gt, -9999.0, tile, tile, None)
    hs = np.full(n, h, np.float32)
    # EC weights on the points nearest the raster edge (the only ones corrected)
    d = ctx.CalcDist2Boundary()
    near = np.argsort(d)[:200]
    _, ew64, _ = orc.ec_weights(x[near], y[near], d[near], hs[near], g, cases.all_valid(side, side, 1.0))
    assert rel(ew[near], ew64).max() <= REL
    r = np.random.default_rng(20260314)
    for _ in range(6):
        k = int(r.integers(0, n))
        c0 = int(min(max(int(x[k]) - 32, 0), side - 64))
        r0 = int(min(max(side - 1 - int(y[k]) - 32, 0), side - 64))
        # keep the window inside one tile so its origin is that tile's
        c0 = min(c0, (c0 // tile) * tile + tile - 64)
        r0 = min(r0, (r0 // tile) * tile + tile - 64)
        win = orc.grid(64, 64, float(c0), float(side - r0 - 64), 1.0)
        sel = orc.tile_point_filter(x, y, hs, win)
        _, o64, _ = orc.density_tile(x[sel], y[sel], w[sel], ew[sel], hs[sel], win, None)
        got = den[r0:r0 + 64, c0:c0 + 64]
        pos = o64 > 0
        assert np.all(got[~pos] == 0.0)
        assert rel(got[pos], o64[pos]).max() <= REL, (c0, r0)


def test_failed_point_upload_keeps_context_consistent(ctx):
    r = np.random.default_rng(5)
    x = r.uniform(0, 50, 100).astype(np.float32)
    y = r.uniform(0, 50, 100).astype(np.float32)
    ctx.points_upload(x, y, None)
    ctx.set_bandwidth(3.0)
    a = ctx.DensityAtPoints()
    bx = r.uniform(0, 50, 200).astype(np.float32)
    by = r.uniform(0, 50, 200).astype(np.float32)
    bx[150] = np.nan
    with pytest.raises(M.MkdeError) as e:
        ctx.points_upload(bx, by, None)
    assert e.value.code == M.MKDE_E_ARG
    # the previous 100 points are still the context's, with their bandwidth
    assert ctx.n == 100
    assert len(ctx.get_bandwidths()) == 100
    assert np.array_equal(ctx.DensityAtPoints(), a)
    with pytest.raises(ValueError):
        ctx.set_bandwidths(np.ones(99, np.float32))
