"""GPU parity of the bandwidth-selection layer and the whole-job driver.

The oracle optimisers (restated control flow of MLE_FixedBandWidth and
hooke_jeeves) are driven by the FP64 oracle log-likelihood; the GPU versions by
the GPU FP64 log-likelihood.  The selected bandwidths must be identical.
"""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REL = 1e-5


@pytest.fixture(scope="module")
def small(orc):
    x, y, w = cases.thomas_points(5, 120, 6.0, 80.0, seed=20260301)
    g = orc.grid(80, 80, 0.0, 0.0, 1.0)
    mask = cases.disk_mask(80, 80, 36, n_holes=2, hole_min=5, hole_max=12, seed=20260301)
    flags, _ = orc.mark_boundary(g, mask)
    dist = orc.dist_to_boundary(x, y, g, flags)
    return dict(x=x, y=y, w=w, g=g, mask=mask, dist=dist)


def mgrid(M, g):
    return M.grid(g.ncols, g.nrows, g.xll, g.yll, g.cs, g.nodata)


def oracle_loglik_fixed(orc, s, h, ec=True):
    n = len(s["x"])
    hs = np.full(n, h, np.float32)
    if ec:
        _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    else:
        ew64 = np.ones(n)
    o = orc.loo_density(s["x"], s["y"], s["w"], ew64.astype(np.float32), hs)
    return orc.loglik_fp64(o["den1_64"])


def oracle_adaptive_h(orc, s, h, alpha, ec=True):
    n = len(s["x"])
    h = np.float32(h)
    if ec:
        _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], np.full(n, h, np.float32), s["g"], s["mask"])
    else:
        ew64 = np.ones(n)
    _, p64 = orc.pilot_density(s["x"], s["y"], s["w"], ew64.astype(np.float32), h)
    g64 = orc.compgml64(p64)
    return orc.abramson64(float(h), float(alpha), g64, p64).astype(np.float32)


def oracle_loglik_adaptive(orc, s, h, alpha, ec=True):
    hs = oracle_adaptive_h(orc, s, h, alpha, ec)
    if ec:
        _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    else:
        ew64 = np.ones(len(hs))
    o = orc.loo_density(s["x"], s["y"], s["w"], ew64.astype(np.float32), hs)
    return orc.loglik_fp64(o["den1_64"])


def test_adaptive_loglik_values(ctx, orc, small):
    import mgpukde_tiled_amd as M

    s = small
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    hh = np.array([3.0, 4.5, 6.0], np.float32)
    aa = np.array([-0.5, -0.3, -0.7], np.float32)
    got = ctx.LogLikelihoodAdaptive(hh, aa, edge_correction=True)
    n = len(s["x"])
    for k in range(3):
        ref = oracle_loglik_adaptive(orc, s, hh[k], aa[k])
        assert abs(got[k] - ref) <= 1e-5 * n, (k, got[k], ref)


def test_mle_fixed_bandwidth_identical(ctx, orc, small):
    import mgpukde_tiled_amd as M

    s = small
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    h0 = ctx.BandWidth_rule_of_thumb()
    hopt, it = ctx.MLE_FixedBandWidth(h0, edge_correction=True)
    ohopt, oit, _ = orc.golden_section(h0, lambda h: oracle_loglik_fixed(orc, s, h))
    assert it == oit
    assert hopt == ohopt


def test_hooke_jeeves_identical(ctx, orc, small):
    import mgpukde_tiled_amd as M

    s = small
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.mask_upload(mgrid(M, s["g"]), s["mask"])
    h0 = ctx.BandWidth_rule_of_thumb()
    step_h = float(np.float32(h0) / np.float32(10))
    (h, a, L), it = ctx.hooke_jeeves(h0, -0.5, step_h, 0.1, edge_correction=True)
    (oh, oa, oL), oit = orc.hooke_jeeves(h0, -0.5, step_h, 0.1,
                                         lambda hh, aa: oracle_loglik_adaptive(orc, s, hh, aa))
    assert (h, a) == (oh, oa)
    assert it == oit
    assert abs(L - oL) <= 1e-5 * abs(oL)


@pytest.mark.parametrize("hoption", [0, 1, -1])
def test_run_job_fixed_options(ctx, orc, small, hoption):
    """The whole GPU block for h_option 0 (rule of thumb), 1 (MLE) and -1
    (bandwidths given): bandwidths, EC weights and the density raster."""
    import mgpukde_tiled_amd as M

    s = small
    n = len(s["x"])
    ctx.points_upload(s["x"], s["y"], s["w"])
    g = mgrid(M, s["g"])
    ctx.mask_upload(g, s["mask"])
    ctx.tile_upload(g, s["mask"])
    h0 = orc.rule_of_thumb(s["x"], s["y"], s["w"])
    if hoption == -1:
        hs = np.random.default_rng(3).uniform(2.0, 7.0, n).astype(np.float32)
        ctx.set_bandwidths(hs)
    st = ctx.run_job(hoption=hoption, edge_correction=True)
    assert st["h0"] == h0
    if hoption == 0:
        hs = np.full(n, h0, np.float32)
    elif hoption == 1:
        ohopt, _, _ = orc.golden_section(h0, lambda h: oracle_loglik_fixed(orc, s, h))
        assert st["h_opt"] == ohopt
        hs = np.full(n, ohopt, np.float32)
    assert np.array_equal(ctx.get_bandwidths(), hs)
    _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    ew = ctx.get_ec_weights()
    assert np.max(np.abs(ew - ew64) / ew64) <= REL
    den = ctx.tile_download()
    _, o64, _ = orc.density_tile(s["x"], s["y"], s["w"], ew, hs, s["g"], s["mask"])
    v = s["mask"] != cases.NODATA
    assert np.all(den[~v] == cases.NODATA)
    ref = o64[v]
    got = den[v]
    assert np.all(got[ref == 0] == 0)
    pos = ref > 0
    assert np.max(np.abs(got[pos] - ref[pos]) / ref[pos]) <= REL


def test_run_job_adaptive_fixed_alpha(ctx, orc, small):
    """h_option 2 without the optimiser: Abramson at the rule-of-thumb h0."""
    import mgpukde_tiled_amd as M

    s = small
    ctx.points_upload(s["x"], s["y"], s["w"])
    g = mgrid(M, s["g"])
    ctx.mask_upload(g, s["mask"])
    ctx.tile_upload(g, s["mask"])
    st = ctx.run_job(hoption=2, edge_correction=True, optimise=False, alpha=-0.5)
    hs = ctx.get_bandwidths()
    href = oracle_adaptive_h(orc, s, st["h0"], -0.5)
    assert np.max(np.abs(hs - href) / href) <= REL
    ew = ctx.get_ec_weights()
    _, ew64, _ = orc.ec_weights(s["x"], s["y"], s["dist"], hs, s["g"], s["mask"])
    assert np.max(np.abs(ew - ew64) / ew64) <= REL
    den = ctx.tile_download()
    _, o64, _ = orc.density_tile(s["x"], s["y"], s["w"], ew, hs, s["g"], s["mask"])
    v = (s["mask"] != cases.NODATA) & (o64 > 0)
    assert np.max(np.abs(den[v] - o64[v]) / o64[v]) <= REL


def test_density_stats_counters(ctx, orc, small):
    import mgpukde_tiled_amd as M

    s = small
    ctx.points_upload(s["x"], s["y"], s["w"])
    ctx.set_bandwidth(9.0)
    ctx.set_ec_weights(None)
    ctx.density_stats(True)
    tile = M.grid(80, 80, 0.0, 0.0, 1.0)
    a = ctx.KernelDensityEstimation_tile(tile, s["mask"])
    st = ctx.density_stats(False)
    b = ctx.KernelDensityEstimation_tile(tile, s["mask"])
    assert np.array_equal(a, b)  # the counting variant computes the same raster
    assert st["candidates"] >= st["full_blocks"] + st["partial_blocks"] > 0
