"""Seeded random configurations through the whole path against the oracle:
grid origin and cell size, mask shape (disk or rectangle with holes, or all
valid), clustered or uniform points (some outside the raster), weights, and a
fixed or per-point bandwidth from well below to well above the cell size.

Each case checks what the reference's path produces, at the bars of the other
parity tests: boundary flags and distances bit-exact, EC pair counts exact and
weights within 1e-5, density-raster pair counts exact and cells within 1e-5 of
the FP64 oracle, leave-one-out densities at points within 1e-5."""
import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu

REL = 1e-5
N_CASES = 40


def rel(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(got - ref) / np.where(ref != 0, np.abs(ref), 1.0)


def make_case(k):
    r = np.random.default_rng(20260500 + k)
    nc, nr = int(r.integers(40, 260)), int(r.integers(40, 260))
    cs = float(np.float32(r.choice([1.0, 0.5, 0.37, 2.0, 0.1])))
    xll = float(np.float32(r.uniform(-1000.0, 1000.0)))
    yll = float(np.float32(r.uniform(-1000.0, 1000.0)))
    kind = ["disk", "rect", "all"][k % 3]
    if kind == "disk":
        mask = cases.disk_mask(nc, nr, min(nc, nr) * r.uniform(0.3, 0.5), n_holes=int(r.integers(0, 4)),
                               hole_min=2, hole_max=max(3, min(nc, nr) // 6), seed=20260600 + k)
    elif kind == "rect":
        mask = cases.rect_mask(nc, nr, n_holes=int(r.integers(1, 5)), hole_min=2, hole_max=max(3, min(nc, nr) // 5),
                               seed=20260600 + k)
    else:
        mask = cases.all_valid(nc, nr)
    n = int(r.integers(50, 1500))
    ext_x, ext_y = nc * cs, nr * cs
    if k % 2:
        x = r.uniform(-0.05 * ext_x, 1.05 * ext_x, n)
        y = r.uniform(-0.05 * ext_y, 1.05 * ext_y, n)
    else:
        centres = r.uniform(0.1, 0.9, (5, 2)) * [ext_x, ext_y]
        pick = r.integers(0, 5, n)
        x = centres[pick, 0] + r.normal(0, 0.05 * ext_x, n)
        y = centres[pick, 1] + r.normal(0, 0.05 * ext_y, n)
    x = (xll + x).astype(np.float32)
    y = (yll + y).astype(np.float32)
    w = r.uniform(0.5, 2.0, n).astype(np.float32)
    lo, hi = cs * r.choice([0.8, 3.0, 10.0]), cs * r.choice([12.0, 40.0])
    adaptive = k % 4 >= 2
    hs = (r.uniform(lo, hi, n) if adaptive else np.full(n, r.uniform(lo, hi))).astype(np.float32)
    return dict(nc=nc, nr=nr, cs=cs, xll=xll, yll=yll, mask=mask, x=x, y=y, w=w, hs=hs)


@pytest.mark.parametrize("k", range(N_CASES))
def test_random_configuration(ctx, orc, k):
    import mgpukde_tiled_amd as M

    c = make_case(k)
    g = orc.grid(c["nc"], c["nr"], c["xll"], c["yll"], c["cs"])
    mg = M.grid(c["nc"], c["nr"], c["xll"], c["yll"], c["cs"])
    x, y, w, hs, mask = c["x"], c["y"], c["w"], c["hs"], c["mask"]
    ctx.points_upload(x, y, w)
    ctx.mask_upload(mg, mask)

    flags, nb = ctx.MarkBoundary()
    oflags, onb = orc.mark_boundary(g, mask)
    assert nb == onb and np.array_equal(flags, oflags)
    d = ctx.CalcDist2Boundary()
    od = orc.dist_to_boundary(x, y, g, oflags)
    assert np.array_equal(d, od), "boundary distances must be bit-exact"

    ctx.set_bandwidths(hs)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    _, ew64, one = orc.ec_weights(x, y, od, hs, g, mask)
    assert ne == one, "EC pair count %d, oracle %d" % (ne, one)
    assert rel(ew, ew64).max() <= REL

    ew32 = ew64.astype(np.float32)
    ctx.set_ec_weights(ew32)
    got = ctx.KernelDensityEstimation_tile(mg, mask)
    _, o64, nd = orc.density_tile(x, y, w, ew32, hs, g, mask)
    valid = mask != cases.NODATA
    assert np.all(got[~valid] == mask[~valid])
    zero = o64[valid] == 0
    assert np.all(got[valid][zero] == 0), "cells outside every cut-off must be exactly 0"
    assert rel(got[valid][~zero], o64[valid][~zero]).max(initial=0.0) <= REL
    assert ctx.count_density_pairs(mg, mask) == nd

    den = ctx.DensityAtPoints(loo=True)
    oden = orc.loo_density(x, y, w, ew32, hs)["den1_64"]
    pos = oden > 0
    assert np.all(np.asarray(den)[~pos] == 0)
    assert rel(np.asarray(den)[pos], oden[pos]).max(initial=0.0) <= REL


@pytest.mark.parametrize("k", range(0, N_CASES, 2))
def test_random_bandwidth_search(ctx, orc, k):
    # the batched leave-one-out log-likelihood of five candidates around the
    # case's bandwidth (EC recomputed per candidate, one pass, captured graph)
    # against the FP64 oracle's EC weights, densities and log sum
    import mgpukde_tiled_amd as M

    c = make_case(k)
    g = orc.grid(c["nc"], c["nr"], c["xll"], c["yll"], c["cs"])
    x, y, w, mask = c["x"], c["y"], c["w"], c["mask"]
    ctx.points_upload(x, y, w)
    ctx.mask_upload(M.grid(c["nc"], c["nr"], c["xll"], c["yll"], c["cs"]), mask)
    h = float(np.median(c["hs"]))
    hc = (h * np.array([0.5, 0.8, 1.0, 1.3, 2.0])).astype(np.float32)
    got = ctx.LogLikelihood(hc, edge_correction=True)
    flags, _ = orc.mark_boundary(g, mask)
    d = orc.dist_to_boundary(x, y, g, flags)
    ref = np.empty(len(hc))
    for j, hj in enumerate(hc):
        hs = np.full(len(x), hj, np.float32)
        _, ew64, _ = orc.ec_weights(x, y, d, hs, g, mask)
        ref[j] = orc.loglik_fp64(orc.loo_density(x, y, w, ew64.astype(np.float32), hs)["den1_64"])
    # each density within 2e-6 relative (FP32 sums folded every 32 terms)
    assert np.max(np.abs(got - ref)) <= 2e-6 * len(x)
