"""The bandwidth search as one captured HIP graph (option "search_graph")
against the host-sized launches: identical log-likelihoods bit for bit, on
every replay of the graph, with other work of the context between replays,
for several candidate sets of one K (the graph is reused with new candidates),
for K that is not a multiple of 16, and with edge correction off."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu


def candidates(h_ref, K):
    return (np.float32(h_ref) * (10.0 ** (-1.0 + 2.0 * np.arange(K) / max(K - 1, 1)))).astype(np.float32)


@pytest.fixture(scope="module")
def scene():
    side = 300
    x, y = cases.mixture_points(6000, float(side), n_centres=16, seed=20260411)
    mask = cases.disk_mask(side, side, 140, n_holes=4, hole_min=8, hole_max=40, seed=20260411)
    return x, y, mask, M.grid(side, side, 0.0, 0.0, 1.0)


def eager_and_graph(ctx, hc, ec=True):
    ctx.set_option("search_graph", 0)
    e = ctx.LogLikelihood(hc, edge_correction=ec)
    ctx.set_option("search_graph", 1)
    g = ctx.LogLikelihood(hc, edge_correction=ec)
    return e, g


def test_graph_replays_equal_eager(ctx, scene):
    x, y, mask, g = scene
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    sets = [candidates(h0, 64), candidates(h0 * 1.13, 64), candidates(h0 * 0.87, 64)]
    ctx.set_option("search_graph", 0)
    eager = [ctx.LogLikelihood(hc, edge_correction=True) for hc in sets]
    ctx.set_option("search_graph", 1)
    for rep in range(3):
        for i, hc in enumerate(sets):
            got = ctx.LogLikelihood(hc, edge_correction=True)
            assert np.array_equal(got, eager[i]), (rep, i)
            # other work of the context between the replays
            ctx.set_bandwidth(float(h0))
            ctx.EdgeCorrectionWeightsExact()
            ctx.DensityAtPoints(loo=True)
    assert ctx.get_option("search_graph") == 1


@pytest.mark.parametrize("K", [1, 5, 16, 21])
def test_graph_partial_passes_equal_eager(ctx, scene, K):
    x, y, mask, g = scene
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    hc = candidates(h0, K)
    e, gr = eager_and_graph(ctx, hc)
    assert np.array_equal(e, gr)
    assert np.array_equal(ctx.LogLikelihood(hc, edge_correction=True), e)   # a second replay


def test_graph_without_edge_correction(ctx, scene):
    x, y, mask, g = scene
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    hc = candidates(h0, 32)
    e, gr = eager_and_graph(ctx, hc, ec=False)
    assert np.array_equal(e, gr)


def test_graph_follows_new_points(ctx, scene):
    # a new upload of the same size must not replay a graph of the old points
    x, y, mask, g = scene
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    hc = candidates(h0, 16)
    ctx.set_option("search_graph", 1)
    a = ctx.LogLikelihood(hc, edge_correction=True)
    x2 = (x[::-1] * np.float32(0.9)).astype(np.float32)
    ctx.points_upload(x2, y, None)
    b = ctx.LogLikelihood(hc, edge_correction=True)
    ctx.set_option("search_graph", 0)
    b_eager = ctx.LogLikelihood(hc, edge_correction=True)
    assert np.array_equal(b, b_eager)
    assert not np.array_equal(a, b)


def test_graph_follows_launch_shape_options(ctx, scene):
    # a changed split target or EC chunk recaptures the graph: each replay
    # equals the host-sized launches with the same options
    x, y, mask, g = scene
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    hc = candidates(h0, 24)
    for opt, val in (("loo_split_target", 8192), ("ec_chunk", 512), ("ec_wide_min", 2048)):
        keep = ctx.get_option(opt)
        ctx.set_option(opt, val)
        e, gr = eager_and_graph(ctx, hc)
        assert np.array_equal(e, gr), opt
        assert np.array_equal(ctx.LogLikelihood(hc, edge_correction=True), e), opt
        ctx.set_option(opt, keep)
