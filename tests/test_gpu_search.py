"""Batched bandwidth search: leave-one-out log-likelihood of 64 log-spaced
bandwidths (0.1x - 10x the rule of thumb) over a clustered mixture with edge
correction, all candidates in batched passes.  The selected bandwidth (argmax)
must equal the oracle's; the log-likelihoods agree with the FP64 oracle."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu


def candidates(h_ref, K=64):
    return (np.float32(h_ref) * (10.0 ** (-1.0 + 2.0 * np.arange(K) / (K - 1)))).astype(np.float32)


def test_search_argmax_matches_oracle(ctx, orc):
    extent = 160.0
    x, y = cases.mixture_points(1200, extent, n_centres=16, seed=20260303)
    w = np.ones(len(x), np.float32)
    g = orc.grid(160, 160, 0.0, 0.0, 1.0)
    mask = cases.all_valid(160, 160)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(M.grid(160, 160, 0.0, 0.0, 1.0), mask)
    h_ref = ctx.BandWidth_rule_of_thumb()
    hc = candidates(h_ref)
    got = ctx.LogLikelihood(hc, edge_correction=True)
    flags, _ = orc.mark_boundary(g, mask)
    d = orc.dist_to_boundary(x, y, g, flags)
    ref64 = np.empty(len(hc))
    ref32 = np.empty(len(hc))
    for k, h in enumerate(hc):
        hs = np.full(len(x), h, np.float32)
        ew32, ew64, _ = orc.ec_weights(x, y, d, hs, g, mask)
        o = orc.loo_density(x, y, w, ew64.astype(np.float32), hs)
        ref64[k] = orc.loglik_fp64(o["den1_64"])
        o32 = orc.loo_density(x, y, w, ew32, hs)
        ref32[k] = orc.loglik_ref32(o32["den1"])
    n = len(x)
    assert np.max(np.abs(got - ref64)) <= 1e-6 * n
    assert int(np.argmax(got)) == int(np.argmax(ref64))
    s = np.sort(ref64)
    margin = s[-1] - s[-2]
    # the reference sums n float logs: its argmax agrees with the FP64 one
    # whenever the top-2 margin exceeds that sum's worst-case rounding error,
    # about n * 2^-24 * sum |log den_i| (>= n * 2^-24 * |logL|)
    bound = n * 2.0 ** -24 * np.max(np.abs(ref64))
    if margin > 4 * bound:
        assert int(np.argmax(ref32)) == int(np.argmax(ref64))


@pytest.mark.parametrize("edge_correction", [True, False])
def test_full_pass_matches_shorter_pass(ctx, edge_correction):
    # A 16-candidate pass stops each source's candidate walk on the wave
    # minimum of the screened distance; a 15-candidate pass uses a wave vote
    # per candidate.  Over the same 15 largest candidates the two passes share
    # the index (built for the largest bandwidth) and the work split, so their
    # common log-likelihoods are identical bit for bit.
    x, y = cases.mixture_points(1500, 160.0, n_centres=12, seed=20260412)
    w = np.linspace(0.5, 2.0, len(x)).astype(np.float32)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(M.grid(160, 160, 0.0, 0.0, 1.0), cases.all_valid(160, 160))
    hc = candidates(ctx.BandWidth_rule_of_thumb(), 16)
    full = ctx.LogLikelihood(hc, edge_correction=edge_correction)
    short = ctx.LogLikelihood(hc[1:], edge_correction=edge_correction)
    assert np.all(np.isfinite(full))
    assert np.array_equal(full[1:], short)
