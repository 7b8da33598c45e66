"""The density raster's cut-off on exact ties: points whose squared distance to
many cell centres equals 9 h^2 exactly, so the inside / outside decision of
those cells rests on the strict comparison fl64(dx2 + dy2) < 9 fl32(h^2)
(kde_cuda_kdtr_tiled.cu, density kernel cut-off).  The circles do not
overlap, so every cell's value comes from at most one point and the set of
cells with a nonzero density is exactly the set of cells inside a cut-off:
it must equal the oracle's cell for cell, on the PARTIAL blocks and wave
sub-blocks where the boundary crosses the block."""
import numpy as np
import pytest

import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu
REL = 1e-5


def tie_points(xll, yll, nc, nr, spacing=40):
    # centres on a lattice of cells, with sub-cell offsets; h^2 exact in FP32
    # and 9 h^2 a sum of two integer squares (225 = 15^2 = 9^2 + 12^2,
    # 81 = 9^2, 36 = 6^2, 100 = 10^2 = 6^2 + 8^2)
    hs_cycle = [5.0, 3.0, 2.0, 10.0 / 3.0]
    offs = [0.0, 0.25, 0.5, -0.5, 0.125]
    x, y, h = [], [], []
    k = 0
    for r in range(spacing // 2, nr - spacing // 2, spacing):
        for c in range(spacing // 2, nc - spacing // 2, spacing):
            ox = offs[k % len(offs)]
            oy = offs[(k // 2) % len(offs)]
            x.append(xll + c + 0.5 + ox)
            y.append(yll + (nr - 0.5 - r) + oy)
            h.append(hs_cycle[k % len(hs_cycle)])
            k += 1
    return np.array(x, np.float32), np.array(y, np.float32), np.array(h, np.float32)


@pytest.mark.parametrize("xll,yll", [(0.0, 0.0), (4096.0, 1024.0), (9.0e6, 3.0e6)])
def test_cutoff_ties_cell_for_cell(orc, ctx, xll, yll):
    nc, nr = 330, 250       # ragged edge blocks on both axes
    x, y, hs = tie_points(xll, yll, nc, nr)
    w = np.linspace(0.5, 2.0, len(x)).astype(np.float32)
    ones = np.ones(len(x), np.float32)
    ctx.points_upload(x, y, w)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ones)
    den = ctx.KernelDensityEstimation_tile(M.grid(nc, nr, xll, yll, 1.0), None)
    _, o64, n_eval = orc.density_tile(x, y, w, ones, hs, orc.grid(nc, nr, xll, yll, 1.0), None)
    inside = o64 > 0
    assert inside.sum() == n_eval          # one point per cell: the pair count
    assert np.array_equal(den > 0, inside)
    assert np.max(np.abs(den[inside] - o64[inside]) / o64[inside]) <= REL


def test_cutoff_ties_with_mask(orc, ctx):
    # the same ties on a masked tile: nodata cells keep the mask value, the
    # rest must match the oracle's inside set exactly
    nc, nr = 260, 200
    x, y, hs = tie_points(0.0, 0.0, nc, nr)
    ones = np.ones(len(x), np.float32)
    ctx.points_upload(x, y, None)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ones)
    mask = np.ones((nr, nc), np.float32)
    mask[::7, :] = -9999.0
    mask[:, ::11] = -9999.0
    den = ctx.KernelDensityEstimation_tile(M.grid(nc, nr, 0.0, 0.0, 1.0), mask)
    _, o64, _ = orc.density_tile(x, y, ones, ones, hs, orc.grid(nc, nr, 0.0, 0.0, 1.0), mask)
    valid = mask != -9999.0
    assert np.all(den[~valid] == -9999.0)
    inside = valid & (o64 > 0)
    assert np.array_equal(valid & (den > 0), inside)
    assert np.max(np.abs(den[inside] - o64[inside]) / o64[inside]) <= REL
