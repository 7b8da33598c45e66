"""GPU parity of the valid-cell runs the library builds on the device from an
uploaded mask (k_run_count / k_run_fill): every row's runs take part in the
edge-correction integral, so the EC pair count and weights of points whose
cut-off covers the whole mask equal the oracle's (which scans the mask cell
by cell, EdgeCorrectionWeightsExact tiled:2694-2807) only when every run of
every row is right.  The masks put runs across the 64-column steps of the
kernels, at the first and last columns, one cell long, in rows with no valid
cell, and at the most runs a row can hold (alternating cells)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NODATA = -9999.0


def masks():
    r = np.random.default_rng(20260311)
    out = {}
    m = np.ones((37, 130), np.float32)
    m[:, 1::2] = NODATA                      # alternating: 65 one-cell runs per row
    m[5] = NODATA                            # a row without valid cells
    out["alternating_130"] = m
    m = np.ones((40, 64), np.float32)
    m[:, 63] = NODATA
    m[::3, 0] = NODATA
    m[7, :] = NODATA
    m[9, 1:63] = NODATA                      # two one-cell runs at the row's ends
    out["one_step_64"] = m
    m = np.where(r.random((50, 193)) < 0.5, 1.0, NODATA).astype(np.float32)
    m[:, 62:66] = 1.0                        # runs across the 64-column step
    m[:, 127:129] = NODATA
    out["random_193"] = m
    m = np.ones((30, 1), np.float32)
    m[::2] = NODATA
    out["one_column"] = m
    m = np.ones((20, 65), np.float32)
    m[:, 64] = NODATA
    m[3, 64] = 1.0
    out["last_column_65"] = m
    return out


@pytest.mark.parametrize("name", sorted(masks()))
def test_runs_from_device_match_oracle_ec(ctx, orc, name):
    import mgpukde_tiled_amd as M

    mask = masks()[name]
    nr, nc = mask.shape
    g = orc.grid(nc, nr, 10.0, 20.0, 1.0, NODATA)
    r = np.random.default_rng(sum(map(ord, name)))
    n = 40
    x = (10.0 + r.uniform(0, nc, n)).astype(np.float32)
    y = (20.0 + r.uniform(0, nr, n)).astype(np.float32)
    # cut-offs cover the whole mask for most points, a few cover part of it
    hs = np.where(np.arange(n) % 4 == 0, 2.5, float(max(nc, nr))).astype(np.float32)
    ctx.points_upload(x, y, None)
    ctx.mask_upload(M.grid(nc, nr, 10.0, 20.0, 1.0, NODATA), mask)
    ctx.set_bandwidths(hs)
    ew, ne = ctx.EdgeCorrectionWeightsExact()
    oflags, _ = orc.mark_boundary(g, mask)
    od = orc.dist_to_boundary(x, y, g, oflags)
    _, ew64, one = orc.ec_weights(x, y, od, hs, g, mask)
    assert ne == one, "EC pair count %d, oracle %d" % (ne, one)
    e = np.abs(ew.astype(np.float64) - ew64) / np.abs(ew64)
    assert e.max() <= 1e-5
