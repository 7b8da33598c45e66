"""The autotuner: it times the resident job (or search) under each value of
the launch-shape options, leaves the chosen values set, records them for the
job class, and the tuned job still agrees with the FP64 oracle (a tuning moves
FP32 rounding at most, never the cut-off sets)."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu
REL = 1e-5


@pytest.fixture(scope="module")
def job(orc):
    x, y, w = cases.thomas_points(8, 300, 9.0, 160.0, seed=20260511)
    g = orc.grid(160, 160, 0.0, 0.0, 1.0)
    mask = cases.disk_mask(160, 160, 74, n_holes=3, hole_min=6, hole_max=20, seed=20260511)
    return dict(x=x, y=y, w=w, g=g, mask=mask)


def mgrid(g):
    return M.grid(g.ncols, g.nrows, g.xll, g.yll, g.cs, g.nodata)


def test_autotune_job_chooses_records_and_keeps_parity(orc, job):
    s = job
    g = mgrid(s["g"])
    with M.Context(0) as c:
        c.points_upload(s["x"], s["y"], s["w"])
        c.mask_upload(g, s["mask"])
        c.tile_upload(g, s["mask"])
        rep = c.autotune_job(hoption=2, edge_correction=True, optimise=False, alpha=-0.5, reps=2)
        assert rep["workload"] == "job" and rep["final_ms"] <= rep["baseline_ms"] * 1.0 + 1e-9
        names = {t["option"] for t in rep["trials"]}
        assert {"density_split_target", "pts_split_target", "ec_chunk", "points_per_bin"} <= names
        for k, v in rep["chosen"].items():
            assert c.get_option(k) == v
        c.run_job(hoption=2, edge_correction=True, optimise=False, alpha=-0.5)
        hs = c.get_bandwidths()
        ew = c.get_ec_weights()
        den = c.tile_download()
    _, o64, _ = orc.density_tile(s["x"], s["y"], s["w"], ew, hs, s["g"], s["mask"])
    v = (s["mask"] != cases.NODATA) & (o64 > 0)
    assert np.max(np.abs(den[v] - o64[v]) / o64[v]) <= REL
    # the choice is recorded for the job class and applied to a new context
    with M.Context(0) as c2:
        assert c2.autotune_lookup(len(s["x"]), 160 * 160, 2, True, False)
        for k, v in rep["chosen"].items():
            assert c2.get_option(k) == v
        assert not c2.autotune_lookup(len(s["x"]) * 64, 160 * 160, 2, True, False)


def test_autotune_search_keeps_the_selected_bandwidth(job):
    s = job
    with M.Context(0) as c:
        c.points_upload(s["x"], s["y"], s["w"])
        c.mask_upload(mgrid(s["g"]), s["mask"])
        h0 = c.BandWidth_rule_of_thumb()
        hc = (np.float32(h0) * (10.0 ** (-1.0 + 2.0 * np.arange(32) / 31.0))).astype(np.float32)
        before = c.LogLikelihood(hc, edge_correction=True)
        rep = c.autotune_search(hc, edge_correction=True, reps=2)
        assert rep["workload"] == "search"
        after = c.LogLikelihood(hc, edge_correction=True)
    n = len(s["x"])
    assert np.max(np.abs(after - before)) <= 1e-6 * n
    assert int(np.argmax(after)) == int(np.argmax(before))


def test_options_reject_bad_values():
    with M.Context(0) as c:
        with pytest.raises(M.MkdeError):
            c.set_option("ec_chunk", 700)
        with pytest.raises(M.MkdeError):
            c.set_option("no_such_option", 1)
        c.set_option("density_split_target", 8192)
        assert c.get_option("density_split_target") == 8192
