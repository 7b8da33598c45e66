"""The drop-in command line end to end on the GPU: the reference's mode 1
example (Redwood points + masks) and mode 0 (generated points), checked
against the oracle pipeline on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
RED_TIF = os.path.join(GOLD, "redwood_mask.tif")
RED_CSV = os.path.join(GOLD, "pntsRedwood.csv")
CLI = os.path.join(os.path.dirname(HERE), "mgpukde_tiled_amd", "mgpukde_tiled")
REL = 1e-5


def run(args, cwd):
    r = subprocess.run([CLI] + [str(a) for a in args], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def read_pnts(path):
    rows = [ln.split(",") for ln in open(path).read().strip().splitlines()[1:]]
    return np.array(rows, np.float64)


def oracle_fixed_pipeline(orc, x, y, w, g_ec, mask_ec, tile, mask_den):
    h0 = orc.rule_of_thumb(x, y, w)
    flags, _ = orc.mark_boundary(g_ec, mask_ec)
    d = orc.dist_to_boundary(x, y, g_ec, flags)
    hs = np.full(len(x), h0, np.float32)
    _, ew64, _ = orc.ec_weights(x, y, d, hs, g_ec, mask_ec)
    _, o64, _ = orc.density_tile(x, y, w, ew64.astype(np.float32), hs, tile, mask_den)
    return h0, ew64, d, o64


def test_cli_mode1_redwood_rule_of_thumb(orc, tmp_path):
    out = run([1, RED_CSV, RED_TIF, RED_TIF, 0, 1, 0, 1, 0, 12, 1, "cpu.tif", "gpu.tif", 0], tmp_path)
    assert "Rule of thumb bandwidth h0" in out
    den = M.tiff_read(str(tmp_path / "gpu.tif"))
    nc, nr, gt, nd, _ = M.tiff_info(str(tmp_path / "gpu.tif"))
    assert (nc, nr) == (128, 128) and nd == -9999.0
    assert np.allclose(gt, M.tiff_info(RED_TIF)[2], rtol=0, atol=0)
    x, y, _, w = M.csv_read_points(RED_CSV)
    mask = M.tiff_read(RED_TIF)
    g = orc.grid(128, 128, 0.0, float(np.float32(1.0 - 128 / 128)), 1 / 128)
    h0, ew64, d, o64 = oracle_fixed_pipeline(orc, x, y, w, g, mask, g, mask)
    pos = o64 > 0
    assert np.all(den[~pos] == 0)
    assert np.max(np.abs(den[pos] - o64[pos]) / o64[pos]) <= REL
    p = read_pnts(str(tmp_path / "pntsCUDA.csv"))
    assert len(p) == 62
    order = np.argsort(d, kind="stable")
    assert np.allclose(p[:, 0], x[order], atol=5e-7) and np.allclose(p[:, 1], y[order], atol=5e-7)
    assert np.allclose(p[:, 2], h0, atol=5e-7)
    assert np.allclose(p[:, 3], ew64[order], rtol=1e-5, atol=5e-7)


def test_cli_mode0_generated_points(orc, tmp_path):
    run([0, 100, 1, 0.1, 0, 1, 0, 1, 0, 12, 1, "cpu.tif", "gpu.tif", 0], tmp_path)
    den = M.tiff_read(str(tmp_path / "gpu.tif"))
    assert den.shape == (1000, 1000)
    nc, nr, gt, nd, _ = M.tiff_info(str(tmp_path / "gpu.tif"))
    assert np.allclose(gt, [0, np.float32(0.1), 0, 100, 0, -np.float32(0.1)], rtol=0, atol=0)
    x, y, _ = orc.msvc_rand_points(100, seed=100)
    w = np.ones(100, np.float32)
    n_ec, _, _, _, cs_ec = cases.mode0_grid(1.0)
    g_ec = orc.grid(n_ec, n_ec, 0.0, 0.0, cs_ec)
    mask_ec = np.zeros((n_ec, n_ec), np.float32)
    ncols, nrows, xll, yll, cs = cases.mode0_density_tile(0.1)
    tile = orc.grid(ncols, nrows, xll, yll, cs)
    h0, ew64, d, o64 = oracle_fixed_pipeline(orc, x, y, w, g_ec, mask_ec, tile, None)
    pos = o64 > 0
    assert np.all(den[~pos] == 0)
    assert np.max(np.abs(den[pos] - o64[pos]) / o64[pos]) <= REL
    p = read_pnts(str(tmp_path / "pntsCUDA.csv"))
    assert np.allclose(p[:, 2], h0, atol=5e-7)


def test_cli_mode1_file_bandwidths_and_mle(orc, tmp_path):
    # h_option -1: bandwidths from the CSV; then h_option 1 (MLE) runs through
    x, y, _, w = M.csv_read_points(RED_CSV)
    hs = np.linspace(0.05, 0.15, len(x)).astype(np.float32)
    pts = tmp_path / "pts_h.csv"
    M.csv_write_points(str(pts), x, y, hs, w)
    run([1, pts, RED_TIF, RED_TIF, -1, 1, 0, 1, 0, 12, 1, "cpu.tif", "gpu.tif", 0], tmp_path)
    den = M.tiff_read(str(tmp_path / "gpu.tif"))
    mask = M.tiff_read(RED_TIF)
    g = orc.grid(128, 128, 0.0, 0.0, 1 / 128)
    flags, _ = orc.mark_boundary(g, mask)
    d = orc.dist_to_boundary(x, y, g, flags)
    hs_rd = M.csv_read_points(str(pts), with_h=True)[2]
    _, ew64, _ = orc.ec_weights(x, y, d, hs_rd, g, mask)
    _, o64, _ = orc.density_tile(x, y, w, ew64.astype(np.float32), hs_rd, g, mask)
    pos = o64 > 0
    assert np.max(np.abs(den[pos] - o64[pos]) / o64[pos]) <= REL
    out = run([1, RED_CSV, RED_TIF, RED_TIF, 1, 1, 0, 1, 0, 12, 1, "cpu.tif", "gpu1.tif", 0], tmp_path)
    assert "cross validated optimal fixed bandwidth hopt" in out
    out = run([1, RED_CSV, RED_TIF, RED_TIF, 2, 1, 0, 1, 0, 12, 1, "cpu.tif", "gpu2.tif", 0], tmp_path)
    assert "alpha0" in out
    den2 = M.tiff_read(str(tmp_path / "gpu2.tif"))
    assert np.all(np.isfinite(den2)) and den2.sum() > 0
