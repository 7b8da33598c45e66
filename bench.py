#!/usr/bin/env python3
"""Benchmark: 1M-point adaptive, edge-corrected KDE on an 8192x8192 raster.

Workload (thomas_adaptive): 200 parents uniform in [0,8192)^2, Poisson(5000)
children each with N(0, 40^2) offsets, clipped; weights U[0.5, 2].  Mask: disk
of radius 3900 cells minus 20 rectangular holes of 50-400 cells, nodata -9999;
the same raster is the edge-correction mask and the density mask (cs = 1).
One step = one whole job on inputs resident in HBM: MarkBoundary +
CalcDist2Boundary, EC(h0), pilot density at the rule-of-thumb h0, geometric
mean, Abramson h_i = h0*(f_i/g)^-0.5, EC(h_i), the 8192^2 density raster.

value = Gaussian kernel evaluations of the job (pairs inside the cut-off, all
phases, counted exactly by separate untimed passes) / seconds per job.

Outside the timed steps the default run also reports:
  * the instruction rates of this GPU (exp, FP32 FMA, FP32 MFMA; mkde_rates)
    and the density kernel's counters from the committed profile;
  * full-size parity of the adaptive job against the FP64 oracle: boundary
    distances and EC weights on point samples, the geometric mean and the
    Abramson bandwidths of every point, raster tiles;
  * "search": the 64-candidate bandwidth search on 200,000 points (ms per
    search, the selected bandwidth, FP64-oracle log-likelihoods of the three
    best candidates, the top-2 margin against the FP32 error bound);
  * "tiles": 10M lognormal-cluster points, 32768^2 as 4096^2 tiles, h = 16,
    EC on (time, HBM GB/s, per-window error against the FP64 oracle);
  * "serving": the mixed-job stream at fixed fractions of its capacity;
  * "cpu_baseline": the reference CPU loop timed on this host's cores.
"""
import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

# MI355X dense FP32 peak: the FP32 MFMA rate (the same figure as the packed
# FP32 VALU rate); the density kernel runs its FULL cells on FP32 MFMA and its
# PARTIAL cells on the VALU
MI355X_FP32_TFLOPS = 157.3
MI355X_HBM_GBPS = 8000.0
FLOP_PER_EVAL = 8  # 2 sub + 2 mul/fma for d^2, 1 mul for the scale, 1 fma to accumulate
NODATA = -9999.0
METRIC = "Gaussian kernel evals/s + %VALU peak, 1M-pt adaptive KDE; p99 job latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def oracle_module():
    p = os.path.join(ROOT, "oracle")
    if p not in sys.path:
        sys.path.insert(0, p)
    import oracle as O

    return O


def workload(scale=1.0):
    import cases

    extent = 8192.0 * scale
    n_parents = max(1, int(round(200 * scale * scale)))
    x, y, w = cases.thomas_points(n_parents, 5000, 40.0, extent, seed=20260002)
    side = int(round(8192 * scale))
    mask = cases.disk_mask(side, side, 3900.0 * scale, n_holes=20, hole_min=max(2, int(50 * scale)),
                           hole_max=max(3, int(400 * scale)), seed=20260002)
    return x, y, w, mask, side


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return np.abs(got - ref) / np.abs(ref)


# ------------------------------------------------------------ roofline ----
def probe_rates():
    """exp / FP32 FMA / FP32 MFMA rates of this GPU (mkde_rates), or None."""
    exe = os.path.join(ROOT, "mgpukde_tiled_amd", "mkde_rates")
    if not os.path.exists(exe):
        return None
    try:
        r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        return json.loads(r.stdout.strip().splitlines()[-1]) if r.returncode == 0 else None
    except (subprocess.SubprocessError, ValueError, IndexError):
        return None


def profile_value(name, key):
    """The value of a `key: value` line of a committed profile summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            v = line.split(":", 1)[1].strip()
            try:
                return json.loads(v)
            except ValueError:
                return v
    return None


def roofline(e_den, ker_ms, rates):
    achieved = FLOP_PER_EVAL * e_den / (ker_ms * 1e-3) / 1e12
    prof = "k_density_tile_thomas_adaptive_1M_8192"
    traffic = profile_value(prof + "_traffic.md", "traffic_bytes_per_launch")
    counters = profile_value(prof + "_counters.md", "counters_json")
    r = {
        "bound": "mfma",
        "achieved": achieved,
        "peak": MI355X_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": achieved / MI355X_FP32_TFLOPS,
        "frac_kind": "FLOP-equivalent: %d FLOP per (point, valid cell) pair inside the cut-off, the work of the "
                     "direct form, over the kernel time; the kernel evaluates the Gaussian separably (two exps per "
                     "point per 64-cell row/column, one FMA per pair on the matrix cores or the VALU), so its "
                     "hardware utilisation is in 'counters'" % FLOP_PER_EVAL,
        "traffic": int(traffic) if traffic is not None else None,
        "kernel": "k_density_tile",
        "work_per_launch": "%d evaluations x %d FLOP" % (e_den, FLOP_PER_EVAL),
    }
    if traffic is not None:
        r["hbm_GBps"] = float(traffic) / (ker_ms * 1e-3) / 1e9
        r["hbm_frac"] = r["hbm_GBps"] / MI355X_HBM_GBPS
    if counters is not None:
        r["counters"] = counters
    if rates is not None:
        r["measured_rates"] = {k: rates[k] for k in ("exp_per_s", "fma_per_s", "fp32_valu_tflops",
                                                      "mfma_f32_32x32x2_tflops", "coissue_1to1", "coissue_1to4")
                               if k in rates}
        # the direct form needs one v_exp_f32 per evaluation; its floor on this
        # GPU from the measured exp rate alone
        r["direct_form_exp_floor_ms"] = e_den / rates["exp_per_s"] * 1e3
    return r


# ------------------------------------------------ full-size parity (C2) ----
def parity_adaptive(ctx, x, y, w, mask, side, h0, hs, ew, den, g_gpu, n_tiles=12, n_ec=256):
    """The adaptive job against the FP64 oracle at full size:
    * boundary distances (oracle brute force) and EC weights at the final h_i
      on a seeded sample of the points whose weight is corrected;
    * the EC(h0) weights of the pilot on a sample; the FP64 pilot density of
      every point (oracle, binned range search) with the job's EC(h0) weights,
      the geometric mean g and the Abramson bandwidth h_i of every point;
    * the raster on seeded 64x64 tiles (oracle over the points that can reach
      each tile, N11)."""
    O = oracle_module()
    n = len(x)
    g = O.grid(side, side, 0.0, 0.0, 1.0)
    rs = np.random.default_rng(20260098)
    t0 = time.perf_counter()
    flags, _ = O.mark_boundary(g, mask)
    d_gpu = ctx.CalcDist2Boundary()
    out = {}
    h2 = (hs * hs).astype(np.float64)   # fl32(h*h), exact in double
    near = np.flatnonzero(d_gpu < 9.0 * h2)
    far = np.flatnonzero(d_gpu >= 9.0 * h2)
    smp = np.concatenate([rs.choice(near, min(n_ec, len(near)), replace=False),
                          rs.choice(far, min(n_ec // 8, len(far)), replace=False)])
    d_or = O.dist_to_boundary(x[smp], y[smp], g, flags)
    out["dist_sample_bit_exact"] = bool(np.array_equal(d_or, d_gpu[smp]))
    _, ew64, _ = O.ec_weights(x[smp], y[smp], d_or, hs[smp], g, mask)
    e = rel_err(ew[smp], ew64)
    out["ec_weights"] = {"points": int(len(smp)), "max_rel_err": float(e.max()), "mean_rel_err": float(e.mean())}
    # the pilot: EC(h0) weights as the job computes them, checked on a sample
    hs0 = np.full(n, h0, np.float32)
    ctx.set_bandwidth(h0)
    ew0, _ = ctx.EdgeCorrectionWeightsExact()
    near0 = np.flatnonzero(d_gpu < 9.0 * float(np.float32(np.float32(h0) * np.float32(h0))))
    smp0 = rs.choice(near0, min(n_ec, len(near0)), replace=False)
    _, ew0_64, _ = O.ec_weights(x[smp0], y[smp0], d_gpu[smp0], hs0[smp0], g, mask)
    out["ec_weights_pilot"] = {"points": int(len(smp0)), "max_rel_err": float(rel_err(ew0[smp0], ew0_64).max())}
    den0_64, pairs = O.point_density_binned(x, y, w, ew0, hs0, loo=False)
    g64 = O.compgml64(den0_64)
    href = O.abramson64(float(h0), -0.5, g64, den0_64)
    eh = rel_err(hs, href)
    out["geometric_mean"] = {"gpu": g_gpu, "oracle_fp64": g64, "rel_err": abs(g_gpu - g64) / g64}
    out["bandwidths_h_i"] = {"points": n, "max_rel_err": float(eh.max()), "mean_rel_err": float(eh.mean()),
                             "pilot_pairs": int(pairs)}
    # raster tiles, seeded, centred on random points (where the density is)
    mx, acc, cnt, per = 0.0, 0.0, 0, []
    for _ in range(n_tiles):
        k = int(rs.integers(0, n))
        c0 = int(min(max(int(x[k]) - 32, 0), side - 64))
        r0 = int(min(max(side - 1 - int(y[k]) - 32, 0), side - 64))
        tile = O.grid(64, 64, float(c0), float(side - r0 - 64), 1.0)
        tm = np.ascontiguousarray(mask[r0:r0 + 64, c0:c0 + 64])
        sel = O.tile_point_filter(x, y, hs, tile)
        _, o64, _ = O.density_tile(x[sel], y[sel], w[sel], ew[sel], hs[sel], tile, tm, want64=True)
        gt = den[r0:r0 + 64, c0:c0 + 64]
        v = (tm != NODATA) & (o64 > 0)
        if np.any(v):
            er = rel_err(gt[v], o64[v])
            mx = max(mx, float(er.max()))
            acc += float(er.sum())
            cnt += int(v.sum())
            per.append(float(er.max()))
    out["raster"] = {"tiles": n_tiles, "cells_checked": cnt, "max_rel_err": mx, "mean_rel_err": acc / max(cnt, 1),
                     "max_rel_err_per_tile": per}
    out["max_rel_err"] = max(mx, out["ec_weights"]["max_rel_err"], out["ec_weights_pilot"]["max_rel_err"],
                             out["bandwidths_h_i"]["max_rel_err"])
    out["oracle_s"] = time.perf_counter() - t0
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    return out, ew0, d_gpu


# -------------------------------------------------------- CPU baseline ----
def cpu_timing_oracle():
    """A second instance of the oracle module bound to a build of the same
    source made on this host with -O3 -march=native (the CPU timing build);
    the parity checks keep the portable -O2 build."""
    import importlib.util

    spec = importlib.util.spec_from_file_location("oracle_timing", os.path.join(ROOT, "oracle", "oracle.py"))
    O = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(O)
    d = tempfile.mkdtemp(prefix="mkde_cpu_")
    try:
        O.use_library(O.build_native(d))
        return O, "oracle/mkde_oracle.c built on this host with gcc -O3 -march=native -fopenmp"
    except (OSError, subprocess.SubprocessError):
        O.lib()
        return O, "oracle/libmkde_oracle.so (-O2)"


def cpu_baseline_adaptive(O, build, x, y, w, ew, hs, h0, ew0, dist, mask, side, valid_cells, target_s=12.0):
    """The reference CPU path of the adaptive job on this host: the density
    loop (brute force over all points per valid cell, float accumulation,
    tiled:707-733) timed on seeded 64x64 tiles and reported as evaluations/s;
    the whole job's CPU time extrapolated from timed samples of each phase:
    the density loop per valid cell, the O(N^2) pilot per query point
    (ComputeFixedDensityAtPoints tiled:3096-3134) and the EC box loop per
    corrected point at h0 and at h_i (EdgeCorrectionWeightsExact
    tiled:2694-2807)."""
    rs = np.random.default_rng(20260099)
    n = len(x)
    g = O.grid(side, side, 0.0, 0.0, 1.0)
    times, e_total, cells, tiles = 0.0, 0, 0, 0
    while times < target_s and tiles < 64:
        r0 = int(rs.integers(0, side // 64)) * 64
        c0 = int(rs.integers(0, side // 64)) * 64
        tile = O.grid(64, 64, float(c0), float(side - r0 - 64), 1.0)
        tm = np.ascontiguousarray(mask[r0:r0 + 64, c0:c0 + 64])
        t0 = time.perf_counter()
        _, _, ne = O.density_tile(x, y, w, ew, hs, tile, tm, want64=False)
        times += time.perf_counter() - t0
        e_total += ne
        cells += int(np.count_nonzero(tm != NODATA))
        tiles += 1
    density_s = times / max(cells, 1) * valid_cells
    q = rs.choice(n, 64, replace=False)
    t0 = time.perf_counter()
    O.pilot_density_queries(x, y, w, ew0, h0, q)
    pilot_s = (time.perf_counter() - t0) / len(q) * n

    def ec_phase(h_of):
        h2 = (h_of * h_of).astype(np.float64)
        corr = np.flatnonzero(dist < 9.0 * h2)
        if len(corr) == 0:
            return 0.0, 0
        s = rs.choice(corr, min(256, len(corr)), replace=False)
        t = time.perf_counter()
        O.ec_weights(x[s], y[s], dist[s], h_of[s], g, mask)
        return (time.perf_counter() - t) / len(s) * len(corr), len(corr)

    ec0_s, n_ec0 = ec_phase(np.full(n, h0, np.float32))
    ec_s, n_ec = ec_phase(hs)
    total = density_s + pilot_s + ec0_s + ec_s
    return {
        "value": e_total / times if times > 0 else None,
        "unit": "evals/s",
        "cores": O.num_threads(),
        "kind": "port",
        "build": build,
        "sample": "%d seeded 64x64 tiles of the 8192^2 raster (%d valid cells, %d evaluations), all %d points, "
                  "reference CPU loop (brute force over all points per cell, float accumulation), %.1f s"
                  % (tiles, cells, e_total, n, times),
        "cells_per_s": cells / times if times > 0 else None,
        "job_s_extrapolated": total,
        "job_phases_s_extrapolated": {"density": density_s, "pilot_O_N2": pilot_s, "ec_h0": ec0_s, "ec_h_i": ec_s},
        "extrapolation": "density from %d tiles per valid cell (%d valid cells); pilot from %d query points x all "
                         "%d sources; EC from 256 corrected points per pass (%d and %d corrected points)"
                         % (tiles, valid_cells, len(q), n, n_ec0, n_ec),
    }


def cpu_plain_small(O):
    """The plain small workload on the CPU in full: EC weights (100 points on
    the 100x100 mask) and the 1000x1000 density loop, reference float path."""
    import cases

    x, y = cases.uniform_points(100, 100.0, seed=20260001)
    w = np.ones(100, np.float32)
    hs = np.full(100, cases.plain_small_bandwidth(x, y), np.float32)
    g_ec = O.grid(100, 100, 0.0, 0.0, 1.0)
    mask_ec = cases.all_valid(100, 100, 0.0)
    nc, nr, xll, yll, cs = cases.mode0_density_tile(0.1)
    tile = O.grid(nc, nr, xll, yll, cs)
    t0 = time.perf_counter()
    flags, _ = O.mark_boundary(g_ec, mask_ec)
    d = O.dist_to_boundary(x, y, g_ec, flags)
    ew32, _, e_ec = O.ec_weights(x, y, d, hs, g_ec, mask_ec)
    _, _, e_den = O.density_tile(x, y, w, ew32, hs, tile, None, want64=False)
    t = time.perf_counter() - t0
    return {"s": t, "E_density": int(e_den), "E_ec": int(e_ec), "evals_per_s": (e_den + e_ec) / t,
            "cells_per_s": nc * nr / t}


# ------------------------------------------------------------- search ----
def search_candidates(h_ref, K=64):
    return (np.float32(h_ref) * (10.0 ** (-1.0 + 2.0 * np.arange(K) / (K - 1)))).astype(np.float32)


def leg_search(args, with_oracle=True, O_cpu=None):
    """Batched bandwidth search: 200,000 points (50% uniform, 50% around 64
    Gaussian centres with sigma 2% of the extent) on an all-valid 4096^2 mask;
    LOO log-likelihood of 64 log-spaced bandwidths 0.1x-10x the rule of thumb,
    EC recomputed per candidate.  The three best candidates are recomputed by
    the FP64 oracle (EC weights of every point, binned range search)."""
    import cases
    import mgpukde_tiled_amd as M

    side = int(round(4096 * args.scale))
    n = int(round(200000 * args.scale * args.scale))
    x, y = cases.mixture_points(n, float(side), n_centres=64, seed=20260003)
    w = np.ones(n, np.float32)
    mask = cases.all_valid(side, side)
    res = {"workload": "search: %d mixture points, %d^2 all-valid mask, 64 candidates 0.1x-10x the rule of thumb, "
                       "EC per candidate" % (n, side)}
    with M.Context(0) as ctx:
        ctx.points_upload(x, y, None)
        ctx.mask_upload(M.grid(side, side, 0.0, 0.0, 1.0), mask)
        h0 = ctx.BandWidth_rule_of_thumb()
        hc = search_candidates(h0)
        def timed(reps=3):
            ctx.LogLikelihood(hc, edge_correction=True)
            tt = []
            for _ in range(reps):
                t0 = time.perf_counter()
                LL = ctx.LogLikelihood(hc, edge_correction=True)
                tt.append(time.perf_counter() - t0)
            return LL, 1e3 * float(np.median(tt))

        # the host-sized launches (a host wait per candidate and per pass),
        # then the search as one captured graph with the default and with the
        # autotuned launch shapes
        ctx.set_option("search_graph", 0)
        L_host, ms_host = timed()
        ctx.set_option("search_graph", 1)
        L, ms_default = timed()
        res["graph_identical_to_host_sized"] = bool(np.array_equal(L, L_host))
        res["ms_per_search_host_sized"] = ms_host
        res["ms_per_search_default_shapes"] = ms_default
        if not args.no_tune:
            res["autotune"] = ctx.autotune_search(hc, edge_correction=True, reps=3)
        L, ms_tuned = timed()
        t = [ms_tuned * 1e-3]
        k = int(np.argmax(L))
        e_loo = e_ec = 0
        for hk in hc:
            ctx.set_bandwidth(float(hk))
            e_loo += range_total(ctx) - n
            e_ec += ctx.EdgeCorrectionWeightsExact()[1]
        d = ctx.CalcDist2Boundary()
    ms = 1e3 * float(np.median(t))
    res.update({"ms_per_search": ms, "h_rule_of_thumb": h0, "argmax": k, "h_selected": float(hc[k]),
                "E_loo": e_loo, "E_ec": e_ec, "evals_per_s": (e_loo + e_ec) / (ms * 1e-3),
                "logL": [float(v) for v in L]})
    srt = np.sort(L)
    res["top2_margin"] = float(srt[-1] - srt[-2])
    # each density at a point is an FP32 sum folded every 32 terms (relative
    # error < 32 * 2^-24 per fold, 2e-6), so |dlogL| < n * 2e-6
    res["logL_fp32_error_bound"] = 2e-6 * n
    if with_oracle:
        O = oracle_module()
        O.lib()
        g = O.grid(side, side, 0.0, 0.0, 1.0)
        t0 = time.perf_counter()
        top = [int(i) for i in np.argsort(-np.asarray(L))[:3]]
        ref = {}
        for kk in top:
            hs = np.full(n, hc[kk], np.float32)
            _, ew64, _ = O.ec_weights(x, y, d, hs, g, mask)
            den64, _ = O.point_density_binned(x, y, w, ew64.astype(np.float32), hs, loo=True)
            ref[kk] = O.loglik_fp64(den64)
        best_or = max(ref, key=lambda i: ref[i])
        res["oracle_top3"] = [{"k": kk, "h": float(hc[kk]), "logL_gpu": float(L[kk]), "logL_oracle_fp64": ref[kk],
                               "abs_err": abs(float(L[kk]) - ref[kk]), "rel_err": abs(float(L[kk]) - ref[kk]) /
                               abs(ref[kk])} for kk in top]
        res["argmax_equal_oracle"] = bool(best_or == k)
        res["oracle_top2_margin"] = float(sorted(ref.values())[-1] - sorted(ref.values())[-2])
        res["max_rel_err"] = max(r["rel_err"] for r in res["oracle_top3"])
        res["oracle_s"] = time.perf_counter() - t0
    if O_cpu is not None:
        # the reference CPU LOO at the selected bandwidth: range search (binned,
        # the same candidate sets as the k-d tree) + scatter, timed on a sample
        # of queries, per candidate
        rs = np.random.default_rng(20260097)
        q = rs.choice(n, 2000, replace=False)
        hs = np.full(n, hc[k], np.float32)
        t0 = time.perf_counter()
        O_cpu.point_density_binned(x, y, w, np.ones(n, np.float32), hs, loo=True, queries=q)
        t_q = (time.perf_counter() - t0) / len(q) * n
        res["cpu_loo_per_candidate_s_extrapolated"] = t_q
    return res


def range_total(ctx):
    import ctypes

    import mgpukde_tiled_amd as M

    tot = ctypes.c_int64()
    M._check(M._lib.mkde_range_candidates(ctx._h, None, None, 0, ctypes.byref(tot)))
    return tot.value


# -------------------------------------------------------------- tiles ----
def leg_tiles(args, with_oracle=True, stats=False, O_cpu=None):
    """Precision stress: 10M points with log-normal cluster intensities over
    1024 centres, a 32768^2 GeoTIFF-style raster at cellsize 1 processed as
    4096^2 tiles (each tile with its own origin) in one pass, fixed h = 16,
    EC on (all-valid mask: the raster edge).  Evaluations/s, HBM GB/s of the
    raster write, and the max / mean relative error against the FP64 oracle on
    one 128^2 window in each of 8 seeded tiles; EC weights of the points
    nearest the raster edge."""
    import cases
    import mgpukde_tiled_amd as M

    side = int(round(32768 * args.scale))
    tile = max(64, int(round(4096 * args.scale)))
    n = int(round(10_000_000 * args.scale * args.scale))
    h = 16.0
    x, y = cases.lognormal_clusters(n, float(side), 1024, 64.0, seed=20260004)
    w = np.ones(n, np.float32)
    gt = [0.0, 1.0, 0.0, float(side), 0.0, -1.0]
    res = {"workload": "tiles: %d lognormal-cluster points, %d^2 raster as %d^2 tiles, h = %g, EC on"
                       % (n, side, tile, h), "points": n, "raster": side, "tile": tile, "h": h}
    with M.Context(0) as ctx:
        ctx.points_upload(x, y, None)
        ctx.set_bandwidth(h)
        log("tiles: %d points, %d^2 raster, %d^2 tiles" % (n, side, tile))
        ctx.mask_upload(M.grid(side, side, 0.0, 0.0, 1.0), cases.all_valid(side, side, 1.0))
        ctx.EdgeCorrectionWeightsExact()
        t0 = time.perf_counter()
        ew, e_ec = ctx.EdgeCorrectionWeightsExact()
        ms_ec = 1e3 * (time.perf_counter() - t0)
        for _ in range(2):
            ctx.KernelDensityEstimation_raster(side, side, gt, NODATA, tile, tile, None, keep_on_device=True)
        t, ker = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            ctx.KernelDensityEstimation_raster(side, side, gt, NODATA, tile, tile, None, keep_on_device=True)
            t.append(time.perf_counter() - t0)
            ker.append(ctx.last_kernel_ms("density"))
        if stats:
            ctx.density_stats(True)
            ctx.KernelDensityEstimation_raster(side, side, gt, NODATA, tile, tile, None, keep_on_device=True)
            res["density_stats"] = ctx.density_stats(False)
        den = ctx.tile_download()
        d = ctx.CalcDist2Boundary()
        e_den = 0
        ntile = (side + tile - 1) // tile
        for ty in range(ntile):
            for tx in range(ntile):
                tc, tr = min(tile, side - tx * tile), min(tile, side - ty * tile)
                g = M.grid(tc, tr, np.float32(tx * tile), np.float32(side - ty * tile - tr), 1.0)
                e_den += ctx.count_density_pairs(g, None)
    kms = float(np.median(ker))
    res.update({"ms_per_raster": 1e3 * float(np.median(t)), "kernel_ms": kms, "ms_ec": ms_ec, "E_density": e_den,
                "E_ec": e_ec, "evals_per_s": e_den / (kms * 1e-3), "cells_per_s": float(side) * side / (kms * 1e-3),
                "raster_write_GBps": 4.0 * side * side / (kms * 1e-3) / 1e9})
    hs = np.full(n, h, np.float32)
    if with_oracle:
        O = oracle_module()
        O.lib()
        t0 = time.perf_counter()
        rs = np.random.default_rng(20260097)
        per_tile = []
        win = min(128, tile, side)
        for _ in range(8):
            k = int(rs.integers(0, n))
            c0 = int(min(max(int(x[k]) - win // 2, 0), side - win))
            r0 = int(min(max(side - 1 - int(y[k]) - win // 2, 0), side - win))
            tx, ty = c0 // tile, r0 // tile
            c0 = min(c0, tx * tile + min(tile, side - tx * tile) - win)
            r0 = min(r0, ty * tile + min(tile, side - ty * tile) - win)
            g = O.grid(win, win, float(c0), float(side - r0 - win), 1.0)
            sel = O.tile_point_filter(x, y, hs, g)
            _, o64, _ = O.density_tile(x[sel], y[sel], w[sel], ew[sel], hs[sel], g, None, want64=True)
            gwin = den[r0:r0 + win, c0:c0 + win].astype(np.float64)
            v = o64 > 0
            if np.any(v):
                e = rel_err(gwin[v], o64[v])
                per_tile.append({"tile": [tx, ty], "window": [c0, r0, win], "max_rel_err": float(e.max()),
                                 "mean_rel_err": float(e.mean()), "cells": int(v.sum()),
                                 "zero_cells_exact": bool(np.all(gwin[~v] == 0.0))})
        near = np.argsort(d)[:256]
        gm = O.grid(side, side, 0.0, 0.0, 1.0)
        _, ew64, _ = O.ec_weights(x[near], y[near], d[near], hs[near], gm, cases.all_valid(side, side, 1.0))
        res["accuracy"] = per_tile
        res["ec_weights_max_rel_err"] = float(rel_err(ew[near], ew64).max())
        res["max_rel_err"] = max([p["max_rel_err"] for p in per_tile] + [res["ec_weights_max_rel_err"]])
        res["oracle_s"] = time.perf_counter() - t0
    if O_cpu is not None:
        # the reference CPU loop per tile: the tile's points (N11 filter), brute
        # force per cell; timed on 64x64 windows of 4 tiles, extrapolated over
        # all tiles by cells x filtered points
        rs = np.random.default_rng(20260096)
        ntile = (side + tile - 1) // tile
        # points each tile's loop visits: those within 3h of the tile (N11)
        cnt = np.zeros((ntile, ntile), np.float64)
        r3 = 3.0 * h
        txl = np.clip(np.floor((x - r3) / tile), 0, ntile - 1).astype(np.int64)
        txh = np.clip(np.floor((x + r3) / tile), 0, ntile - 1).astype(np.int64)
        tyl = np.clip(np.floor((side - y - r3) / tile), 0, ntile - 1).astype(np.int64)
        tyh = np.clip(np.floor((side - y + r3) / tile), 0, ntile - 1).astype(np.int64)
        for ty_, tx_, keep in ((tyl, txl, None), (tyl, txh, txh != txl), (tyh, txl, tyh != tyl),
                               (tyh, txh, (tyh != tyl) & (txh != txl))):
            sel = slice(None) if keep is None else keep
            np.add.at(cnt, (ty_[sel], tx_[sel]), 1.0)
        tcs = np.array([min(tile, side - t * tile) for t in range(ntile)], np.float64)
        work = float((cnt * tcs[:, None] * tcs[None, :]).sum())
        tsum, wsum = 0.0, 0.0
        for _ in range(4):
            tx, ty = int(rs.integers(0, ntile)), int(rs.integers(0, ntile))
            xl, xu = tx * tile - r3, tx * tile + tcs[tx] + r3
            yu, yl = side - ty * tile + r3, side - ty * tile - tcs[ty] - r3
            m = (x >= xl) & (x <= xu) & (y >= yl) & (y <= yu)
            g = O_cpu.grid(64, 64, float(tx * tile), float(side - ty * tile - 64), 1.0)
            t0 = time.perf_counter()
            O_cpu.density_tile(x[m], y[m], w[m], ew[m], hs[m], g, None, want64=False)
            tsum += time.perf_counter() - t0
            wsum += float(m.sum()) * 64 * 64
        res["cpu_density_s_extrapolated"] = tsum / max(wsum, 1.0) * work
    return res


# ------------------------------------------------------------ serving ----
def serve_jobs(count, seed=20260005, max_points=5_000_000, max_side=8192):
    """The serving stream's jobs: points log-uniform 1e3 - 5e6 (uniform or
    Thomas-clustered, 50/50), square rasters with side log-uniform 256 - 8192
    at cellsize 1, random disk or rectangle masks with holes, 50% fixed
    rule-of-thumb, 30% adaptive (Abramson, alpha -0.5), 20% bandwidth search
    over 16 candidates; edge correction on."""
    import cases
    import mgpukde_tiled_amd as M

    r = np.random.default_rng(seed)
    jobs = []
    for i in range(count):
        n = int(round(np.exp(r.uniform(np.log(1e3), np.log(max_points)))))
        side = int(round(np.exp(r.uniform(np.log(256), np.log(max_side)))))
        ext = float(side)
        if r.random() < 0.5:
            x, y = cases.uniform_points(n, ext, seed=seed + 7 * i)
        else:
            mu = 500
            x, y, _ = cases.thomas_points(max(1, n // mu), mu, ext / 200.0, ext, seed=seed + 7 * i, weights=False)
        u = r.random()
        hopt = 0 if u < 0.5 else (2 if u < 0.8 else M.MKDE_H_SEARCH)
        holes = int(r.integers(0, 8))
        if r.random() < 0.5:
            mask = cases.disk_mask(side, side, side * r.uniform(0.3, 0.5), n_holes=holes,
                                   hole_min=max(2, side // 40), hole_max=max(3, side // 10), seed=seed + 7 * i)
        else:
            mask = cases.rect_mask(side, side, n_holes=holes, hole_min=max(2, side // 40),
                                   hole_max=max(3, side // 10), seed=seed + 7 * i)
        jobs.append(dict(x=x, y=y, g=M.grid(side, side, 0.0, 0.0, 1.0), mask=mask, hoption=hopt, n=len(x),
                         side=side))
        if (i + 1) % 200 == 0:
            log("serve:   %d of %d jobs generated" % (i + 1, count))
    return jobs


def _submit(srv, j, plain=False):
    return srv.submit(j["x"], j["y"], j["g"], j["mask"], hoption=j["hoption"], edge_correction=True,
                      optimise=False, alpha=-0.5, n_candidates=16, cand_lo=0.1, cand_hi=10.0,
                      one_candidate_per_pass=plain)


def _digest(a):
    """128-bit hash of a result raster (xxh3: a few ms for a 64 MB raster;
    blake2b, ~30x slower, when xxhash is missing)."""
    try:
        import xxhash
        return xxhash.xxh3_128_hexdigest(memoryview(np.ascontiguousarray(a)).cast("B"))
    except ImportError:
        return hashlib.blake2b(memoryview(np.ascontiguousarray(a)).cast("B"), digest_size=16).hexdigest()


def run_served(jobs, slots, arrivals):
    """Serve the jobs on `slots` slots with the given arrival times (s);
    returns (per-job (result, digest), makespan from the first arrival)."""
    import mgpukde_tiled_amd as M

    results = [None] * len(jobs)
    with M.Server(slots=slots) as srv:
        tickets = [None] * len(jobs)
        submitted = threading.Event()

        def collect():
            for i in range(len(jobs)):
                while tickets[i] is None:
                    submitted.wait(0.001)
                res, out = srv.wait(tickets[i])
                results[i] = (res, _digest(out))
                if (i + 1) % 100 == 0:
                    log("serve:   %d of %d jobs done" % (i + 1, len(jobs)))

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        sub_t = []
        for i, j in enumerate(jobs):
            dt = t0 + arrivals[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            sub_t.append(time.perf_counter() - t0)
            tickets[i] = _submit(srv, j)
        th.join()
        info = srv.info()
    ends = [sub_t[i] + results[i][0]["latency_ms"] * 1e-3 for i in range(len(jobs))]
    return results, max(ends) - sub_t[0], info


def run_sequential(jobs, plain=False):
    import mgpukde_tiled_amd as M

    out = []
    t_hash = 0.0   # hashing the results is the benchmark's check, not serving work
    with M.Server(slots=1) as srv:
        t0 = time.perf_counter()
        for i, j in enumerate(jobs):
            res, o = srv.wait(_submit(srv, j, plain))
            assert res["status"] == 0, res["message"]
            th = time.perf_counter()
            out.append((res, _digest(o)))
            t_hash += time.perf_counter() - th
            if (i + 1) % 100 == 0:
                log("serve:   %d of %d jobs done one at a time" % (i + 1, len(jobs)))
        t = time.perf_counter() - t0 - t_hash
    return out, t


def bench_serving(n_jobs, slots, fractions=(0.5, 0.8, 0.95), plain=True, qps=0.0, detail=None):
    """The mixed stream one job at a time (plain: one candidate per pass;
    batched: 16 candidates per pass) and served on `slots` slots: capacity
    from a burst (every job submitted at once), then Poisson arrivals at fixed
    fractions of that capacity (or at `qps`), with p50/p99 latency at each.
    Every served result is compared by hash with the sequential one."""
    import mgpukde_tiled_amd as M

    jobs = serve_jobs(n_jobs)
    log("serve: %d jobs, %d points, %.3g cells" % (len(jobs), sum(j["n"] for j in jobs),
                                                   sum(j["side"] ** 2 for j in jobs)))
    with M.Server(slots=1) as srv:   # warm the code paths (untimed)
        for j in jobs[:2]:
            srv.wait(_submit(srv, j))
    seq, t_seq = run_sequential(jobs)
    log("serve: sequential %.2f jobs/s" % (len(jobs) / t_seq))
    out = {"jobs": len(jobs), "slots": slots, "sequential_jobs_per_s": len(jobs) / t_seq,
           "mix": {"fixed": sum(j["hoption"] == 0 for j in jobs), "adaptive": sum(j["hoption"] == 2 for j in jobs),
                   "search": sum(j["hoption"] == 3 for j in jobs)}}
    if plain:
        pl, t_plain = run_sequential(jobs, plain=True)
        out["plain_jobs_per_s"] = len(jobs) / t_plain
        out["plain_same_selected_h"] = sum(1 for a, b in zip(pl, seq) if a[0]["h_sel"] == b[0]["h_sel"])
        out["plain_identical_rasters"] = sum(1 for a, b in zip(pl, seq) if a[1] == b[1])
        log("serve: plain %.2f jobs/s" % out["plain_jobs_per_s"])
    burst, t_burst, info = run_served(jobs, slots, np.zeros(len(jobs)))
    cap = len(jobs) / t_burst
    out["capacity_jobs_per_s"] = cap
    out["served_vs_sequential"] = cap / out["sequential_jobs_per_s"]
    if plain:
        out["served_vs_plain"] = cap / out["plain_jobs_per_s"]
    same = sum(1 for a, b in zip(seq, burst) if a[1] == b[1] and a[0]["h_sel"] == b[0]["h_sel"])
    lat = np.array([r[0]["latency_ms"] for r in burst])
    out["burst"] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                    "identical_results": same, "peak_committed_bytes": info["peak_committed"],
                    "budget_bytes": info["budget"]}
    log("serve: capacity %.2f jobs/s (burst)" % cap)
    levels = []
    rates = [qps] if qps > 0 else [f * cap for f in fractions]
    for k, q in enumerate(rates):
        arr = np.cumsum(np.random.default_rng(20260006 + k).exponential(1.0 / q, len(jobs)))
        arr -= arr[0]
        res, t_run, _ = run_served(jobs, slots, arr)
        lat = np.array([r[0]["latency_ms"] for r in res])
        run = np.array([r[0]["run_ms"] for r in res])
        same = sum(1 for a, b in zip(seq, res) if a[1] == b[1] and a[0]["h_sel"] == b[0]["h_sel"])
        levels.append({"qps": q, "fraction_of_capacity": q / cap, "achieved_jobs_per_s": len(jobs) / t_run,
                       "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
                       "run_ms": {"p50": float(np.percentile(run, 50)), "p99": float(np.percentile(run, 99))},
                       "identical_results": same})
        # latency by job size: the jobs with the shorter half of the run times
        # alone (one at a time, batched) and the longer half
        alone = np.array([a[0]["run_ms"] for a in seq])
        small = alone <= np.median(alone)
        for name, m in (("small", small), ("large", ~small)):
            if m.any():
                levels[-1]["latency_ms_" + name] = {"jobs": int(m.sum()),
                                                    "p50": float(np.percentile(lat[m], 50)),
                                                    "p90": float(np.percentile(lat[m], 90)),
                                                    "p99": float(np.percentile(lat[m], 99))}
        log("serve: %.2f QPS: p50 %.0f ms, p99 %.0f ms" % (q, levels[-1]["latency_ms"]["p50"],
                                                          levels[-1]["latency_ms"]["p99"]))
    out["at_fixed_qps"] = levels
    if detail:
        with open(detail, "w") as f:
            json.dump([{"n": j["n"], "side": j["side"], "hoption": j["hoption"], "run_ms_alone": a[0]["run_ms"]}
                       for j, a in zip(jobs, seq)], f)
    return out


# ------------------------------------------------------------ headline ----
def bench_adaptive(args):
    import mgpukde_tiled_amd as M

    rates = probe_rates()
    log("rates: %s" % (json.dumps(rates) if rates else "unavailable"))
    log("workload: generating inputs")
    x, y, w, mask, side = workload(args.scale)
    n = len(x)
    log("workload: %d points, raster %dx%d" % (n, side, side))
    ctx = M.Context(0)
    grid = M.grid(side, side, 0.0, 0.0, 1.0)
    ctx.points_upload(x, y, w)
    ctx.mask_upload(grid, mask)
    ctx.tile_upload(grid, mask)

    def step():
        return ctx.run_job(hoption=2, edge_correction=True, optimise=False, alpha=-0.5, recompute_boundary=True)

    tune = None
    if not args.no_tune:
        # the autotuner picks the launch shapes for this job class on this GPU
        # (untimed, before the warm-up); the timed steps run with its choice
        tune = ctx.autotune_job(hoption=2, edge_correction=True, optimise=False, alpha=-0.5, reps=3)
        log("autotune: %.1f -> %.1f ms, %s" % (tune["baseline_ms"], tune["final_ms"], tune["chosen"]))
    for i in range(args.warmup):
        st = step()
        log("warmup %d: %.1f ms" % (i, st["ms_total"]))
    ctx.synchronize()
    wall, stats, dens_ms = [], [], []
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = step()
        ctx.synchronize()
        wall.append((time.perf_counter() - t0) * 1e3)
        stats.append(st)
        # density kernel time of this step (HIP events on the context's stream)
        dens_ms.append(ctx.last_kernel_ms("density"))
        log("step %d: %.1f ms (device %.1f ms; density %.1f ms)" % (i, wall[-1], st["ms_total"], st["ms_density"]))
    ctx.synchronize()
    t_all = time.perf_counter() - t_all0
    ms_per_step = t_all * 1e3 / args.steps

    dstats = None
    if args.stats:
        ctx.density_stats(True)
        step()
        dstats = ctx.density_stats(False)
        log("density stats: %s" % dstats)

    # --- exact work counts and job outputs (untimed) --------------------
    hs = ctx.get_bandwidths()
    ew = ctx.get_ec_weights()
    den = ctx.tile_download()
    e_den = ctx.count_density_pairs(grid, mask)
    e_ec = stats[-1]["n_ec_pairs"]
    h0 = stats[-1]["h0"]
    g_gpu = stats[-1]["g"]
    ctx.set_bandwidth(h0)
    _, e_ec0 = ctx.EdgeCorrectionWeightsExact()
    e_pilot = range_total(ctx)
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    e_total = e_den + e_ec + e_ec0 + e_pilot
    log("work: E_den=%d E_ec=%d E_ec0=%d E_pilot=%d" % (e_den, e_ec, e_ec0, e_pilot))
    ker_ms = float(np.mean(dens_ms))
    valid_cells = int(np.count_nonzero(mask != NODATA))

    out = {
        "metric": METRIC,
        "value": e_total / (ms_per_step * 1e-3),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "thomas_adaptive: %d Thomas-process points (200 parents, Poisson(5000), sigma 40), "
                        "weights U[0.5,2], %dx%d raster cs 1, disk r=3900 minus 20 holes, Abramson alpha -0.5 "
                        "at the rule-of-thumb h0, edge correction on" % (n, side, side),
            "points": n,
            "raster": [side, side],
            "parallelism": "single GPU",
        },
        "latency_ms": {"p50": float(np.percentile(wall, 50)), "p99": float(np.percentile(wall, 99)),
                       "max": float(np.max(wall))},
        "work": {"E_density": e_den, "E_ec_final": e_ec, "E_ec_pilot": e_ec0, "E_pilot": e_pilot,
                 "valid_cells": valid_cells,
                 "note": "EC pairs are the reference's per-cell evaluations; this build integrates them from FP64 "
                         "prefix sums in O(rows + columns) per point"},
        "phases_ms": {k: float(np.median([s[k] for s in stats]))
                      for k in ("ms_boundary", "ms_ec", "ms_bandwidth", "ms_density", "ms_total")},
        "density_kernel": {"evals_per_s": e_den / (ker_ms * 1e-3), "cells_per_s": valid_cells / (ker_ms * 1e-3),
                           "ms": ker_ms},
        "roofline": roofline(e_den, ker_ms, rates),
        "autotune": tune,
    }
    if dstats is not None:
        out["density_kernel"]["classification"] = dstats
    ew0 = dist = None
    if not args.no_accuracy:
        acc, ew0, dist = parity_adaptive(ctx, x, y, w, mask, side, h0, hs, ew, den, g_gpu)
        out["accuracy_vs_fp64_oracle"] = acc
        log("parity: max rel err %.3g (oracle %.1f s)" % (acc["max_rel_err"], acc["oracle_s"]))
    ctx.close()
    O_cpu = None
    if not args.no_cpu:
        O_cpu, build = cpu_timing_oracle()
        if ew0 is None:
            with M.Context(0) as c2:
                c2.points_upload(x, y, w)
                c2.mask_upload(grid, mask)
                c2.set_bandwidth(h0)
                ew0, _ = c2.EdgeCorrectionWeightsExact()
                dist = c2.CalcDist2Boundary()
        cb = cpu_baseline_adaptive(O_cpu, build, x, y, w, ew, hs, h0, ew0, dist, mask, side, valid_cells)
        cb["plain_small_full"] = cpu_plain_small(O_cpu)
        out["cpu_baseline"] = cb
        log("cpu: %.3g evals/s, job extrapolated %.0f s" % (cb["value"], cb["job_s_extrapolated"]))
    else:
        out["cpu_baseline"] = None
    return out, O_cpu


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workloads (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the oracle checks")
    ap.add_argument("--no-legs", action="store_true", help="skip the search and tiles workloads")
    ap.add_argument("--no-tune", action="store_true", help="run with the default launch shapes (no autotuner)")
    ap.add_argument("--stats", action="store_true", help="report density-kernel classification counters")
    ap.add_argument("--jobs", type=int, default=120, help="--workload serve: number of jobs")
    ap.add_argument("--serve-jobs", type=int, default=40,
                    help="jobs of the serving sample in the benchmark line (0: none)")
    ap.add_argument("--slots", type=int, default=4, help="serve: worker slots")
    ap.add_argument("--qps", type=float, default=0.0, help="serve: one arrival rate instead of fractions of capacity")
    ap.add_argument("--detail", default="", help="serve: write per-job sizes and run times to this file")
    ap.add_argument("--no-plain", action="store_true", help="serve: skip the one-candidate-per-pass baseline")
    ap.add_argument("--fractions", default="0.5,0.8,0.95",
                    help="serve: arrival rates as fractions of the measured capacity (empty: capacity only)")
    ap.add_argument("--workload", default="adaptive", choices=["adaptive", "search", "tiles", "serve", "copy", "rates"],
                    help="adaptive: the benchmark line; the others measure one workload on its own")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "rates":
        print(json.dumps(probe_rates()), flush=True)
        return
    if args.workload == "copy":
        return bench_copy()
    if args.workload == "search":
        print(json.dumps(leg_search(args, with_oracle=not args.no_accuracy)), flush=True)
        return
    if args.workload == "tiles":
        print(json.dumps(leg_tiles(args, with_oracle=not args.no_accuracy, stats=args.stats)), flush=True)
        return
    if args.workload == "serve":
        fr = tuple(float(v) for v in args.fractions.split(",") if v.strip())
        print(json.dumps(bench_serving(args.jobs, args.slots, fractions=fr, qps=args.qps, plain=not args.no_plain,
                                       detail=args.detail or None)),
              flush=True)
        return
    out, O_cpu = bench_adaptive(args)
    if not args.no_legs:
        log("search leg")
        out["search"] = leg_search(args, with_oracle=not args.no_accuracy, O_cpu=O_cpu)
        log("search: %.1f ms per search, argmax %d" % (out["search"]["ms_per_search"], out["search"]["argmax"]))
        log("tiles leg")
        out["tiles"] = leg_tiles(args, with_oracle=not args.no_accuracy, O_cpu=O_cpu)
        log("tiles: %.1f ms per raster" % out["tiles"]["ms_per_raster"])
    if args.serve_jobs > 0:
        out["serving"] = bench_serving(args.serve_jobs, args.slots, qps=args.qps)
    print(json.dumps(out), flush=True)


def bench_copy():
    """Calibration of the HBM byte counters: three device-to-device copies of
    1 GiB (FETCH_SIZE / WRITE_SIZE of this copy against its known 1 GiB read
    and 1 GiB write give the counters' scale)."""
    import torch

    n = 1 << 28
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"workload": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms": dt * 1e3,
                      "GBps": 8.0 * n / dt / 1e9}), flush=True)


if __name__ == "__main__":
    main()
