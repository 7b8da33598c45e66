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
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

# MI355X dense FP32 peak: the FP32 MFMA rate (the same figure as the packed
# FP32 VALU rate); the density kernel runs its FULL cells on FP32 MFMA and its
# PARTIAL cells on the VALU
MI355X_FP32_TFLOPS = 157.3
FLOP_PER_EVAL = 8  # 2 sub + 2 mul/fma for d^2, 1 mul for the scale, 1 fma to accumulate


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(scale=1.0):
    import cases

    extent = 8192.0 * scale
    n_parents = max(1, int(round(200 * scale * scale)))
    x, y, w = cases.thomas_points(n_parents, 5000, 40.0, extent, seed=20260002)
    side = int(round(8192 * scale))
    mask = cases.disk_mask(side, side, 3900.0 * scale, n_holes=20, hole_min=max(2, int(50 * scale)),
                           hole_max=max(3, int(400 * scale)), seed=20260002)
    return x, y, w, mask, side


def density_traffic():
    """HBM bytes per launch of the density kernel on this workload, from the
    two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) scaled by the same
    counters on a 1 GiB copy; None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "k_density_tile_thomas_adaptive_1M_8192_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(float(line.split(":")[1]))
    return None


def cpu_baseline(x, y, w, ew, hs, mask, side, target_s=15.0):
    """The oracle's restatement of the reference CPU loop (brute force over all
    points per cell, float accumulation) on a seeded sample of 64x64 tiles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rs = np.random.default_rng(20260099)
    nthreads = O.num_threads()
    times = 0.0
    e_total = 0
    cells = 0
    tiles = 0
    while times < target_s and tiles < 64:
        r0 = int(rs.integers(0, side // 64)) * 64
        c0 = int(rs.integers(0, side // 64)) * 64
        tile = O.grid(64, 64, float(c0), float(side - r0 - 64), 1.0)
        tm = mask[r0:r0 + 64, c0:c0 + 64]
        t0 = time.perf_counter()
        _, _, ne = O.density_tile(x, y, w, ew, hs, tile, tm, want64=False)
        times += time.perf_counter() - t0
        e_total += ne
        cells += int(np.count_nonzero(tm != -9999.0))
        tiles += 1
    return {
        "value": e_total / times if times > 0 else None,
        "unit": "evals/s",
        "cores": nthreads,
        "kind": "port",
        "sample": "%d seeded 64x64 tiles of the 8192^2 raster (%d valid cells, %d evaluations), all %d points, "
                  "reference CPU loop (brute force over all points per cell, float accumulation), %.1f s"
                  % (tiles, cells, e_total, len(x), times),
        "cells_per_s": cells / times if times > 0 else None,
    }


def accuracy(x, y, w, ew, hs, mask, side, den, ntiles=3):
    """Max / mean relative error of the GPU raster against the FP64 oracle on a
    few seeded 64x64 tiles."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    rs = np.random.default_rng(20260098)
    mx, acc, cnt = 0.0, 0.0, 0
    for _ in range(ntiles):
        r0 = int(rs.integers(0, side // 64)) * 64
        c0 = int(rs.integers(0, side // 64)) * 64
        tile = O.grid(64, 64, float(c0), float(side - r0 - 64), 1.0)
        tm = mask[r0:r0 + 64, c0:c0 + 64]
        _, o64, _ = O.density_tile(x, y, w, ew, hs, tile, tm, want64=True)
        g = den[r0:r0 + 64, c0:c0 + 64]
        v = (tm != -9999.0) & (o64 > 0)
        if np.any(v):
            e = np.abs(g[v].astype(np.float64) - o64[v]) / o64[v]
            mx = max(mx, float(e.max()))
            acc += float(e.sum())
            cnt += int(v.sum())
    return {"max_rel_err": mx, "mean_rel_err": acc / max(cnt, 1), "cells_checked": cnt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=float, default=1.0, help="shrink the workload (testing only)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--stats", action="store_true", help="report density-kernel classification counters")
    ap.add_argument("--jobs", type=int, default=120, help="serve: number of jobs")
    ap.add_argument("--serve-jobs", type=int, default=40,
                    help="jobs of the serving sample added to the benchmark line (0: none)")
    ap.add_argument("--slots", type=int, default=4, help="serve: worker slots")
    ap.add_argument("--qps", type=float, default=0.0, help="serve: arrival rate (0: 4x the sequential rate)")
    ap.add_argument("--plain", action="store_true",
                    help="serve: also time the plain form (one job at a time, one candidate per pass)")
    ap.add_argument("--workload", default="adaptive", choices=["adaptive", "search", "tiles", "serve", "copy"],
                    help="adaptive: the benchmark line; search / tiles: measurements of the other workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (--gpus %d requested)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import mgpukde_tiled_amd as M

    if args.workload == "search":
        return bench_search(args)
    if args.workload == "tiles":
        return bench_tiles(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "copy":
        return bench_copy(args)
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

    for i in range(args.warmup):
        st = step()
        log("warmup %d: %.1f ms" % (i, st["ms_total"]))
    ctx.synchronize()
    wall = []
    stats = []
    dens_ms = []   # density kernel time of every timed step (HIP events on the context's stream)
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        t0 = time.perf_counter()
        st = step()
        ctx.synchronize()
        wall.append((time.perf_counter() - t0) * 1e3)
        stats.append(st)
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

    # --- exact work counts (untimed) ---------------------------------------
    hs = ctx.get_bandwidths()
    ew = ctx.get_ec_weights()
    e_den = ctx.count_density_pairs(grid, mask)
    e_ec = stats[-1]["n_ec_pairs"]
    h0 = stats[-1]["h0"]
    # pilot pairs and EC(h0) pairs at the rule-of-thumb bandwidth
    ctx.set_bandwidth(h0)
    _, e_ec0 = ctx.EdgeCorrectionWeightsExact()
    offs, _ = ctx_range_total(ctx)
    e_pilot = offs
    ctx.set_bandwidths(hs)
    ctx.set_ec_weights(ew)
    e_total = e_den + e_ec + e_ec0 + e_pilot
    log("work: E_den=%d E_ec=%d E_ec0=%d E_pilot=%d" % (e_den, e_ec, e_ec0, e_pilot))

    ker_ms = float(np.mean(dens_ms))
    achieved_tflops = FLOP_PER_EVAL * e_den / (ker_ms * 1e-3) / 1e12
    valid_cells = int(np.count_nonzero(mask != -9999.0))

    out = {
        "metric": "Gaussian kernel evals/s + %VALU peak, 1M-pt adaptive KDE; p99 job latency",
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
                 "valid_cells": valid_cells},
        "phases_ms": {k: float(np.median([s[k] for s in stats]))
                      for k in ("ms_boundary", "ms_ec", "ms_bandwidth", "ms_density", "ms_total")},
        "density_kernel": {"evals_per_s": e_den / (ker_ms * 1e-3), "cells_per_s": valid_cells / (ker_ms * 1e-3),
                           "ms": ker_ms},
        "roofline": {
            "bound": "mfma",
            "achieved": achieved_tflops,
            "peak": MI355X_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / MI355X_FP32_TFLOPS,
            "traffic": density_traffic(),
            "kernel": "k_density_tile",
            "work_per_launch": "%d evaluations x %d FLOP" % (e_den, FLOP_PER_EVAL),
        },
    }
    if dstats is not None:
        out["density_kernel"]["classification"] = dstats
    if not args.no_accuracy:
        den = ctx.tile_download()
        out["accuracy_vs_fp64_oracle"] = accuracy(x, y, w, ew, hs, mask, side, den)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, y, w, ew, hs, mask, side)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    if args.serve_jobs > 0:
        # job latency of the served mixed stream (a bounded sample of it)
        args.jobs = args.serve_jobs
        args.plain = True
        sv = bench_serve(args, emit=False)
        out["serving"] = {k: sv[k] for k in ("jobs", "slots", "qps", "sequential_jobs_per_s", "served_jobs_per_s",
                                             "speedup", "served_latency_ms", "served_run_ms", "identical_results",
                                             "plain")}
    print(json.dumps(out), flush=True)


def bench_search(args):
    """Batched bandwidth search: 200,000 points (50% uniform, 50% around 64
    Gaussian centres with sigma 2% of the extent) on an all-valid 4096^2 mask;
    LOO log-likelihood of 64 log-spaced bandwidths 0.1x-10x the rule of thumb,
    edge correction recomputed per candidate."""
    import cases
    import mgpukde_tiled_amd as M

    side = int(round(4096 * args.scale))
    n = int(round(200000 * args.scale * args.scale))
    x, y = cases.mixture_points(n, float(side), n_centres=64, seed=20260003)
    mask = cases.all_valid(side, side)
    ctx = M.Context(0)
    ctx.points_upload(x, y, None)
    ctx.mask_upload(M.grid(side, side, 0.0, 0.0, 1.0), mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    hc = (np.float32(h0) * (10.0 ** (-1.0 + 2.0 * np.arange(64) / 63.0))).astype(np.float32)
    for _ in range(args.warmup):
        ctx.LogLikelihood(hc, edge_correction=True)
    t = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        L = ctx.LogLikelihood(hc, edge_correction=True)
        t.append(time.perf_counter() - t0)
    k = int(np.argmax(L))
    s = np.sort(L)
    # work (untimed): leave-one-out pairs and EC pairs of every candidate
    e_loo = e_ec = 0
    for hk in hc:
        ctx.set_bandwidth(float(hk))
        e_loo += ctx_range_total(ctx)[0] - n
        e_ec += ctx.EdgeCorrectionWeightsExact()[1]
    print(json.dumps({"workload": "search: %d mixture points, %d^2 all-valid mask, 64 candidates" % (n, side),
                      "ms_per_search": 1e3 * float(np.median(t)), "h_rule_of_thumb": h0,
                      "argmax": k, "h_selected": float(hc[k]), "top2_margin": float(s[-1] - s[-2]),
                      "E_loo": e_loo, "E_ec": e_ec,
                      "evals_per_s": (e_loo + e_ec) / float(np.median(t)),
                      "logL": [float(v) for v in L]}), flush=True)
    ctx.close()


def bench_tiles(args):
    """Precision stress: 10M points with log-normal cluster intensities over
    1024 centres, a 32768^2 GeoTIFF-style raster at cellsize 1 processed as
    4096^2 tiles (each tile with its own origin) in one pass, fixed h = 16,
    EC on.  Reports evaluations/s, HBM GB/s of the raster write, and the
    max / mean relative error against the FP64 oracle on one 128^2 window in
    each of several seeded tiles."""
    import cases
    import mgpukde_tiled_amd as M

    side = int(round(32768 * args.scale))
    tile = max(64, int(round(4096 * args.scale)))
    n = int(round(10_000_000 * args.scale * args.scale))
    h = 16.0
    x, y = cases.lognormal_clusters(n, float(side), 1024, 64.0, seed=20260004)
    w = np.ones(n, np.float32)
    gt = [0.0, 1.0, 0.0, float(side), 0.0, -1.0]
    ctx = M.Context(0)
    ctx.points_upload(x, y, None)
    ctx.set_bandwidth(h)
    log("tiles: %d points, %d^2 raster, %d^2 tiles" % (n, side, tile))
    ctx.mask_upload(M.grid(side, side, 0.0, 0.0, 1.0), cases.all_valid(side, side, 1.0))
    t0 = time.perf_counter()
    ew, e_ec = ctx.EdgeCorrectionWeightsExact()
    ms_ec = 1e3 * (time.perf_counter() - t0)
    for _ in range(args.warmup + 1):
        ctx.KernelDensityEstimation_raster(side, side, gt, -9999.0, tile, tile, None, keep_on_device=True)
    t, ker = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.KernelDensityEstimation_raster(side, side, gt, -9999.0, tile, tile, None, keep_on_device=True)
        t.append(time.perf_counter() - t0)
        ker.append(ctx.last_kernel_ms("density"))
    den = ctx.tile_download()
    log("tiles: density %.2f ms (kernel %.2f ms)" % (1e3 * float(np.median(t)), float(np.median(ker))))
    dstats = None
    if args.stats:
        ctx.density_stats(True)
        ctx.KernelDensityEstimation_raster(side, side, gt, -9999.0, tile, tile, None, keep_on_device=True)
        dstats = ctx.density_stats(False)
        log("density stats: %s" % dstats)
    # exact work: pairs counted tile by tile in each tile's own geometry
    e_den = 0
    ntile = (side + tile - 1) // tile
    for ty in range(ntile):
        for tx in range(ntile):
            tc, tr = min(tile, side - tx * tile), min(tile, side - ty * tile)
            g = M.grid(tc, tr, np.float32(tx * tile), np.float32(side - ty * tile - tr), 1.0)
            e_den += ctx.count_density_pairs(g, None)
    kms = float(np.median(ker))
    res = {"workload": "tiles", "points": n, "raster": side, "tile": tile, "h": h,
           "ms_per_raster": 1e3 * float(np.median(t)), "kernel_ms": kms, "ms_ec": ms_ec,
           "E_density": e_den, "E_ec": e_ec, "evals_per_s": e_den / (kms * 1e-3),
           "raster_write_GBps": 4.0 * side * side / (kms * 1e-3) / 1e9, "density_stats": dstats}
    if not args.no_accuracy:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        rs = np.random.default_rng(20260097)
        hs = np.full(n, h, np.float32)
        per_tile = []
        for _ in range(6):
            wc, wr = min(128, tile, side), min(128, tile, side)
            # centre the window on a random point, so it lands where the density is
            k = int(rs.integers(0, n))
            c0 = int(min(max(int(x[k]) - wc // 2, 0), side - wc))
            r0 = int(min(max(side - 1 - int(y[k]) - wr // 2, 0), side - wr))
            tx, ty = c0 // tile, r0 // tile
            c0 = min(c0, tx * tile + min(tile, side - tx * tile) - wc)
            r0 = min(r0, ty * tile + min(tile, side - ty * tile) - wr)
            win = O.grid(wc, wr, float(c0), float(side - r0 - wr), 1.0)
            sel = O.tile_point_filter(x, y, hs, win)
            _, o64, _ = O.density_tile(x[sel], y[sel], w[sel], ew[sel], hs[sel], win, None, want64=True)
            gwin = den[r0:r0 + wr, c0:c0 + wc].astype(np.float64)
            v = o64 > 0
            if np.any(v):
                e = np.abs(gwin[v] - o64[v]) / o64[v]
                per_tile.append({"tile": [tx, ty], "max_rel_err": float(e.max()), "mean_rel_err": float(e.mean()),
                                 "cells": int(v.sum())})
            assert np.all(gwin[~v] == 0.0)
        res["accuracy"] = per_tile
        res["max_rel_err"] = max([p["max_rel_err"] for p in per_tile] + [0.0])
    print(json.dumps(res), flush=True)
    ctx.close()


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
    return jobs


def bench_serve(args, emit=True):
    """Serving stream: the same job list run one job at a time (one slot,
    back to back) and served (several slots, Poisson arrivals at a fixed QPS);
    jobs/s, p50/p99 latency and bit-equality of every result."""
    import hashlib
    import threading

    import mgpukde_tiled_amd as M

    log("serve: generating %d jobs" % args.jobs)
    jobs = serve_jobs(args.jobs)
    log("serve: %d points, %.3g cells in total" % (sum(j["n"] for j in jobs), sum(j["side"] ** 2 for j in jobs)))

    def submit(srv, j, plain=False):
        return srv.submit(j["x"], j["y"], j["g"], j["mask"], hoption=j["hoption"], edge_correction=True,
                          optimise=False, alpha=-0.5, n_candidates=16, cand_lo=0.1, cand_hi=10.0,
                          one_candidate_per_pass=plain)

    def digest(a):
        return hashlib.blake2b(a.tobytes(), digest_size=16).hexdigest()

    # warm the slots' buffers and code paths on a few jobs (untimed)
    with M.Server(slots=1) as srv:
        for j in jobs[:2]:
            srv.wait(submit(srv, j))
    # one job at a time
    seq = []
    with M.Server(slots=1) as srv:
        t0 = time.perf_counter()
        for j in jobs:
            res, out = srv.wait(submit(srv, j))
            assert res["status"] == 0, res["message"]
            seq.append((res, digest(out)))
        t_seq = time.perf_counter() - t0
    seq_rate = len(jobs) / t_seq
    log("serve: sequential %.2f jobs/s (%.1f s)" % (seq_rate, t_seq))
    # the plain form: one job at a time and one candidate bandwidth per
    # leave-one-out pass (the reference's per-candidate LogLikelihood calls)
    plain = []
    if args.plain:
        with M.Server(slots=1) as srv:
            t0 = time.perf_counter()
            for j in jobs:
                res, out = srv.wait(submit(srv, j, plain=True))
                assert res["status"] == 0, res["message"]
                plain.append((res, out.copy() if j["hoption"] == 3 else None, digest(out)))
            t_plain = time.perf_counter() - t0
        log("serve: plain %.2f jobs/s (%.1f s)" % (len(jobs) / t_plain, t_plain))
    qps = args.qps if args.qps > 0 else 4.0 * seq_rate
    arr = np.cumsum(np.random.default_rng(20260006).exponential(1.0 / qps, len(jobs)))
    results = [None] * len(jobs)
    with M.Server(slots=args.slots) as srv:
        tickets = [None] * len(jobs)
        submitted = threading.Event()

        def collect():
            for i in range(len(jobs)):
                while tickets[i] is None:
                    submitted.wait(0.001)
                res, out = srv.wait(tickets[i])
                results[i] = (res, digest(out))

        th = threading.Thread(target=collect)
        th.start()
        t0 = time.perf_counter()
        sub_t = []
        for i, j in enumerate(jobs):
            dt = t0 + arr[i] - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            sub_t.append(time.perf_counter() - t0)
            tickets[i] = submit(srv, j)
        th.join()
        info = srv.info()
    ends = [sub_t[i] + results[i][0]["latency_ms"] * 1e-3 for i in range(len(jobs))]
    t_served = max(ends) - sub_t[0]
    served_rate = len(jobs) / t_served
    lat = np.array([r[0]["latency_ms"] for r in results])
    run = np.array([r[0]["run_ms"] for r in results])
    seq_run = np.array([r[0]["run_ms"] for r in seq])
    same = sum(1 for a, b in zip(seq, results) if a[1] == b[1] and a[0]["h_sel"] == b[0]["h_sel"])
    plain_info = None
    if plain:
        # plain vs batched: the same selected bandwidths; rasters bit-identical
        # except where the summation order of a candidate pass differs
        same_h = sum(1 for a, b in zip(plain, seq) if a[0]["h_sel"] == b[0]["h_sel"])
        same_r = sum(1 for a, b in zip(plain, seq) if a[2] == b[1])
        plain_info = {"plain_jobs_per_s": len(jobs) / t_plain, "served_vs_plain": (len(jobs) / t_served) /
                      (len(jobs) / t_plain), "plain_same_selected_h": same_h, "plain_identical_rasters": same_r}
    out = {"workload": "serve", "jobs": len(jobs), "slots": args.slots, "qps": qps,
           "sequential_jobs_per_s": seq_rate, "served_jobs_per_s": served_rate, "speedup": served_rate / seq_rate,
           "served_latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99))},
           "served_run_ms": {"p50": float(np.percentile(run, 50)), "p99": float(np.percentile(run, 99))},
           "sequential_run_ms": {"p50": float(np.percentile(seq_run, 50)), "p99": float(np.percentile(seq_run, 99))},
           "identical_results": same, "plain": plain_info,
           "peak_committed_bytes": info["peak_committed"], "budget_bytes": info["budget"],
           "mix": {"fixed": sum(j["hoption"] == 0 for j in jobs), "adaptive": sum(j["hoption"] == 2 for j in jobs),
                   "search": sum(j["hoption"] == 3 for j in jobs)},
           "sequential_ms_by_kind": {k: float(sum(r[0]["run_ms"] for j, r in zip(jobs, seq) if j["hoption"] == h))
                                     for k, h in (("fixed", 0), ("adaptive", 2), ("search", 3))},
           "slowest_sequential": [{"points": jobs[i]["n"], "side": jobs[i]["side"], "hoption": jobs[i]["hoption"],
                                   "run_ms": seq[i][0]["run_ms"]} for i in np.argsort(-seq_run)[:6]]}
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_copy(args):
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


def ctx_range_total(ctx):
    import ctypes

    import mgpukde_tiled_amd as M

    tot = ctypes.c_int64()
    M._check(M._lib.mkde_range_candidates(ctx._h, None, None, 0, ctypes.byref(tot)))
    return tot.value, None


if __name__ == "__main__":
    main()
