"""The job server: every result the same alone and in the mix (bit-identical:
jobs share no device state and every kernel is deterministic), parity of a
served job with the oracle, HBM admission (refusal, waiting, peak within the
budget), failure isolation and cancellation, and the footprint bound against
measured context memory."""
import numpy as np
import pytest

import cases
import mgpukde_tiled_amd as M

pytestmark = pytest.mark.gpu
REL = 1e-5


def make_jobs(seed=5, count=10):
    r = np.random.default_rng(seed)
    jobs = []
    kinds = [("fixed", 0, False), ("adaptive", 2, False), ("search", M.MKDE_H_SEARCH, False), ("mle", 1, False),
             ("file", -1, False), ("hj", 2, True)]
    for i in range(count):
        name, hopt, opt = kinds[i % len(kinds)]
        side = int(r.integers(64, 300))
        n = int(r.integers(200, 3000))
        if i % 2:
            x, y = cases.uniform_points(n, float(side), seed=seed * 100 + i)
        else:
            x, y, _ = cases.thomas_points(max(1, n // 200), 200, side / 30.0, float(side), seed=seed * 100 + i)
        w = r.uniform(0.5, 2.0, len(x)).astype(np.float32)
        mask = cases.disk_mask(side, side, side * 0.45, n_holes=3, hole_min=3, hole_max=max(4, side // 8),
                               seed=seed * 100 + i)
        h = r.uniform(side / 60.0, side / 20.0, len(x)).astype(np.float32) if hopt == -1 else None
        jobs.append(dict(name=name, x=x, y=y, w=w, h=h, g=M.grid(side, side, 0.0, 0.0, 1.0), mask=mask,
                         hoption=hopt, optimise=opt))
    return jobs


def submit(srv, j):
    return srv.submit(j["x"], j["y"], j["g"], j["mask"], w=j["w"], h=j["h"], hoption=j["hoption"],
                      edge_correction=True, optimise=j["optimise"], alpha=-0.5, n_candidates=16)


def test_alone_equals_in_the_mix():
    jobs = make_jobs()
    alone = []
    with M.Server(slots=1) as srv:
        for j in jobs:
            res, out = srv.wait(submit(srv, j))
            assert res["status"] == 0, res["message"]
            alone.append((res, out.copy()))
    with M.Server(slots=4) as srv:
        tickets = [submit(srv, j) for j in jobs]
        mixed = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert info["finished"] == len(jobs) and info["peak_committed"] <= info["budget"]
    slots_used = set()
    for j, (ra, oa), (rm, om) in zip(jobs, alone, mixed):
        assert rm["status"] == 0, rm["message"]
        assert np.array_equal(oa, om), j["name"]
        assert ra["h_sel"] == rm["h_sel"] and ra["alpha_sel"] == rm["alpha_sel"] and ra["h0"] == rm["h0"]
        slots_used.add(rm["slot"])
    assert len(slots_used) > 1


def test_served_job_equals_context_path_and_oracle(orc, ctx):
    side = 160
    x, y, w = cases.thomas_points(4, 300, 6.0, float(side), seed=77)
    mask = cases.disk_mask(side, side, 70.0, n_holes=2, hole_min=5, hole_max=20, seed=77)
    g = M.grid(side, side, 0.0, 0.0, 1.0)
    with M.Server(slots=2) as srv:
        res, out = srv.wait(srv.submit(x, y, g, mask, w=w, hoption=0, edge_correction=True))
    assert res["status"] == 0, res["message"]
    ctx.points_upload(x, y, w)
    ctx.mask_upload(g, mask)
    ctx.tile_from_mask()
    st = ctx.run_job(hoption=0, edge_correction=True, optimise=False, recompute_boundary=False)
    assert np.array_equal(ctx.tile_download(), out)
    assert st["h0"] == res["h0"]
    og = orc.grid(side, side, 0.0, 0.0, 1.0)
    h0 = orc.rule_of_thumb(x, y, w)
    assert np.float32(h0) == np.float32(res["h0"])
    flags, _ = orc.mark_boundary(og, mask)
    d = orc.dist_to_boundary(x, y, og, flags)
    hs = np.full(len(x), h0, np.float32)
    _, ew64, _ = orc.ec_weights(x, y, d, hs, og, mask)
    _, o64, _ = orc.density_tile(x, y, w, ew64.astype(np.float32), hs, og, mask)
    valid = mask != -9999.0
    assert np.all(out[~valid] == mask[~valid])
    pos = valid & (o64 > 0)
    assert np.all(out[valid & ~pos] == 0)
    assert np.max(np.abs(out[pos] - o64[pos]) / o64[pos]) <= REL


def test_search_job_selects_the_context_argmax(ctx):
    x, y = cases.mixture_points(1500, 200.0, n_centres=8, seed=31)
    mask = cases.all_valid(200, 200)
    g = M.grid(200, 200, 0.0, 0.0, 1.0)
    with M.Server(slots=2) as srv:
        res, out = srv.wait(srv.submit(x, y, g, mask, hoption=M.MKDE_H_SEARCH, n_candidates=16, cand_lo=0.1,
                                       cand_hi=10.0))
    assert res["status"] == 0, res["message"]
    ctx.points_upload(x, y, None)
    ctx.mask_upload(g, mask)
    h0 = ctx.BandWidth_rule_of_thumb()
    # the job descriptor carries lo / hi as floats: h_k = (float)(h0 * lo * (hi/lo)^(k/15)) in double
    lo, hi = float(np.float32(0.1)), float(np.float32(10.0))
    hc = np.array([np.float32(float(h0) * lo * (hi / lo) ** (k / 15.0)) for k in range(16)], np.float32)
    L = ctx.LogLikelihood(hc, edge_correction=True)
    assert res["h_sel"] == hc[int(np.argmax(L))]


def test_admission_budget_refusal_and_waiting():
    jobs = make_jobs(seed=9, count=6)
    bounds = [M.job_footprint(len(j["x"]), j["g"].ncols, j["g"].nrows, j["hoption"], j["optimise"], 16)
              for j in jobs]
    # room for about one job at a time beyond what the slots keep
    budget = int(max(bounds) * 2.2)
    with M.Server(slots=3, hbm_budget=budget) as srv:
        big = np.zeros(10, np.float32)
        assert M.job_footprint(10, 1000, 1000) > budget
        with pytest.raises(M.MkdeError) as e:
            srv.submit(big, big, M.grid(1000, 1000, 0, 0, 1.0), None, edge_correction=False)
        assert e.value.code == M.MKDE_E_NOMEM
        tickets = [submit(srv, j) for j in jobs]
        outs = [srv.wait(t) for t in tickets]
        info = srv.info()
    assert info["refused"] == 1 and info["finished"] == len(jobs)
    assert info["peak_committed"] <= budget
    with M.Server(slots=1) as srv:
        ref = [srv.wait(submit(srv, j))[1] for j in jobs]
    for (res, out), r in zip(outs, ref):
        assert res["status"] == 0, res["message"]
        assert np.array_equal(out, r)


def test_failed_and_cancelled_jobs_leave_the_server_working():
    jobs = make_jobs(seed=13, count=4)
    bad = dict(jobs[0])
    bad["x"] = bad["x"].copy()
    bad["x"][3] = np.nan                  # rejected inside the worker (points upload)
    with M.Server(slots=1) as srv:
        t_bad = submit(srv, bad)
        t_ok = [submit(srv, j) for j in jobs]
        t_cancel = submit(srv, jobs[1])
        srv.cancel(t_cancel)
        rb, _ = srv.wait(t_bad)
        rc, _ = srv.wait(t_cancel)
        ok = [srv.wait(t) for t in t_ok]
    assert rb["status"] == 1 and "non-finite" in rb["message"]
    assert rc["status"] == M.MKDE_E_CANCELLED and rc["slot"] == -1
    with M.Server(slots=2) as srv:
        again = [srv.wait(submit(srv, j)) for j in jobs]
    for (r1, o1), (r2, o2) in zip(ok, again):
        assert r1["status"] == 0 and r2["status"] == 0
        assert np.array_equal(o1, o2)


def test_footprint_bound_holds_for_each_job_kind():
    for j in make_jobs(seed=21, count=6):
        c = M.Context(0)
        c.points_upload(j["x"], j["y"], j["w"])
        c.mask_upload(j["g"], j["mask"])
        c.tile_from_mask()
        if j["hoption"] == -1:
            c.set_bandwidths(j["h"])
        if j["hoption"] == M.MKDE_H_SEARCH:
            h0 = c.BandWidth_rule_of_thumb()
            c.LogLikelihood(np.float32(h0) * np.logspace(-1, 1, 16).astype(np.float32), edge_correction=True)
            c.set_bandwidth(h0)
            c.run_job(hoption=-1, edge_correction=True, optimise=False, recompute_boundary=False)
        else:
            c.run_job(hoption=j["hoption"], edge_correction=True, optimise=j["optimise"], recompute_boundary=False)
        c.tile_download()
        used = c.memory()
        bound = M.job_footprint(len(j["x"]), j["g"].ncols, j["g"].nrows, j["hoption"], j["optimise"], 16)
        assert 0 < used <= bound, (j["name"], used, bound)
        c.release()
        assert c.memory() == 0
        c.close()


def test_one_candidate_per_pass_selects_the_same_bandwidth():
    # the plain form (one leave-one-out pass per candidate) and the batched
    # form (16 candidates per pass) select the same bandwidth; the rasters
    # agree to the FP32 sum tolerance
    x, y = cases.mixture_points(2500, 240.0, n_centres=10, seed=43)
    mask = cases.disk_mask(240, 240, 110.0, n_holes=3, hole_min=5, hole_max=30, seed=43)
    g = M.grid(240, 240, 0.0, 0.0, 1.0)
    with M.Server(slots=2) as srv:
        rb, ob = srv.wait(srv.submit(x, y, g, mask, hoption=M.MKDE_H_SEARCH, n_candidates=16))
        rp, op = srv.wait(srv.submit(x, y, g, mask, hoption=M.MKDE_H_SEARCH, n_candidates=16,
                                     one_candidate_per_pass=True))
    assert rb["status"] == 0 and rp["status"] == 0
    assert rb["h_sel"] == rp["h_sel"]
    valid = mask != -9999.0
    pos = valid & (ob > 0)
    assert np.all(op[~valid] == ob[~valid])
    assert np.max(np.abs(op[pos] - ob[pos]) / ob[pos]) <= 1e-5
