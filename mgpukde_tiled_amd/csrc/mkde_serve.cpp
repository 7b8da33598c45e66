// mkde_serve.cpp -- the job server: a stream of KDE jobs of different point
// counts, raster sizes and bandwidth options served by worker slots on one
// device.  No reference counterpart (the reference runs one job per process,
// kde_cuda_kdtr_tiled.cu main()).
//
// Each slot is a host thread that owns one mkde_ctx, hence one HIP stream and
// its own device buffers, grown to the largest job the slot has run and
// reused (no hipMalloc/hipFree in steady state).  While one slot's job runs
// kernels, another's does its host work (input checks, copies, waits),
// and small jobs' kernels fill the device side by side.  Jobs start in
// submission order when a slot is free and the HBM admission test passes:
//
//   held (device bytes of all slots) + sum of bounds of running jobs
//       + bound of the next job  <=  budget
//
// where a job's bound (mkde_job_footprint) caps what it can add to its slot's
// buffers.  A job further back that fits may start while the head does not,
// until MKDE_SKIP_MAX jobs have passed the head; then the head waits for
// memory to free up.  When the head does not fit and nothing runs, idle slots
// drop their buffers.  Jobs share no device state, so each result is
// bit-identical to the same job run alone, in any order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "mkde.h"
#include "mkde_device.h"

extern "C" int mkde_internal_fail(int code, const char *msg);
namespace mkde_dev {
Tuning default_tuning();
}

namespace {

using Clock = std::chrono::steady_clock;

// Jobs that may start ahead of a queue head that does not fit in the budget.
constexpr int MKDE_SKIP_MAX = 16;

int sfail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
int sfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    return mkde_internal_fail(code, buf);
}

#define SCHK(expr)                    \
    do {                              \
        int _s = (expr);              \
        if (_s != MKDE_OK) return _s; \
    } while (0)

enum JobState { QUEUED, RUNNING, DONE };

struct Job {
    mkde_serve_job d{};
    int64_t bound = 0;
    Clock::time_point t_submit, t_start, t_end;
    JobState state = QUEUED;
    mkde_serve_result res{};
};

struct Slot {
    std::thread th;
    mkde_ctx *ctx = nullptr;
    int64_t held = 0;   // device bytes of ctx after its last job
};

double ms_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

struct mkde_server {
    int device = 0;
    int64_t budget = 0;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::unordered_map<int64_t, Job> jobs;   // element references stay valid across inserts
    int64_t next_ticket = 1;
    int head_skips = 0;   // jobs started ahead of the current queue head
    int64_t held_total = 0;
    int64_t committed_running = 0;
    int64_t peak = 0;
    int64_t submitted = 0, finished = 0, refused = 0, releases = 0, running = 0;
    int ready = 0, init_status = MKDE_OK;
    std::string init_msg;
    bool stop = false;
};

extern "C" int64_t mkde_job_footprint(int64_t n, int64_t ncols, int64_t nrows, int hoption, int optimise,
                                      int n_candidates)
{
    (void)n_candidates;   // candidates are evaluated 16 at a time
    if (n < 0 || ncols < 0 || nrows < 0) return -1;
    const int64_t nbins = n / 32 + 8200;   // uniform bins: <= n/32 + 2*4096 + 1
    int64_t b = (int64_t)1 << 20;          // counters, per-candidate constants, reduction partials
    // coordinates (original and bin order), weights, permutation, point
    // records, bandwidths, EC weights, distances, densities, work lists, and
    // the bin numbers and radix-sort storage that build the bin order
    b += 128 * n + 28 * nbins;
    const bool loo = hoption == 1 || hoption == MKDE_H_SEARCH || (hoption == 2 && optimise);
    // saved bandwidths / EC weights and the per-candidate coefficient and
    // density arrays of the batched leave-one-out pass
    if (loo) b += 148 * n;
    // mask, valid flags, boundary flags and list, row runs, density output;
    // per row and column: run pointers and counts, cell-centre tables
    b += 24 * ncols * nrows + 32 * (nrows + ncols);
    // FP64 partials of split density blocks (8 B per cell): at most twice the
    // split target and 2 * cost / 2048 + 1 per block (cost <= n)
    const Tuning tune = mkde_dev::default_tuning();
    const double nblk = (double)((ncols + MKDE_DT_BW - 1) / MKDE_DT_BW) * (double)((nrows + 63) / 64);
    b += (int64_t)std::min(2.0 * (double)tune.dt_target, 2.0 * (double)n * nblk / 2048.0 + nblk) *
         (8 * MKDE_DT_CELLS);
    // work lists of the point kernels at their largest (point_work_bounds):
    // items (20 B each) and FP64 partials of split 64-query waves, 512 B per
    // part for densities at points and 8 KB per part for a 16-candidate pass
    // (a bandwidth search sizes them so for its captured graph)
    int64_t items = 0, slots = 0;
    point_work_bounds(n, tune.pts_target, &items, &slots);
    b += items * 20 + slots * 512;
    if (loo) {
        point_work_bounds(n, tune.loo_target, &items, &slots);
        b += items * 20 + slots * 8192;
    }
    return b;
}

static int check_job(const mkde_serve_job *j)
{
    if (!j) return sfail(MKDE_E_ARG, "job is NULL");
    if (j->n < 0 || (j->n > 0 && (!j->x || !j->y))) return sfail(MKDE_E_ARG, "job has no point coordinates");
    if (!j->out) return sfail(MKDE_E_ARG, "job has no output raster");
    if (j->grid.ncols <= 0 || j->grid.nrows <= 0 || !(j->grid.cellsize > 0.0f))
        return sfail(MKDE_E_ARG, "job raster geometry is invalid");
    if (j->hoption < -1 || j->hoption > MKDE_H_SEARCH) return sfail(MKDE_E_ARG, "hoption %d unknown", j->hoption);
    if (j->hoption == -1 && !j->h) return sfail(MKDE_E_ARG, "hoption -1 needs bandwidths");
    if (j->edge_correction && !j->mask) return sfail(MKDE_E_ARG, "edge correction needs a mask raster");
    if (j->hoption == MKDE_H_SEARCH) {
        if (j->n_candidates < 1 || j->n_candidates > 1024)
            return sfail(MKDE_E_RANGE, "n_candidates %d not in [1, 1024]", j->n_candidates);
        if (!(j->cand_lo > 0.0f) || !(j->cand_hi >= j->cand_lo) || !std::isfinite(j->cand_hi))
            return sfail(MKDE_E_ARG, "candidate range [%g, %g] invalid", j->cand_lo, j->cand_hi);
    }
    return MKDE_OK;
}

// One job on a slot's context: the sequence of the reference's GPU block
// (tiled:877-1517) for one raster that is both the EC mask and the density
// tile, or the candidate search then the density at the selected bandwidth.
static int run_job(mkde_ctx *c, const mkde_serve_job &j, mkde_serve_result &r)
{
    SCHK(mkde_points_upload(c, j.n, j.x, j.y, j.w));
    if (j.mask) {
        SCHK(mkde_mask_upload(c, &j.grid, j.mask));
        SCHK(mkde_tile_from_mask(c));
    } else {
        SCHK(mkde_tile_upload(c, &j.grid, nullptr));
    }
    if (j.hoption == -1) SCHK(mkde_set_bandwidths(c, j.h));
    mkde_job_stats st{};
    if (j.hoption == MKDE_H_SEARCH) {
        float h0 = 0.0f;
        SCHK(mkde_rule_of_thumb(c, &h0));
        const int K = j.n_candidates;
        std::vector<float> hc(K);
        std::vector<double> L(K);
        const double lo = j.cand_lo, ratio = (double)j.cand_hi / (double)j.cand_lo;
        for (int k = 0; k < K; ++k)
            hc[k] = (float)((double)h0 * lo * std::pow(ratio, K > 1 ? (double)k / (double)(K - 1) : 0.0));
        if (j.one_candidate_per_pass) {
            for (int k = 0; k < K; ++k) SCHK(mkde_loo_loglik(c, 1, &hc[k], j.edge_correction, &L[k]));
        } else {
            SCHK(mkde_loo_loglik(c, K, hc.data(), j.edge_correction, L.data()));
        }
        int best = 0;
        for (int k = 1; k < K; ++k)
            if (L[k] > L[best]) best = k;
        SCHK(mkde_set_bandwidth_fixed(c, hc[best]));
        mkde_job_opts o{-1, j.edge_correction, 0, 0.0f, 0};
        SCHK(mkde_run_job(c, &o, &st));
        r.h0 = h0;
        r.h_sel = hc[best];
        r.alpha_sel = 0.0f;
    } else {
        mkde_job_opts o{j.hoption, j.edge_correction, j.optimise, j.alpha, 0};
        SCHK(mkde_run_job(c, &o, &st));
        r.h0 = st.h0;
        r.h_sel = st.h_opt;
        r.alpha_sel = st.alpha_opt;
    }
    return mkde_tile_download(c, j.out);
}

static void worker(mkde_server *s, int i)
{
    Slot &me = s->slots[i];
    mkde_ctx *c = nullptr;
    const int st = mkde_ctx_create(s->device, &c);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (st != MKDE_OK && s->init_status == MKDE_OK) {
            s->init_status = st;
            s->init_msg = mkde_last_error();
        }
        me.ctx = c;
        s->ready++;
    }
    s->cv_done.notify_all();
    if (st != MKDE_OK) return;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        if (s->stop) break;
        Job *job = nullptr;
        int64_t ticket = 0;
        if (!s->queue.empty()) {
            const int64_t room = s->budget - s->held_total - s->committed_running;
            // the head if it fits, else the first later job that fits while
            // the head has been passed fewer than MKDE_SKIP_MAX times
            size_t pick = s->queue.size();
            if (s->jobs.at(s->queue.front()).bound <= room) {
                pick = 0;
            } else if (s->head_skips < MKDE_SKIP_MAX) {
                for (size_t q = 1; q < s->queue.size(); ++q)
                    if (s->jobs.at(s->queue[q]).bound <= room) {
                        pick = q;
                        break;
                    }
            }
            if (pick < s->queue.size()) {
                ticket = s->queue[pick];
                Job &j = s->jobs.at(ticket);
                s->queue.erase(s->queue.begin() + (std::ptrdiff_t)pick);
                s->head_skips = pick == 0 ? 0 : s->head_skips + 1;
                j.state = RUNNING;
                j.t_start = Clock::now();
                j.res.slot = i;
                s->committed_running += j.bound;
                s->running++;
                s->peak = std::max(s->peak, s->held_total + s->committed_running);
                job = &j;
            } else if (s->committed_running == 0) {
                if (me.held > 0) {
                    // nothing runs and the next job does not fit: drop this slot's buffers
                    lk.unlock();
                    (void)mkde_ctx_release(c);
                    int64_t now = 0;
                    (void)mkde_ctx_memory(c, &now);
                    lk.lock();
                    s->held_total += now - me.held;
                    me.held = now;
                    s->releases++;
                    continue;
                }
                s->cv_work.notify_all();   // let the other idle slots release theirs
            }
        }
        if (!job) {
            s->cv_work.wait(lk);
            continue;
        }
        lk.unlock();
        mkde_serve_result r = job->res;
        const int rc = run_job(c, job->d, r);
        r.status = rc;
        if (rc != MKDE_OK) {
            snprintf(r.message, sizeof r.message, "%s", mkde_last_error());
            (void)mkde_synchronize(c);
        }
        int64_t now = 0;
        (void)mkde_ctx_memory(c, &now);
        lk.lock();
        job->t_end = Clock::now();
        r.queue_ms = ms_between(job->t_submit, job->t_start);
        r.run_ms = ms_between(job->t_start, job->t_end);
        r.latency_ms = ms_between(job->t_submit, job->t_end);
        job->res = r;
        job->state = DONE;
        s->held_total += now - me.held;
        me.held = now;
        s->committed_running -= job->bound;
        s->running--;
        s->finished++;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    lk.unlock();
    (void)mkde_ctx_destroy(c);
}

extern "C" int mkde_server_create(int device, int slots, int64_t hbm_budget, mkde_server **out)
{
    if (!out) return sfail(MKDE_E_ARG, "out is NULL");
    *out = nullptr;
    if (slots < 1 || slots > 64) return sfail(MKDE_E_RANGE, "slots %d not in [1, 64]", slots);
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return sfail(MKDE_E_HIP, "no HIP device available");
    if (device < 0 || device >= ndev) return sfail(MKDE_E_ARG, "device %d out of range [0,%d)", device, ndev);
    if (hbm_budget <= 0) {
        int prev = 0;
        (void)hipGetDevice(&prev);
        size_t fr = 0, tot = 0;
        e = hipSetDevice(device);
        if (e == hipSuccess) e = hipMemGetInfo(&fr, &tot);
        (void)hipSetDevice(prev);
        if (e != hipSuccess) return sfail(MKDE_E_HIP, "hipMemGetInfo failed: %s", hipGetErrorString(e));
        hbm_budget = (int64_t)((double)fr * 0.8);
    }
    mkde_server *s = new (std::nothrow) mkde_server();
    if (!s) return sfail(MKDE_E_NOMEM, "out of host memory");
    s->device = device;
    s->budget = hbm_budget;
    s->slots.resize(slots);
    for (int i = 0; i < slots; ++i) s->slots[i].th = std::thread(worker, s, i);
    int st;
    std::string msg;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->ready == slots; });
        st = s->init_status;
        msg = s->init_msg;
    }
    if (st != MKDE_OK) {
        (void)mkde_server_destroy(s);
        return mkde_internal_fail(st, msg.c_str());
    }
    *out = s;
    return MKDE_OK;
}

extern "C" int mkde_server_submit(mkde_server *s, const mkde_serve_job *job, int64_t *ticket)
{
    if (!s || !ticket) return sfail(MKDE_E_ARG, "null argument");
    SCHK(check_job(job));
    const int64_t bound = mkde_job_footprint(job->n, job->grid.ncols, job->grid.nrows, job->hoption, job->optimise,
                                             job->n_candidates);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return sfail(MKDE_E_STATE, "server is shutting down");
    if (bound > s->budget) {
        s->refused++;
        return sfail(MKDE_E_NOMEM, "job needs up to %lld device bytes, over the server budget of %lld",
                     (long long)bound, (long long)s->budget);
    }
    const int64_t t = s->next_ticket++;
    Job &j = s->jobs[t];
    j.d = *job;
    j.bound = bound;
    j.t_submit = Clock::now();
    j.res.status = MKDE_E_PENDING;
    j.res.slot = -1;
    j.res.footprint = bound;
    s->queue.push_back(t);
    s->submitted++;
    *ticket = t;
    s->cv_work.notify_all();
    return MKDE_OK;
}

extern "C" int mkde_server_wait(mkde_server *s, int64_t ticket, double timeout_ms, mkde_serve_result *res)
{
    if (!s || !res) return sfail(MKDE_E_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->jobs.find(ticket) == s->jobs.end()) return sfail(MKDE_E_ARG, "unknown ticket %lld", (long long)ticket);
    // the job is looked up again after every wake: another thread waiting on
    // the same ticket may have retired it meanwhile
    auto done = [&] {
        auto f = s->jobs.find(ticket);
        return f == s->jobs.end() || f->second.state == DONE;
    };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), done)) {
        return MKDE_E_PENDING;
    }
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end())
        return sfail(MKDE_E_ARG, "ticket %lld was retired by another waiter", (long long)ticket);
    *res = it->second.res;
    s->jobs.erase(it);
    s->cv_done.notify_all();   // a second waiter on this ticket wakes and reports it retired
    return MKDE_OK;
}

extern "C" int mkde_server_cancel(mkde_server *s, int64_t ticket)
{
    if (!s) return sfail(MKDE_E_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return sfail(MKDE_E_ARG, "unknown ticket %lld", (long long)ticket);
    Job &j = it->second;
    if (j.state == RUNNING) return sfail(MKDE_E_STATE, "job %lld is running", (long long)ticket);
    if (j.state == QUEUED) {
        auto qit = std::find(s->queue.begin(), s->queue.end(), ticket);
        if (qit == s->queue.begin()) s->head_skips = 0;   // a new head
        s->queue.erase(qit);
        j.state = DONE;
        j.t_end = Clock::now();
        j.res.status = MKDE_E_CANCELLED;
        snprintf(j.res.message, sizeof j.res.message, "cancelled before it started");
        j.res.latency_ms = ms_between(j.t_submit, j.t_end);
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    return MKDE_OK;
}

extern "C" int mkde_server_info_get(mkde_server *s, mkde_server_info *info)
{
    if (!s || !info) return sfail(MKDE_E_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    info->slots = (int)s->slots.size();
    info->budget = s->budget;
    info->held = s->held_total;
    info->peak_committed = s->peak;
    info->submitted = s->submitted;
    info->finished = s->finished;
    info->refused = s->refused;
    info->releases = s->releases;
    info->queued = (int64_t)s->queue.size();
    info->running = s->running;
    return MKDE_OK;
}

extern "C" int mkde_server_destroy(mkde_server *s)
{
    if (!s) return MKDE_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (int64_t t : s->queue) {
            Job &j = s->jobs.at(t);
            j.state = DONE;
            j.res.status = MKDE_E_CANCELLED;
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (Slot &sl : s->slots)
        if (sl.th.joinable()) sl.th.join();
    delete s;
    return MKDE_OK;
}
