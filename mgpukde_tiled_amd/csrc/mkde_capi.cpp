// mkde_capi.cpp -- host side of libmkde: device context, the cell-binned point
// index, the job drivers behind the C ABI of include/mkde.h.
//
// The reference drives its GPU path from main() through host functions that
// exit() on every error and stage every intermediate through the host
// (Reform* round trips, tiled:4297-4646).  Here all state stays resident on
// one device in a context; host copies happen only where the caller asks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <functional>
#include <map>
#include <mutex>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "mkde.h"
#include "mkde_device.h"

namespace mkde_dev {
hipError_t prep_points(hipStream_t, int64_t, const int *, const float *, const float *, const float *, const float *,
                       const float *, float, PtRec *);
hipError_t bin_stats(hipStream_t, int, const int *, const PtRec *, BinInfo *, double *);
int64_t raster_blocks(const GridDesc &, const Tiling &);
hipError_t density_plan(hipStream_t, const GridDesc &, const float *, const PointIndex &, const Tiling &,
                        const DensityWork &, const Tuning &);
hipError_t density_items(hipStream_t, int, const DensityWork &, int, int4 *, int4 *, int *, const unsigned *);
hipError_t density_tile(hipStream_t, const GridDesc &, const float *, float *, const PointIndex &,
                        unsigned long long *, const int *, const int4 *, const unsigned *, double *, int,
                        const Tiling &);
hipError_t density_merge(hipStream_t, const GridDesc &, const Tiling &, const float *, int, const int4 *,
                         const double *, float *);
hipError_t point_plan(hipStream_t, const PointIndex &, int, float, const double *, int, const DensityWork &,
                      const Tuning &);
hipError_t point_density(hipStream_t, const PointIndex &, int, float *, float, const int *, const int4 *,
                         const unsigned *, double *, int);
hipError_t loo_multi(hipStream_t, const PointIndex &, const float *, const double *, const double *, int, float *,
                     const int *, const int4 *, const unsigned *, double *, int, const unsigned *);
hipError_t point_merge(hipStream_t, int64_t, int, int, int, const int4 *, const double *, float *, const unsigned *);
hipError_t range_count(hipStream_t, const PointIndex &, int64_t *);
hipError_t range_fill(hipStream_t, const PointIndex &, const int64_t *, int32_t *);
hipError_t bin_sort_bytes(int64_t, int, size_t *);
hipError_t bin_points(hipStream_t, int64_t, const float *, const float *, const BinGrid &, int, unsigned *, unsigned *,
                      int *, int *, void *, size_t, float *, float *, int *);
hipError_t valid_cells(hipStream_t, int64_t, const float *, float, uint8_t *);
hipError_t runs_count(hipStream_t, const GridDesc &, const uint8_t *, unsigned *, unsigned *, unsigned *, unsigned *,
                      int64_t *);
hipError_t runs_fill(hipStream_t, const GridDesc &, const uint8_t *, const unsigned *, int2 *);
hipError_t boundary(hipStream_t, const GridDesc &, const uint8_t *, uint8_t *, float2 *, unsigned long long *);
hipError_t dist_boundary(hipStream_t, int64_t, const float *, const float *, const float2 *, int64_t, double *);
hipError_t bt_build(hipStream_t, const GridDesc &, const uint8_t *, uint4 *, float4 *, const BndTree &);
hipError_t dist_tree(hipStream_t, int64_t, const float *, const float *, const int *, const GridDesc &,
                     const BndTree &, double *);
hipError_t ec_active(hipStream_t, int64_t, const GridDesc &, const float *, const float *, const double *,
                     const float *, float, const float *, float *, int *, int *, int, int);
int ec_wide_max();
int ec_box_cols(const GridDesc &, float);
Tuning default_tuning();
hipError_t centres(hipStream_t, const GridDesc &, float *, float *);
hipError_t ec(hipStream_t, int64_t, int, int, int, const GridDesc &, const MaskRuns &, const float *, const float *,
              const float *, const float *, const float *, float, const int *, float *, unsigned long long *, int,
              int);
hipError_t ec_dev(hipStream_t, int64_t, const GridDesc &, const MaskRuns &, const float *, const float *, const float *,
                  const float *, const float *, const int *, const int *, float *, int, int, int, int);
hipError_t count_pairs(hipStream_t, const GridDesc &, const MaskRuns &, const PointIndex &, unsigned long long *);
int64_t slot_words();
hipError_t fold_slots(hipStream_t, unsigned long long *);
hipError_t reach(hipStream_t, int, const double *, double *);
hipError_t gml(hipStream_t, const double *, int64_t, double *);
hipError_t sumlog(hipStream_t, int64_t, int, const float *, int, double *, int, double *);
hipError_t abramson(hipStream_t, int64_t, const int *, const float *, const double *, float, float, float *);
hipError_t unsort_f32(hipStream_t, int64_t, const int *, const float *, float *);
hipError_t sort_f32(hipStream_t, int64_t, const int *, const float *, float *);
hipError_t coef_k(hipStream_t, int64_t, const int *, const float *, const float *, double, float *, int);
hipError_t coef_k_dev(hipStream_t, int64_t, const int *, const float *, const float *, const float *, float *, int);
hipError_t fill_f32(hipStream_t, int64_t, float, float *);
hipError_t scratch_address(int, void **);
}  // namespace mkde_dev

// ------------------------------------------------------------ errors ----

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                                  \
    do {                                                                                              \
        hipError_t _e = (expr);                                                                       \
        if (_e != hipSuccess) return fail(MKDE_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

#define CHK(expr)                    \
    do {                             \
        int _s = (expr);             \
        if (_s != MKDE_OK) return _s; \
    } while (0)

// Device bytes held by the context whose entry point runs on this thread
// (set by set_device); DevBuf allocations are charged to it.
static thread_local std::atomic<int64_t> *t_acct = nullptr;

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    std::atomic<int64_t> *acct = nullptr;   // the owning context's device-byte counter
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) {
            (void)hipFree(p);
            if (acct) *acct -= (int64_t)(n * sizeof(T));
        }
        p = nullptr;
        n = 0;
    }
    int ensure(size_t cnt)
    {
        if (cnt == 0) cnt = 1;
        if (p && cnt <= n) return MKDE_OK;
        release();
        hipError_t e = hipMalloc((void **)&p, cnt * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MKDE_E_NOMEM, "hipMalloc(%zu bytes) failed: %s", cnt * sizeof(T), hipGetErrorString(e));
        }
        n = cnt;
        acct = t_acct;
        if (acct) *acct += (int64_t)(cnt * sizeof(T));
        return MKDE_OK;
    }
};

// -------------------------------------------------------------- context --

struct RunsHost {
    std::vector<int64_t> ptr;
    std::vector<int2> se;
};

struct mkde_ctx {
    int device = 0;
    std::atomic<int64_t> dev_bytes{0};   // device memory held by the buffers below
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // the device's static scratch array (allocation-free entry point)
    unsigned char *scratch = nullptr;
    // events of a job's phase timers, created once and reused by every job
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    float last_density_ms = 0.0f, last_ec_ms = 0.0f, last_points_ms = 0.0f;

    // points, original order (host copies kept for the rule of thumb)
    int64_t n = 0;
    std::vector<float> hx, hy, hw;
    DevBuf<float> x, y, w;

    // cell-binned index
    BinGrid bg{};
    int nbins = 0;
    std::vector<int> hperm;     // host copy of perm (host_perm), empty until needed
    DevBuf<int> perm;           // bin-order position -> original index
    DevBuf<unsigned> bin_key, bin_skey;   // building the bin order
    DevBuf<int> bin_idx;
    DevBuf<uint8_t> bin_tmp;
    DevBuf<int> bin_start;      // nbins + 1
    DevBuf<float> xs, ys;       // bin order
    DevBuf<PtRec> pts;          // bin order, per-job constants
    DevBuf<BinInfo> bin_box;
    DevBuf<double> bin_thr;
    DevBuf<double> reach;       // 1 value
    bool pts_dirty = true;
    Tuning tune{};              // launch shapes (mkde_set_option, autotuner)

    // per-job arrays, original order
    DevBuf<float> h, ew;
    DevBuf<double> dist;
    bool h_set = false;
    float h_fixed = 0.0f;       // > 0: every point has this bandwidth
    bool dist_set = false;

    // edge-correction mask
    bool mask_set = false;
    GridDesc mg{};
    int mg_cx32 = 0;   // the mask's column centres are exact in FP32 (grid_cx_fp32)
    DevBuf<float> mask;
    DevBuf<uint8_t> valid;
    DevBuf<int64_t> run_ptr;
    DevBuf<int2> run_se;
    DevBuf<unsigned> run_cnt, run_first, run_chunks, run_total;   // building the runs
    DevBuf<float> colx, rowy;   // cell-centre tables of the mask grid
    DevBuf<uint8_t> bflags;
    DevBuf<float2> blist;
    DevBuf<unsigned long long> bcount;   // counter slots, or one counter with the list
    bool blist_on = false;               // boundary centre list built (brute-force distances)
    bool bflags_ok = false;    // boundary flags / list enqueued for the current mask
    int64_t nboundary = -1;    // boundary cell count once read back
    // boundary hierarchy (nearest-boundary search)
    DevBuf<uint4> bt_bits;
    DevBuf<float4> bt_box;
    BndTree bt{};
    bool bt_ok = false;

    // tile
    bool tile_set = false;
    bool tile_has_mask = false;
    bool tile_is_ec_mask = false;   // the density tile is the resident EC mask raster
    GridDesc tg{};
    Tiling tl{};   // tl.tw == 0: tg is one tile; else tg is a raster of tiles
    DevBuf<float> tmask, tout;

    // scratch
    DevBuf<float> den, den2;
    DevBuf<int> list;
    DevBuf<int> icount;
    DevBuf<unsigned long long> ucount;
    DevBuf<double> partial, dsum, gval;
    DevBuf<float> coefk, ewk, den1k;
    DevBuf<double> ak, tk;
    DevBuf<float> hsave, ewsave;   // bandwidths / EC weights saved across a likelihood scan
    // the batched bandwidth search as one captured graph (mkde_loo_loglik):
    // candidate bandwidths, slot constants and log-likelihoods on the device,
    // the graph and the key of everything it captured
    bool search_graph_on = true;
    DevBuf<float> s_h;          // K candidates, ascending
    DevBuf<double> s_a, s_t;    // 16 slot constants per pass
    DevBuf<double> s_logl;      // K
    hipGraphExec_t s_exec = nullptr;
    std::vector<uint64_t> s_key;
    // work list of the point kernels (64-query waves, heavy ones split)
    DevBuf<unsigned> pw_cost, pw_parts, pw_item0, pw_slot0, pw_sidx, pw_sidx2, pw_chunks, pw_counts, pw_hist;
    DevBuf<unsigned long long> pw_total;
    DevBuf<int4> pw_items, pw_splits;
    DevBuf<int> pw_order;
    DevBuf<double> pw_partials;
    // density kernel counters (mkde_density_stats)
    bool dt_stats_on = false;
    DevBuf<unsigned long long> dt_stats;
    // work list of the density kernel: heavy blocks split into parts,
    // items visited heaviest first
    bool dt_order_on = true;
    int pts_per_bin = 64;   // target points per bin of the point index
    DevBuf<unsigned> blk_cost, blk_parts, blk_item0, blk_slot0, blk_sidx, blk_sidx2, blk_chunks, blk_counts,
        blk_hist;
    DevBuf<unsigned long long> blk_total;
    DevBuf<int4> dt_items, dt_splits;
    DevBuf<int> dt_order;
    DevBuf<double> dt_partials;
};

extern "C" const char *mkde_last_error(void) { return g_err.c_str(); }

// error channel shared with the I/O translation unit
extern "C" __attribute__((visibility("hidden"))) int mkde_internal_fail(int code, const char *msg)
{
    g_err = msg ? msg : "";
    return code;
}

extern "C" const char *mkde_version(void) { return "mkde 0.1 (gfx950)"; }

static int set_device(mkde_ctx *c)
{
    t_acct = &c->dev_bytes;
    HIPCHK(hipSetDevice(c->device));
    return MKDE_OK;
}

extern "C" int mkde_ctx_create(int device, mkde_ctx **out)
{
    if (!out) return fail(MKDE_E_ARG, "mkde_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) return fail(MKDE_E_HIP, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= ndev) return fail(MKDE_E_ARG, "device %d out of range [0,%d)", device, ndev);
    mkde_ctx *c = new (std::nothrow) mkde_ctx();
    if (!c) return fail(MKDE_E_NOMEM, "out of host memory");
    c->device = device;
    c->tune = mkde_dev::default_tuning();
    // tuning knob: MKDE_DENSITY_ORDER=0 visits density blocks in raster order
    const char *ord = getenv("MKDE_DENSITY_ORDER");
    if (ord && ord[0] == '0') c->dt_order_on = false;
    const char *ppb = getenv("MKDE_PTS_PER_BIN");
    if (ppb && atoi(ppb) > 0) c->pts_per_bin = atoi(ppb);
    const char *sg = getenv("MKDE_SEARCH_GRAPH");
    if (sg && sg[0] == '0') c->search_graph_on = false;
    e = hipSetDevice(device);
    // the static scratch array's address on this device, resolved once per
    // device with the device current
    if (e == hipSuccess) e = mkde_dev::scratch_address(device, (void **)&c->scratch);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    if (e != hipSuccess) {
        delete c;
        return fail(MKDE_E_HIP, "context set-up failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return MKDE_OK;
}

extern "C" int mkde_ctx_destroy(mkde_ctx *c)
{
    if (!c) return MKDE_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->s_exec) (void)hipGraphExecDestroy(c->s_exec);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    t_acct = nullptr;
    return MKDE_OK;
}

extern "C" int mkde_ctx_memory(mkde_ctx *c, int64_t *bytes)
{
    if (!c || !bytes) return fail(MKDE_E_ARG, "null argument");
    *bytes = c->dev_bytes.load();
    return MKDE_OK;
}

extern "C" int mkde_ctx_release(mkde_ctx *c)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(set_device(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    // keep the stream, events and settings; drop every buffer and all state
    const int dev = c->device;
    hipStream_t s = c->stream;
    hipEvent_t e0 = c->ev0, e1 = c->ev1;
    unsigned char *scr = c->scratch;
    std::vector<hipEvent_t> pool;
    pool.swap(c->ev_pool);
    const bool order_on = c->dt_order_on, graph_on = c->search_graph_on;
    const Tuning tune = c->tune;
    const int ppb = c->pts_per_bin;
    if (c->s_exec) (void)hipGraphExecDestroy(c->s_exec);
    c->s_exec = nullptr;
    c->~mkde_ctx();
    new (c) mkde_ctx();
    c->device = dev;
    c->stream = s;
    c->ev0 = e0;
    c->ev1 = e1;
    c->scratch = scr;
    c->ev_pool.swap(pool);
    c->dt_order_on = order_on;
    c->search_graph_on = graph_on;
    c->tune = tune;
    c->pts_per_bin = ppb;
    t_acct = &c->dev_bytes;
    return MKDE_OK;
}

extern "C" int mkde_synchronize(mkde_ctx *c)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(set_device(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MKDE_OK;
}

static int copy_h2d(mkde_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return MKDE_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MKDE_OK;
}

static int copy_d2h(mkde_ctx *c, void *dst, const void *src, size_t bytes)
{
    if (bytes == 0) return MKDE_OK;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MKDE_OK;
}

// ---------------------------------------------------------- point index --
//
// Uniform bins over the points' bounding box, about pts_per_bin points per
// bin, at most 4M bins.  The grid is sized here; each point's bin, the bin
// order (a stable sort: ties keep the caller's order, so every downstream
// summation order is deterministic), the bin starts and the bin-order
// coordinates are built on the device (mkde_bins.hip).

static BinGrid bin_grid(int64_t n, float x0, float x1, float y0, float y1, int pts_per_bin)
{
    if (n == 0) { x0 = y0 = 0.0f; x1 = y1 = 1.0f; }
    double wx = std::max((double)x1 - x0, 1e-6), wy = std::max((double)y1 - y0, 1e-6);
    double target = std::max(1.0, (double)n / (double)pts_per_bin);
    double size = std::sqrt(wx * wy / target);
    size = std::max(size, std::max(wx, wy) / 4096.0);
    int nbx, nby;
    for (;;) {
        nbx = (int)std::floor(wx / size) + 1;
        nby = (int)std::floor(wy / size) + 1;
        if ((int64_t)nbx * nby <= (4 << 20)) break;
        size *= 1.25;
    }
    return BinGrid{(double)x0, (double)y0, size, nbx, nby};
}

// Drop every per-point state of the context (a failed upload leaves it
// without points, so later calls report MKDE_E_STATE instead of using
// buffers of another size).
static void forget_points(mkde_ctx *c)
{
    c->n = 0;
    c->h_set = false;
    c->h_fixed = 0.0f;
    c->dist_set = false;
    c->pts_dirty = true;
}

extern "C" int mkde_points_upload(mkde_ctx *c, int64_t n, const float *x, const float *y, const float *w)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!x || !y))) return fail(MKDE_E_ARG, "mkde_points_upload: bad arguments");
    if (n > (int64_t)1 << 30) return fail(MKDE_E_RANGE, "too many points (%lld > 2^30)", (long long)n);
    // validate (and take the bounding box) before anything of the context
    // changes: on this error the previous points stay usable
    float x0 = FLT_MAX, x1 = -FLT_MAX, y0 = FLT_MAX, y1 = -FLT_MAX;
    for (int64_t i = 0; i < n; ++i) {
        if (!std::isfinite(x[i]) || !std::isfinite(y[i]))
            return fail(MKDE_E_ARG, "point %lld has a non-finite coordinate", (long long)i);
        x0 = std::min(x0, x[i]);
        x1 = std::max(x1, x[i]);
        y0 = std::min(y0, y[i]);
        y1 = std::max(y1, y[i]);
    }
    CHK(set_device(c));
    const BinGrid bg = bin_grid(n, x0, x1, y0, y1, c->pts_per_bin);
    const int nbins = bg.nbx * bg.nby;
    // host copies (the rule of thumb runs on the host in point order); the
    // context takes the new points only when all of the upload succeeded
    std::vector<float> hx, hy, hw;
    try {
        hx.assign(x, x + n);
        hy.assign(y, y + n);
        if (w) hw.assign(w, w + n);
        else hw.assign(n, 1.0f);
    } catch (const std::bad_alloc &) {
        return fail(MKDE_E_NOMEM, "out of host memory");
    }
    size_t sort_bytes = 0;
    HIPCHK(mkde_dev::bin_sort_bytes(n, nbins, &sort_bytes));
    forget_points(c);   // from here on a failure leaves the context without points
    HIPCHK(hipStreamSynchronize(c->stream));   // no queued work still reads the old buffers
    CHK(c->x.ensure(n));
    CHK(c->y.ensure(n));
    CHK(c->w.ensure(n));
    CHK(c->xs.ensure(n));
    CHK(c->ys.ensure(n));
    CHK(c->perm.ensure(n));
    CHK(c->bin_key.ensure(n));
    CHK(c->bin_skey.ensure(n));
    CHK(c->bin_idx.ensure(n));
    CHK(c->bin_tmp.ensure(std::max<size_t>(sort_bytes, 1)));
    CHK(c->bin_start.ensure(nbins + 1));
    CHK(c->pts.ensure(n));
    CHK(c->bin_box.ensure(nbins));
    CHK(c->bin_thr.ensure(nbins));
    CHK(c->reach.ensure(1));
    CHK(c->h.ensure(n));
    CHK(c->ew.ensure(n));
    CHK(c->dist.ensure(n));
    CHK(c->den.ensure(n));
    CHK(copy_h2d(c, c->x.p, hx.data(), n * sizeof(float)));
    CHK(copy_h2d(c, c->y.p, hy.data(), n * sizeof(float)));
    CHK(copy_h2d(c, c->w.p, hw.data(), n * sizeof(float)));
    HIPCHK(mkde_dev::bin_points(c->stream, n, c->x.p, c->y.p, bg, nbins, c->bin_key.p, c->bin_skey.p, c->bin_idx.p,
                                c->perm.p, c->bin_tmp.p, sort_bytes, c->xs.p, c->ys.p, c->bin_start.p));
    HIPCHK(mkde_dev::fill_f32(c->stream, n, 1.0f, c->ew.p));
    HIPCHK(mkde_dev::fill_f32(c->stream, n, 0.0f, c->h.p));
    HIPCHK(hipStreamSynchronize(c->stream));
    // commit
    c->hx.swap(hx);
    c->hy.swap(hy);
    c->hw.swap(hw);
    c->hperm.clear();   // read back from the device when needed (host_perm)
    c->bg = bg;
    c->nbins = nbins;
    c->n = n;
    return MKDE_OK;
}

// The bin order's permutation on the host (bin-order position -> caller's
// index), read back once per upload.
static int host_perm(mkde_ctx *c)
{
    if ((int64_t)c->hperm.size() == c->n) return MKDE_OK;
    try {
        c->hperm.resize(c->n);
    } catch (const std::bad_alloc &) {
        return fail(MKDE_E_NOMEM, "out of host memory");
    }
    return copy_d2h(c, c->hperm.data(), c->perm.p, c->n * sizeof(int));
}

extern "C" int mkde_num_points(mkde_ctx *c, int64_t *n)
{
    if (!c || !n) return fail(MKDE_E_ARG, "null argument");
    *n = c->n;
    return MKDE_OK;
}

// Rebuild the bin-order point records after bandwidths or EC weights changed.
static int prep(mkde_ctx *c)
{
    if (!c->pts_dirty) return MKDE_OK;
    HIPCHK(mkde_dev::prep_points(c->stream, c->n, c->perm.p, c->xs.p, c->ys.p, c->w.p, c->h.p, c->ew.p, c->h_fixed,
                                 c->pts.p));
    HIPCHK(mkde_dev::bin_stats(c->stream, c->nbins, c->bin_start.p, c->pts.p, c->bin_box.p, c->bin_thr.p));
    HIPCHK(mkde_dev::reach(c->stream, c->nbins, c->bin_thr.p, c->reach.p));
    c->pts_dirty = false;
    return MKDE_OK;
}

static PointIndex index_of(mkde_ctx *c)
{
    PointIndex I;
    I.pts = c->pts.p;
    I.bin_start = c->bin_start.p;
    I.bin_box = c->bin_box.p;
    I.bin_thr = c->bin_thr.p;
    I.bg = c->bg;
    I.reach_p = c->reach.p;
    I.n = c->n;
    return I;
}

static int need_points(mkde_ctx *c)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    if (c->n <= 0 || !c->x.p) return fail(MKDE_E_STATE, "no points uploaded");
    return set_device(c);
}

// ------------------------------------------------------- rule of thumb --
// MeanCenter (tiled:1563-1578): float sums in point order; StandardDistance2
// (tiled:1581-1595): double sums of w * Distance2; BandWidth2 (tiled:64-68).
// Separate roundings on purpose: no contraction into fused multiply-adds.
#pragma clang fp contract(off)
static float rule_of_thumb_host(const std::vector<float> &x, const std::vector<float> &y, const std::vector<float> &w)
{
    const size_t n = x.size();
    float sum_x = 0.0f, sum_y = 0.0f, sum_w = 0.0f;
    for (size_t p = 0; p < n; ++p) {
        float wp = w[p];
        sum_w = sum_w + wp;
        float tx = wp * x[p];
        sum_x = sum_x + tx;
        float ty = wp * y[p];
        sum_y = sum_y + ty;
    }
    float mx = sum_x / sum_w, my = sum_y / sum_w;
    double sum2 = 0.0, sw = 0.0;
    for (size_t p = 0; p < n; ++p) {
        float wp = w[p];
        sw += wp;
        double dx = (double)(float)(x[p] - mx);
        double dy = (double)(float)(y[p] - my);
        double dxx = dx * dx;
        double dyy = dy * dy;
        double d2 = dxx + dyy;
        double t = (double)wp * d2;
        sum2 = sum2 + t;
    }
    double d2 = sum2 / sw;
    float f = sqrtf(2.0f / (float)(3u * (unsigned)n));
    double bw2 = (double)f * d2;
    return sqrtf((float)bw2);
}
#pragma clang fp contract(on)

extern "C" int mkde_rule_of_thumb(mkde_ctx *c, float *h0)
{
    CHK(need_points(c));
    if (!h0) return fail(MKDE_E_ARG, "h0 is NULL");
    *h0 = rule_of_thumb_host(c->hx, c->hy, c->hw);
    return MKDE_OK;
}

// ----------------------------------------------------------------- mask --

static void build_runs(const float *v, int64_t ncols, int64_t nrows, float nodata, RunsHost &R)
{
    R.ptr.assign(nrows + 1, 0);
    R.se.clear();
    for (int64_t r = 0; r < nrows; ++r) {
        const float *row = v + r * ncols;
        int64_t c = 0;
        while (c < ncols) {
            while (c < ncols && row[c] == nodata) ++c;
            if (c >= ncols) break;
            int64_t s = c;
            while (c < ncols && row[c] != nodata) ++c;
            R.se.push_back(int2{(int)s, (int)(c - 1)});
        }
        R.ptr[r + 1] = (int64_t)R.se.size();
    }
}

static int check_grid(const mkde_grid *g)
{
    if (!g) return fail(MKDE_E_ARG, "grid is NULL");
    if (g->ncols <= 0 || g->nrows <= 0) return fail(MKDE_E_ARG, "grid has no cells");
    if (g->ncols > INT32_MAX - 1 || g->nrows > INT32_MAX - 1)
        return fail(MKDE_E_RANGE, "grid dimension exceeds 2^31-1");
    if (!(g->cellsize > 0.0f) || !std::isfinite(g->cellsize)) return fail(MKDE_E_ARG, "cellsize must be > 0");
    return MKDE_OK;
}

static GridDesc to_desc(const mkde_grid *g)
{
    return GridDesc{g->ncols, g->nrows, g->xll, g->yll, g->cellsize, g->nodata};
}

// Whether every column centre of grid g, fl32(xll + (c + 0.5) cs) with the
// product and the sum rounded in FP64 (cell_x), equals fl32(fma(c + 0.5, cs,
// xll)) computed in FP32: then the edge-correction kernels use the shorter
// FP32 form.  Checked column by column (c + 0.5 is exact in FP32 below 2^23).
#pragma clang fp contract(off)
static bool grid_cx_fp32(const mkde_grid *g)
{
    if (g->ncols > ((int64_t)1 << 23)) return false;
    for (int64_t col = 0; col < g->ncols; ++col) {
        const double prod = ((double)col + 0.5) * (double)g->cellsize;
        const double sum = (double)g->xll + prod;
        const float f64 = (float)sum;
        const float f32 = std::fma((float)col + 0.5f, g->cellsize, g->xll);
        if (__builtin_bit_cast(uint32_t, f64) != __builtin_bit_cast(uint32_t, f32)) return false;
    }
    return true;
}
#pragma clang fp contract(on)

extern "C" int mkde_mask_upload(mkde_ctx *c, const mkde_grid *g, const float *values)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(check_grid(g));
    if (!values) return fail(MKDE_E_ARG, "mask values are NULL");
    // run indices are 32-bit: a row has at most ceil(ncols / 2) runs
    if ((uint64_t)g->nrows * (uint64_t)((g->ncols + 1) / 2) > UINT32_MAX)
        return fail(MKDE_E_RANGE, "mask of %lld x %lld cells exceeds the run index (2^32 runs)",
                    (long long)g->ncols, (long long)g->nrows);
    CHK(set_device(c));
    const int64_t nc = g->ncols * g->nrows;
    // from here on a failure leaves the context without a mask (and without a
    // density tile that was this mask), never with buffers of another grid
    c->mask_set = false;
    c->bflags_ok = false;
    c->nboundary = -1;
    c->bt_ok = false;
    c->dist_set = false;
    if (c->tile_is_ec_mask) c->tile_set = c->tile_is_ec_mask = false;
    HIPCHK(hipStreamSynchronize(c->stream));   // no queued work still reads the old buffers
    CHK(c->mask.ensure(nc));
    CHK(c->valid.ensure(nc));
    CHK(c->bflags.ensure(nc));
    CHK(c->run_ptr.ensure(g->nrows + 1));
    CHK(c->run_cnt.ensure(g->nrows));
    CHK(c->run_first.ensure(g->nrows));
    CHK(c->run_chunks.ensure((g->nrows + 1023) / 1024));
    CHK(c->run_total.ensure(1));
    CHK(c->bcount.ensure(1));
    CHK(c->colx.ensure(g->ncols));
    CHK(c->rowy.ensure(g->nrows));
    CHK(copy_h2d(c, c->mask.p, values, nc * sizeof(float)));
    const GridDesc gd = to_desc(g);
    // valid flags, then the valid-cell runs of every row, built on the device
    HIPCHK(mkde_dev::valid_cells(c->stream, nc, c->mask.p, g->nodata, c->valid.p));
    HIPCHK(mkde_dev::runs_count(c->stream, gd, c->valid.p, c->run_cnt.p, c->run_first.p, c->run_chunks.p,
                                c->run_total.p, c->run_ptr.p));
    unsigned nruns = 0;
    CHK(copy_d2h(c, &nruns, c->run_total.p, sizeof nruns));
    CHK(c->run_se.ensure(std::max<int64_t>(nruns, 1)));
    HIPCHK(mkde_dev::runs_fill(c->stream, gd, c->valid.p, c->run_first.p, c->run_se.p));
    HIPCHK(mkde_dev::centres(c->stream, gd, c->colx.p, c->rowy.p));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->mg = gd;
    c->mg_cx32 = grid_cx_fp32(g) ? 1 : 0;
    c->mask_set = true;
    return MKDE_OK;
}

static int need_mask(mkde_ctx *c)
{
    if (!c->mask_set) return fail(MKDE_E_STATE, "no edge-correction mask uploaded");
    return MKDE_OK;
}

// Boundary flags and list are enqueued without a host wait; the count is
// read back only when a caller needs it (boundary_count).
// The boundary hierarchy covers rasters up to 16384 leaves (262,144 cells) on
// each side.
static bool btree_fits(const mkde_ctx *c)
{
    const int64_t nx = (c->mg.ncols + MKDE_BT_S - 1) / MKDE_BT_S, ny = (c->mg.nrows + MKDE_BT_S - 1) / MKDE_BT_S;
    return nx <= 16384 && ny <= 16384;
}

static int ensure_boundary(mkde_ctx *c)
{
    CHK(need_mask(c));
    if (c->bflags_ok) return MKDE_OK;
    const int64_t nc = c->mg.ncols * c->mg.nrows;
    // the list of boundary centres only feeds the brute-force distances
    // (rasters too large for the hierarchy); otherwise only the count is
    // kept, in counter slots (one atomic per wave to one address serialises
    // across the device)
    c->blist_on = !btree_fits(c);
    const int64_t nw = mkde_dev::slot_words();
    CHK(c->bcount.ensure(nw));
    HIPCHK(hipMemsetAsync(c->bcount.p, 0, nw * sizeof(unsigned long long), c->stream));
    if (c->blist_on) {
        // upper bound of the boundary list: every valid cell
        CHK(c->blist.ensure(nc));
        HIPCHK(mkde_dev::boundary(c->stream, c->mg, c->valid.p, c->bflags.p, c->blist.p, c->bcount.p));
    } else {
        HIPCHK(mkde_dev::boundary(c->stream, c->mg, c->valid.p, c->bflags.p, nullptr, c->bcount.p));
        HIPCHK(mkde_dev::fold_slots(c->stream, c->bcount.p));
    }
    c->bflags_ok = true;
    c->nboundary = -1;
    return MKDE_OK;
}

static int boundary_count(mkde_ctx *c)
{
    CHK(ensure_boundary(c));
    if (c->nboundary >= 0) return MKDE_OK;
    unsigned long long nb = 0;
    CHK(copy_d2h(c, &nb, c->bcount.p + (c->blist_on ? 0 : mkde_dev::slot_words() - 1), sizeof nb));
    c->nboundary = (int64_t)nb;
    return MKDE_OK;
}

extern "C" int mkde_mark_boundary(mkde_ctx *c, uint8_t *flags, int64_t *n_boundary)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(set_device(c));
    CHK(boundary_count(c));
    if (flags) CHK(copy_d2h(c, flags, c->bflags.p, (size_t)(c->mg.ncols * c->mg.nrows)));
    if (n_boundary) *n_boundary = c->nboundary;
    return MKDE_OK;
}

// Hierarchy over the boundary flags of the EC mask: 16x16-cell leaves up to
// one root.  Rasters wider or taller than 16384 leaves use the brute-force
// kernel instead.
static int ensure_btree(mkde_ctx *c)
{
    if (c->bt_ok) return MKDE_OK;
    BndTree &T = c->bt;
    T = BndTree{};
    if (!btree_fits(c)) return MKDE_OK;
    int nx = (int)((c->mg.ncols + MKDE_BT_S - 1) / MKDE_BT_S), ny = (int)((c->mg.nrows + MKDE_BT_S - 1) / MKDE_BT_S);
    int64_t total = 0;
    for (;;) {
        T.nx[T.nlev] = nx;
        T.ny[T.nlev] = ny;
        T.off[T.nlev] = total;
        total += (int64_t)nx * ny;
        T.nlev++;
        if (nx == 1 && ny == 1) break;
        nx = (nx + 1) / 2;
        ny = (ny + 1) / 2;
    }
    CHK(c->bt_bits.ensure(2 * (size_t)T.nx[0] * T.ny[0]));
    CHK(c->bt_box.ensure(total));
    HIPCHK(mkde_dev::bt_build(c->stream, c->mg, c->bflags.p, c->bt_bits.p, c->bt_box.p, T));
    T.bits = c->bt_bits.p;
    T.box = c->bt_box.p;
    c->bt_ok = true;
    return MKDE_OK;
}

static int ensure_dist(mkde_ctx *c)
{
    if (c->dist_set) return MKDE_OK;
    CHK(ensure_boundary(c));
    CHK(ensure_btree(c));
    if (c->bt.nlev > 0) {
        CHK(prep(c));
        HIPCHK(mkde_dev::dist_tree(c->stream, c->n, c->xs.p, c->ys.p, c->perm.p, c->mg, c->bt, c->dist.p));
    } else {
        CHK(boundary_count(c));
        HIPCHK(mkde_dev::dist_boundary(c->stream, c->n, c->x.p, c->y.p, c->blist.p, c->nboundary, c->dist.p));
    }
    c->dist_set = true;
    return MKDE_OK;
}

extern "C" int mkde_dist_to_boundary(mkde_ctx *c, double *dist)
{
    CHK(need_points(c));
    CHK(ensure_dist(c));
    if (dist) CHK(copy_d2h(c, dist, c->dist.p, c->n * sizeof(double)));
    return MKDE_OK;
}

extern "C" int mkde_set_dist_to_boundary(mkde_ctx *c, const double *dist)
{
    CHK(need_points(c));
    if (!dist) return fail(MKDE_E_ARG, "dist is NULL");
    CHK(copy_h2d(c, c->dist.p, dist, c->n * sizeof(double)));
    c->dist_set = true;
    return MKDE_OK;
}

// ------------------------------------------------------------ bandwidths --

extern "C" int mkde_set_bandwidth_fixed(mkde_ctx *c, float h)
{
    CHK(need_points(c));
    if (!std::isfinite(h)) return fail(MKDE_E_ARG, "bandwidth must be finite");
    HIPCHK(mkde_dev::fill_f32(c->stream, c->n, h, c->h.p));
    c->h_fixed = h > 0.0f ? h : 0.0f;
    c->h_set = true;
    c->pts_dirty = true;
    return MKDE_OK;
}

extern "C" int mkde_set_bandwidths(mkde_ctx *c, const float *h)
{
    CHK(need_points(c));
    if (!h) return fail(MKDE_E_ARG, "h is NULL");
    for (int64_t i = 0; i < c->n; ++i)
        if (!std::isfinite(h[i])) return fail(MKDE_E_ARG, "bandwidth of point %lld is not finite", (long long)i);
    CHK(copy_h2d(c, c->h.p, h, c->n * sizeof(float)));
    c->h_fixed = 0.0f;
    c->h_set = true;
    c->pts_dirty = true;
    return MKDE_OK;
}

extern "C" int mkde_get_bandwidths(mkde_ctx *c, float *h)
{
    CHK(need_points(c));
    if (!h) return fail(MKDE_E_ARG, "h is NULL");
    return copy_d2h(c, h, c->h.p, c->n * sizeof(float));
}

static int need_h(mkde_ctx *c)
{
    if (!c->h_set) return fail(MKDE_E_STATE, "bandwidths are not set");
    return MKDE_OK;
}

// ------------------------------------------------------- edge correction --

// Widest box the wide EC kernel takes: ec_wide_max() columns of LDS, or for
// a fixed bandwidth hf > 0 the bound of its boxes, so that a host-sized and a
// graph-captured pass classify every point alike and the graph's wide launch
// reserves no more LDS than the bandwidth can use.
static int ec_wide_cols(const mkde_ctx *c, float hf)
{
    const int wmax = mkde_dev::ec_wide_max();
    return hf > 0.0f ? std::min(wmax, mkde_dev::ec_box_cols(c->mg, hf)) : wmax;
}

// EC weights for bandwidth array h (original order) or the fixed value hf > 0,
// written to ew_out (original order).
static int run_ec(mkde_ctx *c, const float *h, float hf, float *ew_out, unsigned long long *n_eval,
                  bool fold_count = false)
{
    CHK(ensure_dist(c));
    CHK(c->list.ensure(c->n));
    CHK(c->icount.ensure(3));
    const int64_t nw = mkde_dev::slot_words();
    CHK(c->ucount.ensure(nw));
    HIPCHK(hipMemsetAsync(c->icount.p, 0, 3 * sizeof(int), c->stream));
    HIPCHK(hipMemsetAsync(c->ucount.p, 0, nw * sizeof(unsigned long long), c->stream));
    HIPCHK(mkde_dev::ec_active(c->stream, c->n, c->mg, c->x.p, c->y.p, c->dist.p, h, hf, nullptr, ew_out, c->list.p,
                               c->icount.p, std::min(c->tune.ec_wide_min, mkde_dev::ec_wide_max()),
                               ec_wide_cols(c, hf)));
    int cnt3[3] = {0, 0, 0};   // narrow points, wide points, widest box
    CHK(copy_d2h(c, cnt3, c->icount.p, sizeof cnt3));
    MaskRuns R{c->run_ptr.p, c->run_se.p};
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(mkde_dev::ec(c->stream, c->n, cnt3[0], cnt3[1], cnt3[2], c->mg, R, c->colx.p, c->rowy.p, c->x.p, c->y.p,
                        h, hf, c->list.p, ew_out, c->ucount.p, c->tune.ec_chunk, c->mg_cx32));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    // the pair count: folded, and read back unless the caller reads it later
    // (fold_count: from the last word of ucount, after its own wait)
    if (n_eval || fold_count) HIPCHK(mkde_dev::fold_slots(c->stream, c->ucount.p));
    if (n_eval) CHK(copy_d2h(c, n_eval, c->ucount.p + nw - 1, sizeof *n_eval));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_ec_ms, c->ev0, c->ev1));
    return MKDE_OK;
}

extern "C" int mkde_ec_weights(mkde_ctx *c, float *ew, int64_t *n_eval)
{
    CHK(need_points(c));
    CHK(need_mask(c));
    CHK(need_h(c));
    unsigned long long ne = 0;
    CHK(run_ec(c, c->h.p, c->h_fixed, c->ew.p, &ne));
    c->pts_dirty = true;
    if (n_eval) *n_eval = (int64_t)ne;
    if (ew) CHK(copy_d2h(c, ew, c->ew.p, c->n * sizeof(float)));
    return MKDE_OK;
}

extern "C" int mkde_set_ec_weights(mkde_ctx *c, const float *ew)
{
    CHK(need_points(c));
    if (ew) CHK(copy_h2d(c, c->ew.p, ew, c->n * sizeof(float)));
    else HIPCHK(mkde_dev::fill_f32(c->stream, c->n, 1.0f, c->ew.p));
    c->pts_dirty = true;
    return MKDE_OK;
}

extern "C" int mkde_get_ec_weights(mkde_ctx *c, float *ew)
{
    CHK(need_points(c));
    if (!ew) return fail(MKDE_E_ARG, "ew is NULL");
    return copy_d2h(c, ew, c->ew.p, c->n * sizeof(float));
}

// -------------------------------------------------------- range search --

extern "C" int mkde_range_candidates(mkde_ctx *c, int64_t *offsets, int32_t *idx, int64_t capacity, int64_t *total)
{
    CHK(need_points(c));
    CHK(need_h(c));
    CHK(prep(c));
    const int64_t n = c->n;
    CHK(host_perm(c));
    DevBuf<int64_t> counts;
    CHK(counts.ensure(n));
    PointIndex I = index_of(c);
    HIPCHK(mkde_dev::range_count(c->stream, I, counts.p));
    std::vector<int64_t> hc(n);
    CHK(copy_d2h(c, hc.data(), counts.p, n * sizeof(int64_t)));
    std::vector<int64_t> soff(n + 1, 0);
    for (int64_t i = 0; i < n; ++i) soff[i + 1] = soff[i] + hc[i];
    const int64_t tot = soff[n];
    if (total) *total = tot;
    if (offsets) {
        // offsets in the caller's point order
        std::vector<int64_t> cnt_orig(n);
        for (int64_t i = 0; i < n; ++i) cnt_orig[c->hperm[i]] = hc[i];
        offsets[0] = 0;
        for (int64_t i = 0; i < n; ++i) offsets[i + 1] = offsets[i] + cnt_orig[i];
    }
    if (!idx) return MKDE_OK;
    if (!offsets) return fail(MKDE_E_ARG, "offsets required with idx");
    if (capacity < tot) return fail(MKDE_E_RANGE, "capacity %lld < %lld candidates", (long long)capacity, (long long)tot);
    DevBuf<int64_t> doff;
    DevBuf<int32_t> didx;
    CHK(doff.ensure(n + 1));
    CHK(didx.ensure(tot));
    CHK(copy_h2d(c, doff.p, soff.data(), (n + 1) * sizeof(int64_t)));
    HIPCHK(mkde_dev::range_fill(c->stream, I, doff.p, didx.p));
    std::vector<int32_t> hs(tot);
    CHK(copy_d2h(c, hs.data(), didx.p, tot * sizeof(int32_t)));
    for (int64_t i = 0; i < n; ++i) {
        const int o = c->hperm[i];
        int32_t *dst = idx + offsets[o];
        for (int64_t k = soff[i]; k < soff[i + 1]; ++k) *dst++ = c->hperm[hs[k]];
        std::sort(idx + offsets[o], idx + offsets[o + 1]);
    }
    return MKDE_OK;
}

// ------------------------------------------------- densities at points --

// Work list of the point kernels: plan on the device, read the three counts,
// then items, heaviest-first order and partial slots (m values per query).
struct PointWork {
    const int *order = nullptr;
    const int4 *items = nullptr;
    int n_items = 0, n_split = 0;
};

static int plan_points(mkde_ctx *c, const PointIndex &I, int kind, float h_fixed, const double *thr_k, int K, int m,
                       PointWork &P)
{
    const int64_t nw = (c->n + 63) / 64;
    if (nw == 0) return MKDE_OK;
    CHK(c->pw_cost.ensure(nw));
    CHK(c->pw_parts.ensure(nw));
    CHK(c->pw_item0.ensure(nw));
    CHK(c->pw_slot0.ensure(nw));
    CHK(c->pw_sidx.ensure(nw));
    CHK(c->pw_sidx2.ensure(nw));
    CHK(c->pw_chunks.ensure((nw + 1023) / 1024));
    CHK(c->pw_counts.ensure(3));
    CHK(c->pw_hist.ensure(32));
    CHK(c->pw_total.ensure(1));
    DensityWork W{c->pw_cost.p,  c->pw_parts.p,  c->pw_item0.p, c->pw_slot0.p, c->pw_sidx.p,
                  c->pw_sidx2.p, c->pw_chunks.p, c->pw_total.p, c->pw_counts.p, c->pw_hist.p};
    HIPCHK(mkde_dev::point_plan(c->stream, I, kind, h_fixed, thr_k, K, W, c->tune));
    unsigned cnt[3];
    CHK(copy_d2h(c, cnt, c->pw_counts.p, sizeof cnt));
    P.n_items = (int)cnt[0];
    P.n_split = (int)cnt[2];
    CHK(c->pw_items.ensure(P.n_items));
    CHK(c->pw_order.ensure(P.n_items));
    CHK(c->pw_splits.ensure(P.n_split));
    CHK(c->pw_partials.ensure((size_t)cnt[1] * 64 * m));
    HIPCHK(mkde_dev::density_items(c->stream, (int)nw, W, P.n_items, c->pw_items.p, c->pw_splits.p, c->pw_order.p, nullptr));
    P.order = c->pw_order.p;
    P.items = c->pw_items.p;
    return MKDE_OK;
}

static int density_points_sorted(mkde_ctx *c, int loo, float *den_sorted)
{
    CHK(prep(c));
    PointIndex I = index_of(c);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    PointWork P;
    CHK(plan_points(c, I, 0, c->h_fixed, nullptr, 0, 1, P));
    HIPCHK(mkde_dev::point_density(c->stream, I, loo, den_sorted, c->h_fixed, P.order, P.items, c->pw_cost.p,
                                   c->pw_partials.p, P.n_items));
    HIPCHK(mkde_dev::point_merge(c->stream, c->n, 1, 1, P.n_split, c->pw_splits.p, c->pw_partials.p, den_sorted, nullptr));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return MKDE_OK;
}

extern "C" int mkde_density_at_points(mkde_ctx *c, int loo, float *den)
{
    CHK(need_points(c));
    CHK(need_h(c));
    CHK(c->den2.ensure(c->n));
    CHK(density_points_sorted(c, loo ? 1 : 0, c->den.p));
    HIPCHK(mkde_dev::unsort_f32(c->stream, c->n, c->perm.p, c->den.p, c->den2.p));
    if (den) CHK(copy_d2h(c, den, c->den2.p, c->n * sizeof(float)));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_points_ms, c->ev0, c->ev1));
    return MKDE_OK;
}

// Allocation-free variant of mkde_density_at_points: the same kernels, with
// the work list, the FP64 partial sums of split waves and the bin-order
// densities carved from the device's static scratch array (MKDE_SCRATCH_BYTES)
// instead of context buffers, and the result written to the caller's device
// buffer.  Calls no hipMalloc / hipFree.  Sizes beyond the scratch return
// MKDE_E_RANGE before any kernel runs.
namespace {
struct Carve {
    unsigned char *base;
    size_t cap, used = 0;
    bool ok = true;
    Carve(unsigned char *b, size_t c) : base(b), cap(c) {}
    template <class T>
    T *take(size_t count)
    {
        const size_t off = (used + 255) & ~(size_t)255;   // 256-byte aligned pieces
        const size_t bytes = (count ? count : 1) * sizeof(T);
        if (!ok || off > cap || bytes > cap - off) {
            ok = false;
            used = off + bytes;
            return nullptr;
        }
        used = off + bytes;
        return (T *)(base + off);
    }
};
}  // namespace

extern "C" int mkde_density_at_points_noalloc(mkde_ctx *c, int loo, float *den_dev)
{
    CHK(need_points(c));
    CHK(need_h(c));
    if (!den_dev) return fail(MKDE_E_ARG, "den_dev is NULL");
    if (!c->scratch) return fail(MKDE_E_STATE, "the scratch array of device %d is not resolved", c->device);
    const int64_t n = c->n, nw = (n + 63) / 64;
    Carve cv(c->scratch, MKDE_SCRATCH_BYTES);
    float *den_s = cv.take<float>((size_t)n);
    unsigned *cost = cv.take<unsigned>(nw), *parts = cv.take<unsigned>(nw), *item0 = cv.take<unsigned>(nw);
    unsigned *slot0 = cv.take<unsigned>(nw), *sidx = cv.take<unsigned>(nw), *sidx2 = cv.take<unsigned>(nw);
    unsigned *chunks = cv.take<unsigned>((nw + 1023) / 1024), *counts = cv.take<unsigned>(3);
    unsigned *hist = cv.take<unsigned>(32);
    unsigned long long *total = cv.take<unsigned long long>(1);
    if (!cv.ok)
        return fail(MKDE_E_RANGE, "%lld points need %zu bytes of scratch, more than the %u available", (long long)n,
                    cv.used, MKDE_SCRATCH_BYTES);
    CHK(prep(c));
    PointIndex I = index_of(c);
    DensityWork W{cost, parts, item0, slot0, sidx, sidx2, chunks, total, counts, hist};
    HIPCHK(mkde_dev::point_plan(c->stream, I, 0, c->h_fixed, nullptr, 0, W, c->tune));
    unsigned cnt[3];   // items, split slots, split waves
    CHK(copy_d2h(c, cnt, counts, sizeof cnt));
    int4 *items = cv.take<int4>(cnt[0]);
    int *order = cv.take<int>(cnt[0]);
    int4 *splits = cv.take<int4>(cnt[2]);
    double *partials = cv.take<double>((size_t)cnt[1] * 64);
    if (!cv.ok)
        return fail(MKDE_E_RANGE,
                    "the work list of %lld points (%u items, %u split parts) needs %zu bytes of scratch, more than "
                    "the %u available",
                    (long long)n, cnt[0], cnt[1], cv.used, MKDE_SCRATCH_BYTES);
    HIPCHK(mkde_dev::density_items(c->stream, (int)nw, W, (int)cnt[0], items, splits, order, nullptr));
    HIPCHK(mkde_dev::point_density(c->stream, I, loo ? 1 : 0, den_s, c->h_fixed, order, items, cost, partials,
                                   (int)cnt[0]));
    HIPCHK(mkde_dev::point_merge(c->stream, n, 1, 1, (int)cnt[2], splits, partials, den_s, nullptr));
    HIPCHK(mkde_dev::unsort_f32(c->stream, n, c->perm.p, den_s, den_dev));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MKDE_OK;
}

// FP64 sum of log(max(den, 1e-14f)) (clamp) or log(den) over K rows of n.
static int sum_logs(mkde_ctx *c, const float *den, int K, int clamp, double *out_dev)
{
    const int nb = (int)std::min<int64_t>(1024, (c->n + 255) / 256);
    CHK(c->partial.ensure((size_t)nb * K));
    HIPCHK(mkde_dev::sumlog(c->stream, c->n, K, den, clamp, c->partial.p, nb, out_dev));
    return MKDE_OK;
}

// --------------------------------------------------------------- LOO ----

// Restores the context's bandwidth state when a likelihood scan returns, on
// success and on every error path: the fixed bandwidth and its flag always,
// and the bandwidth and EC-weight arrays from hsave / ewsave when the scan
// overwrites them (the adaptive scan).
struct BandwidthGuard {
    mkde_ctx *c;
    float hf;
    bool hset;
    bool arrays = false;
    explicit BandwidthGuard(mkde_ctx *ctx) : c(ctx), hf(ctx->h_fixed), hset(ctx->h_set) {}
    ~BandwidthGuard()
    {
        if (arrays) {
            const size_t bytes = (size_t)c->n * sizeof(float);
            (void)hipMemcpyAsync(c->h.p, c->hsave.p, bytes, hipMemcpyDeviceToDevice, c->stream);
            (void)hipMemcpyAsync(c->ew.p, c->ewsave.p, bytes, hipMemcpyDeviceToDevice, c->stream);
        }
        (void)hipStreamSynchronize(c->stream);
        c->h_fixed = hf;
        c->h_set = hset;
        c->pts_dirty = true;
    }
};

// ----------------------------------------------- the search as a graph --
// A bandwidth search of K candidates runs as one HIP graph: for every
// candidate the EC pass (active list, narrow and wide kernels, all sized on
// the device) and its coefficient slot, then per 16 candidates the work list,
// the batched leave-one-out pass, the merge of split waves and the FP64 sums
// of logs into s_logl -- no host wait between the first launch and the
// read-back of the K log-likelihoods.  Buffers are sized for the largest
// work list a pass can have (waves + split target items and split parts), so
// nothing depends on a value only the device knows.  The candidates'
// bandwidths and slot constants are read from device arrays written before
// each launch, so the graph captured once serves every later search on the
// same resident inputs with the same K (the key lists every pointer and size
// the graph holds).  The kernels and their order are those of the host-sized
// path, so the log-likelihoods are the same bit for bit.
static const int64_t MKDE_SEARCH_GRAPH_MAX_PARTIALS = (int64_t)1 << 30;   // bytes of split-part sums

static int search_enqueue(mkde_ctx *c, int K, int ec_on, const PointIndex &I, int64_t items_max, int64_t nw,
                          const std::vector<int> &wcols)
{
    const int64_t n = c->n;
    MaskRuns R{c->run_ptr.p, c->run_se.p};
    DensityWork W{c->pw_cost.p,  c->pw_parts.p,  c->pw_item0.p, c->pw_slot0.p, c->pw_sidx.p,
                  c->pw_sidx2.p, c->pw_chunks.p, c->pw_total.p, c->pw_counts.p, c->pw_hist.p};
    if (!ec_on) HIPCHK(mkde_dev::fill_f32(c->stream, n, 1.0f, c->ewk.p));
    const int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    for (int g0 = 0, gi = 0; g0 < K; g0 += MKDE_LM_K, ++gi) {
        const int kg = std::min(MKDE_LM_K, K - g0);
        for (int j = 0; j < kg; ++j) {
            const float *hd = c->s_h.p + g0 + j;
            if (ec_on) {
                const int cw_narrow = std::min(c->tune.ec_wide_min, mkde_dev::ec_wide_max());
                HIPCHK(mkde_dev::fill_f32(c->stream, 3, 0.0f, (float *)c->icount.p));   // zero counts (int bits)
                HIPCHK(mkde_dev::ec_active(c->stream, n, c->mg, c->x.p, c->y.p, c->dist.p, nullptr, 0.0f, hd,
                                           c->ewk.p, c->list.p, c->icount.p, cw_narrow, wcols[g0 + j]));
                HIPCHK(mkde_dev::ec_dev(c->stream, n, c->mg, R, c->colx.p, c->rowy.p, c->x.p, c->y.p, hd,
                                        c->icount.p, c->list.p, c->ewk.p, c->tune.ec_chunk, cw_narrow,
                                        wcols[g0 + j], c->mg_cx32));
            }
            HIPCHK(mkde_dev::coef_k_dev(c->stream, n, c->perm.p, c->w.p, c->ewk.p, hd,
                                        c->coefk.p + (MKDE_LM_K - kg) + j, MKDE_LM_K));
        }
        const double *a = c->s_a.p + (int64_t)gi * MKDE_LM_K, *t = c->s_t.p + (int64_t)gi * MKDE_LM_K;
        HIPCHK(mkde_dev::point_plan(c->stream, I, 1, 0.0f, t, MKDE_LM_K, W, c->tune));
        HIPCHK(mkde_dev::density_items(c->stream, (int)nw, W, (int)items_max, c->pw_items.p, c->pw_splits.p,
                                       c->pw_order.p, c->pw_counts.p));
        HIPCHK(mkde_dev::loo_multi(c->stream, I, c->coefk.p, a, t, kg, c->den1k.p, c->pw_order.p, c->pw_items.p,
                                   c->pw_cost.p, c->pw_partials.p, (int)items_max, c->pw_counts.p));
        HIPCHK(mkde_dev::point_merge(c->stream, n, MKDE_LM_K, kg, (int)nw, c->pw_splits.p, c->pw_partials.p,
                                     c->den1k.p, c->pw_counts.p + 2));
        HIPCHK(mkde_dev::sumlog(c->stream, n, kg, c->den1k.p, 1, c->partial.p, nb, c->s_logl.p + g0));
    }
    return MKDE_OK;
}

// Runs the search as a graph when its buffers fit; *done = false leaves the
// search to the host-sized path.
static int loo_loglik_graph(mkde_ctx *c, int K, const std::vector<int> &ord, const float *h_cand, int ec_on,
                            double *logL, bool *done)
{
    *done = false;
    const int64_t n = c->n, nw = (n + 63) / 64;
    int64_t items_max = 0, slots_max = 0;
    point_work_bounds(n, c->tune.loo_target, &items_max, &slots_max);
    if (slots_max * 64 * MKDE_LM_K * (int64_t)sizeof(double) > MKDE_SEARCH_GRAPH_MAX_PARTIALS) return MKDE_OK;
    const int G = (K + MKDE_LM_K - 1) / MKDE_LM_K;
    // every buffer the graph touches, at its final size, before any capture
    if (ec_on) {
        CHK(ensure_dist(c));
        CHK(c->list.ensure(n));
        CHK(c->icount.ensure(3));
    }
    CHK(c->ewk.ensure(n));
    CHK(c->coefk.ensure((size_t)MKDE_LM_K * n));
    CHK(c->den1k.ensure((size_t)MKDE_LM_K * n));
    CHK(c->s_h.ensure(K));
    CHK(c->s_a.ensure((size_t)G * MKDE_LM_K));
    CHK(c->s_t.ensure((size_t)G * MKDE_LM_K));
    CHK(c->s_logl.ensure(K));
    CHK(c->pw_cost.ensure(nw));
    CHK(c->pw_parts.ensure(nw));
    CHK(c->pw_item0.ensure(nw));
    CHK(c->pw_slot0.ensure(nw));
    CHK(c->pw_sidx.ensure(nw));
    CHK(c->pw_sidx2.ensure(nw));
    CHK(c->pw_chunks.ensure((nw + 1023) / 1024));
    CHK(c->pw_counts.ensure(3));
    CHK(c->pw_hist.ensure(32));
    CHK(c->pw_total.ensure(1));
    CHK(c->pw_items.ensure(items_max));
    CHK(c->pw_order.ensure(items_max));
    CHK(c->pw_splits.ensure(nw));
    CHK(c->pw_partials.ensure((size_t)slots_max * 64 * MKDE_LM_K));
    const int nb = (int)std::min<int64_t>(1024, (n + 255) / 256);
    CHK(c->partial.ensure((size_t)nb * MKDE_LM_K));
    // the device inputs of this search: candidates ascending, slot constants
    std::vector<float> hs(K);
    std::vector<double> a((size_t)G * MKDE_LM_K, 0.0), t((size_t)G * MKDE_LM_K, -1.0);
    for (int k = 0; k < K; ++k) hs[k] = h_cand[ord[k]];
    // each candidate's widest wide EC box: the LDS of its wide launch (part
    // of the key, so new candidates that need other sizes recapture)
    std::vector<int> wcols(K);
    for (int k = 0; k < K; ++k) wcols[k] = ec_wide_cols(c, hs[k]);
    for (int g0 = 0, gi = 0; g0 < K; g0 += MKDE_LM_K, ++gi) {
        const int kg = std::min(MKDE_LM_K, K - g0);
        for (int j = 0; j < kg; ++j) {
            const double h2 = (double)(float)(hs[g0 + j] * hs[g0 + j]);
            a[(size_t)gi * MKDE_LM_K + MKDE_LM_K - kg + j] = -1.4426950408889634073599 / (2.0 * h2);
            t[(size_t)gi * MKDE_LM_K + MKDE_LM_K - kg + j] = 9.0 * h2;
        }
    }
    CHK(copy_h2d(c, c->s_h.p, hs.data(), K * sizeof(float)));
    CHK(copy_h2d(c, c->s_a.p, a.data(), a.size() * sizeof(double)));
    CHK(copy_h2d(c, c->s_t.p, t.data(), t.size() * sizeof(double)));
    PointIndex I = index_of(c);
    const uint64_t key_v[] = {(uint64_t)n, (uint64_t)K, (uint64_t)ec_on, (uint64_t)c->nbins,
                              (uint64_t)(uintptr_t)I.pts, (uint64_t)(uintptr_t)I.bin_start,
                              (uint64_t)(uintptr_t)I.bin_box, (uint64_t)(uintptr_t)I.bin_thr,
                              (uint64_t)(uintptr_t)I.reach_p, (uint64_t)(uintptr_t)c->x.p,
                              (uint64_t)(uintptr_t)c->y.p, (uint64_t)(uintptr_t)c->w.p,
                              (uint64_t)(uintptr_t)c->perm.p, (uint64_t)(uintptr_t)c->dist.p,
                              (uint64_t)(uintptr_t)c->mask.p, (uint64_t)(uintptr_t)c->run_ptr.p,
                              (uint64_t)(uintptr_t)c->run_se.p, (uint64_t)(uintptr_t)c->colx.p,
                              (uint64_t)(uintptr_t)c->rowy.p, (uint64_t)c->mg.ncols, (uint64_t)c->mg.nrows,
                              (uint64_t)__builtin_bit_cast(uint32_t, c->mg.xll),
                              (uint64_t)__builtin_bit_cast(uint32_t, c->mg.yll),
                              (uint64_t)__builtin_bit_cast(uint32_t, c->mg.cs), (uint64_t)c->mg_cx32,
                              (uint64_t)c->bg.nbx,
                              (uint64_t)c->bg.nby, __builtin_bit_cast(uint64_t, c->bg.x0),
                              __builtin_bit_cast(uint64_t, c->bg.y0), __builtin_bit_cast(uint64_t, c->bg.size),
                              (uint64_t)(uintptr_t)c->list.p, (uint64_t)(uintptr_t)c->icount.p,
                              (uint64_t)(uintptr_t)c->ewk.p, (uint64_t)(uintptr_t)c->coefk.p,
                              (uint64_t)(uintptr_t)c->den1k.p, (uint64_t)(uintptr_t)c->s_h.p,
                              (uint64_t)(uintptr_t)c->s_a.p, (uint64_t)(uintptr_t)c->s_t.p,
                              (uint64_t)(uintptr_t)c->s_logl.p, (uint64_t)(uintptr_t)c->pw_cost.p,
                              (uint64_t)(uintptr_t)c->pw_parts.p, (uint64_t)(uintptr_t)c->pw_item0.p,
                              (uint64_t)(uintptr_t)c->pw_slot0.p, (uint64_t)(uintptr_t)c->pw_sidx.p,
                              (uint64_t)(uintptr_t)c->pw_sidx2.p, (uint64_t)(uintptr_t)c->pw_chunks.p,
                              (uint64_t)(uintptr_t)c->pw_counts.p, (uint64_t)(uintptr_t)c->pw_hist.p,
                              (uint64_t)(uintptr_t)c->pw_total.p, (uint64_t)(uintptr_t)c->pw_items.p,
                              (uint64_t)(uintptr_t)c->pw_order.p, (uint64_t)(uintptr_t)c->pw_splits.p,
                              (uint64_t)(uintptr_t)c->pw_partials.p, (uint64_t)(uintptr_t)c->partial.p,
                              (uint64_t)c->tune.ec_chunk, (uint64_t)c->tune.ec_wide_min, c->tune.loo_target};
    std::vector<uint64_t> key(std::begin(key_v), std::end(key_v));
    if (ec_on) key.insert(key.end(), wcols.begin(), wcols.end());
    if (!c->s_exec || c->s_key != key) {
        if (c->s_exec) {
            (void)hipGraphExecDestroy(c->s_exec);
            c->s_exec = nullptr;
        }
        c->s_key.clear();
        hipGraph_t graph = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int st = search_enqueue(c, K, ec_on, I, items_max, nw, wcols);
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (st != MKDE_OK) {
            if (graph) (void)hipGraphDestroy(graph);
            return st;
        }
        if (e != hipSuccess) return fail(MKDE_E_HIP, "capturing the search graph failed: %s", hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&c->s_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) {
            c->s_exec = nullptr;
            return fail(MKDE_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(ei));
        }
        c->s_key = key;
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(hipGraphLaunch(c->s_exec, c->stream));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    std::vector<double> res(K);
    CHK(copy_d2h(c, res.data(), c->s_logl.p, K * sizeof(double)));
    HIPCHK(hipEventElapsedTime(&c->last_points_ms, c->ev0, c->ev1));
    for (int k = 0; k < K; ++k) logL[ord[k]] = res[k];
    *done = true;
    return MKDE_OK;
}

extern "C" int mkde_loo_loglik(mkde_ctx *c, int K, const float *h_cand, int ec_on, double *logL)
{
    CHK(need_points(c));
    if (K <= 0 || K > 1024 || !h_cand || !logL) return fail(MKDE_E_ARG, "mkde_loo_loglik: bad arguments");
    if (ec_on) CHK(need_mask(c));
    for (int k = 0; k < K; ++k)
        if (!(h_cand[k] > 0.0f) || !std::isfinite(h_cand[k])) return fail(MKDE_E_ARG, "candidate %d is not > 0", k);
    const int64_t n = c->n;
    // candidates in ascending order of h2f
    std::vector<int> ord(K);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return h_cand[a] < h_cand[b]; });
    // positions for the index: prep with the largest candidate so that the bin
    // thresholds cover every candidate
    float hmax = h_cand[ord[K - 1]];
    // the scan sets the fixed bandwidth to hmax for the index (the bandwidth
    // array itself is never written); the guard restores the caller's state
    BandwidthGuard guard(c);
    c->h_fixed = hmax;
    c->pts_dirty = true;
    CHK(prep(c));
    if (c->search_graph_on) {
        bool done = false;
        CHK(loo_loglik_graph(c, K, ord, h_cand, ec_on, logL, &done));
        if (done) return MKDE_OK;
    }
    PointIndex I = index_of(c);
    const int G = std::min(K, MKDE_LM_K);
    CHK(c->coefk.ensure((size_t)MKDE_LM_K * n));   // [n][MKDE_LM_K]: a point's candidate slots contiguous
    CHK(c->den1k.ensure((size_t)G * n));
    CHK(c->ewk.ensure(n));
    CHK(c->ak.ensure(MKDE_LM_K));
    CHK(c->tk.ensure(MKDE_LM_K));
    CHK(c->dsum.ensure(K));
    std::vector<double> res(K);
    for (int g0 = 0; g0 < K; g0 += MKDE_LM_K) {
        const int kg = std::min(MKDE_LM_K, K - g0);
        // the pass's kg candidates (ascending) in slots MKDE_LM_K - kg ..
        // MKDE_LM_K - 1; the slots below hold a threshold no pair passes
        std::vector<double> a(MKDE_LM_K, 0.0), t(MKDE_LM_K, -1.0);
        for (int j = 0; j < kg; ++j) {
            const float hk = h_cand[ord[g0 + j]];
            const double h2 = (double)(float)(hk * hk);
            a[MKDE_LM_K - kg + j] = -1.4426950408889634073599 / (2.0 * h2);
            t[MKDE_LM_K - kg + j] = 9.0 * h2;
            if (ec_on) CHK(run_ec(c, nullptr, hk, c->ewk.p, nullptr));
            else HIPCHK(mkde_dev::fill_f32(c->stream, n, 1.0f, c->ewk.p));
            // candidate j of this pass in coefficient slot MKDE_LM_K - kg + j
            HIPCHK(mkde_dev::coef_k(c->stream, n, c->perm.p, c->w.p, c->ewk.p, h2, c->coefk.p + (MKDE_LM_K - kg) + j,
                                    MKDE_LM_K));
        }
        CHK(copy_h2d(c, c->ak.p, a.data(), MKDE_LM_K * sizeof(double)));
        CHK(copy_h2d(c, c->tk.p, t.data(), MKDE_LM_K * sizeof(double)));
        HIPCHK(hipEventRecord(c->ev0, c->stream));
        PointWork P;
        // the work estimate culls with the largest threshold (the top slot)
        CHK(plan_points(c, I, 1, 0.0f, c->tk.p, MKDE_LM_K, MKDE_LM_K, P));
        HIPCHK(mkde_dev::loo_multi(c->stream, I, c->coefk.p, c->ak.p, c->tk.p, kg, c->den1k.p, P.order, P.items,
                                   c->pw_cost.p, c->pw_partials.p, P.n_items, nullptr));
        HIPCHK(mkde_dev::point_merge(c->stream, n, MKDE_LM_K, kg, P.n_split, c->pw_splits.p, c->pw_partials.p,
                                     c->den1k.p, nullptr));
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        CHK(sum_logs(c, c->den1k.p, kg, 1, c->dsum.p));
        std::vector<double> part(kg);
        CHK(copy_d2h(c, part.data(), c->dsum.p, kg * sizeof(double)));
        HIPCHK(hipEventElapsedTime(&c->last_points_ms, c->ev0, c->ev1));
        for (int j = 0; j < kg; ++j) res[ord[g0 + j]] = part[j];
    }
    for (int k = 0; k < K; ++k) logL[k] = res[k];
    return MKDE_OK;
}

// Abramson pipeline on the context: EC(h0) if ec_on, pilot at h0, g, h_i.
// Leaves h = h_i and ew = EC(h0) (or 1).
static int abramson_core(mkde_ctx *c, float h0, float alpha, int ec_on)
{
    CHK(mkde_set_bandwidth_fixed(c, h0));
    if (ec_on) {
        CHK(run_ec(c, nullptr, h0, c->ew.p, nullptr));
    } else {
        HIPCHK(mkde_dev::fill_f32(c->stream, c->n, 1.0f, c->ew.p));
    }
    c->pts_dirty = true;
    CHK(density_points_sorted(c, 0, c->den.p));
    CHK(c->dsum.ensure(1));
    CHK(c->gval.ensure(1));
    CHK(sum_logs(c, c->den.p, 1, 0, c->dsum.p));
    HIPCHK(mkde_dev::gml(c->stream, c->dsum.p, c->n, c->gval.p));
    HIPCHK(mkde_dev::abramson(c->stream, c->n, c->perm.p, c->den.p, c->gval.p, h0, alpha, c->h.p));
    c->h_fixed = 0.0f;
    c->h_set = true;
    c->pts_dirty = true;
    return MKDE_OK;
}

extern "C" int mkde_abramson(mkde_ctx *c, float h0, float alpha, int ec_on, float *h_out, float *den0_out,
                             double *g_out)
{
    CHK(need_points(c));
    if (!(h0 > 0.0f) || !std::isfinite(h0) || !std::isfinite(alpha)) return fail(MKDE_E_ARG, "bad h0/alpha");
    if (ec_on) CHK(need_mask(c));
    CHK(abramson_core(c, h0, alpha, ec_on));
    if (h_out) CHK(copy_d2h(c, h_out, c->h.p, c->n * sizeof(float)));
    if (den0_out) {
        CHK(c->den2.ensure(c->n));
        HIPCHK(mkde_dev::unsort_f32(c->stream, c->n, c->perm.p, c->den.p, c->den2.p));
        CHK(copy_d2h(c, den0_out, c->den2.p, c->n * sizeof(float)));
    }
    if (g_out) CHK(copy_d2h(c, g_out, c->gval.p, sizeof(double)));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MKDE_OK;
}

extern "C" int mkde_loo_loglik_adaptive(mkde_ctx *c, int K, const float *h, const float *alpha, int ec_on,
                                        double *logL)
{
    CHK(need_points(c));
    if (K <= 0 || !h || !alpha || !logL) return fail(MKDE_E_ARG, "mkde_loo_loglik_adaptive: bad arguments");
    if (ec_on) CHK(need_mask(c));
    const int64_t n = c->n;
    CHK(c->hsave.ensure(n));
    CHK(c->ewsave.ensure(n));
    CHK(c->dsum.ensure(1));
    HIPCHK(hipMemcpyAsync(c->hsave.p, c->h.p, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->ewsave.p, c->ew.p, n * sizeof(float), hipMemcpyDeviceToDevice, c->stream));
    // each candidate overwrites h and ew; the guard copies them back on return
    BandwidthGuard guard(c);
    guard.arrays = true;
    for (int k = 0; k < K; ++k) {
        if (!(h[k] > 0.0f) || !std::isfinite(h[k]) || !std::isfinite(alpha[k])) {
            logL[k] = -INFINITY;
            continue;
        }
        CHK(abramson_core(c, h[k], alpha[k], ec_on));
        if (ec_on) CHK(run_ec(c, c->h.p, 0.0f, c->ew.p, nullptr));
        c->pts_dirty = true;
        CHK(density_points_sorted(c, 1, c->den.p));
        CHK(sum_logs(c, c->den.p, 1, 1, c->dsum.p));
        CHK(copy_d2h(c, &logL[k], c->dsum.p, sizeof(double)));
    }
    return MKDE_OK;
}

// ------------------------------------------------------------ optimisers --
// MLE_FixedBandWidth (tiled:3044-3093) with the float control flow of the
// reference; the two probes of an iteration are one batched LOO call and the
// comparison is on the FP64 log-likelihoods.
#pragma clang fp contract(off)
extern "C" int mkde_mle_fixed_bandwidth(mkde_ctx *c, float h, int ec_on, float *hopt, int *n_iter)
{
    CHK(need_points(c));
    if (!hopt || !(h > 0.0f)) return fail(MKDE_E_ARG, "mkde_mle_fixed_bandwidth: bad arguments");
    float hA = h / 20;
    float hD = 2 * h;
    float width = hD - hA;
    float epsilon = width / 100;
    float factor = 1 + sqrtf(5.0f);
    int iteration = 0;
    while (width > epsilon && iteration < 30) {
        float hB = hA + width / factor;
        float hC = hD - width / factor;
        float hs[2] = {hB, hC};
        double L[2];
        CHK(mkde_loo_loglik(c, 2, hs, ec_on, L));
        if (L[0] > L[1]) hD = hC;
        else hA = hB;
        width = hD - hA;
        iteration += 1;
    }
    *hopt = (hA + hD) / 2;
    if (n_iter) *n_iter = iteration;
    return MKDE_OK;
}

// hooke_jeeves + hj_likelihood (tiled:3677-3790), bookkeeping as in the
// reference (likelihoods stored as float, running maximum).
extern "C" int mkde_hooke_jeeves(mkde_ctx *c, float h0, float alpha0, float stepH, float stepA, int ec_on, float *out3,
                                 int *n_iter)
{
    CHK(need_points(c));
    if (!out3) return fail(MKDE_E_ARG, "out3 is NULL");
    float Ls[5] = {0, 0, 0, 0, 0};
    auto probe = [&](int last) -> int {
        float hh[5], aa[5];
        int which[5], m = 0;
        if (last == -1) { hh[m] = h0; aa[m] = alpha0; which[m++] = 0; }
        if (last != 2) { hh[m] = h0 - stepH; aa[m] = alpha0; which[m++] = 1; }
        if (last != 1) { hh[m] = h0 + stepH; aa[m] = alpha0; which[m++] = 2; }
        if (last != 4) { hh[m] = h0; aa[m] = alpha0 + stepA; which[m++] = 3; }
        if (last != 3) { hh[m] = h0; aa[m] = alpha0 - stepA; which[m++] = 4; }
        double L[5];
        int s = mkde_loo_loglik_adaptive(c, m, hh, aa, ec_on, L);
        if (s != MKDE_OK) return s;
        for (int i = 0; i < m; ++i) Ls[which[i]] = (float)L[i];
        return MKDE_OK;
    };
    CHK(probe(-1));
    float Lmax = Ls[0];
    float s = stepH / 10;
    float a = stepA / 10;
    int iteration = 0;
    while ((stepH > s || stepA > a) && (h0 - stepH > 0) && iteration <= 30) {
        int dmax = 0;
        for (int i = 0; i < 5; i++) {
            if (Ls[i] > Lmax) {
                Lmax = Ls[i];
                dmax = i;
            }
        }
        if (dmax == 0) {
            stepH = stepH / 2;
            stepA = stepA / 2;
        } else {
            if (dmax == 1) { h0 = h0 - stepH; Ls[2] = Ls[0]; Ls[0] = Ls[1]; }
            if (dmax == 2) { h0 = h0 + stepH; Ls[1] = Ls[0]; Ls[0] = Ls[2]; }
            if (dmax == 3) { alpha0 = alpha0 + stepA; Ls[3] = Ls[0]; Ls[0] = Ls[4]; }
            if (dmax == 4) { alpha0 = alpha0 - stepA; Ls[3] = Ls[0]; Ls[0] = Ls[4]; }
        }
        CHK(probe(dmax));
        iteration++;
    }
    out3[0] = h0;
    out3[1] = alpha0;
    out3[2] = Lmax;
    if (n_iter) *n_iter = iteration;
    return MKDE_OK;
}
#pragma clang fp contract(on)

// ------------------------------------------------------------ density ----

extern "C" int mkde_tile_upload(mkde_ctx *c, const mkde_grid *tile, const float *mask)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(check_grid(tile));
    CHK(set_device(c));
    const int64_t nc = tile->ncols * tile->nrows;
    // a failure below leaves the context without a tile
    c->tile_set = false;
    c->tile_is_ec_mask = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    CHK(c->tout.ensure(nc));
    if (mask) {
        CHK(c->tmask.ensure(nc));
        CHK(copy_h2d(c, c->tmask.p, mask, nc * sizeof(float)));
    }
    c->tg = to_desc(tile);
    c->tl = Tiling{};
    c->tile_has_mask = mask != nullptr;
    c->tile_set = true;
    return MKDE_OK;
}

extern "C" int mkde_tile_from_mask(mkde_ctx *c)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(need_mask(c));
    CHK(set_device(c));
    c->tile_set = false;
    c->tile_is_ec_mask = false;
    HIPCHK(hipStreamSynchronize(c->stream));
    CHK(c->tout.ensure(c->mg.ncols * c->mg.nrows));
    c->tg = c->mg;
    c->tl = Tiling{};
    c->tile_has_mask = true;
    c->tile_is_ec_mask = true;
    c->tile_set = true;
    return MKDE_OK;
}

extern "C" int mkde_raster_upload(mkde_ctx *c, int64_t ncols, int64_t nrows, const double *gt, float nodata,
                                  int64_t tile_cols, int64_t tile_rows, const float *mask)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    if (!gt) return fail(MKDE_E_ARG, "geotransform is NULL");
    if (tile_cols <= 0 || tile_rows <= 0) return fail(MKDE_E_ARG, "tile size must be positive");
    for (int k = 0; k < 6; ++k)
        if (!std::isfinite(gt[k])) return fail(MKDE_E_ARG, "geotransform[%d] is not finite", k);
    // tiles never exceed the raster; blocks per tile and in total must fit an int
    if (tile_cols > ncols) tile_cols = ncols;
    if (tile_rows > nrows) tile_rows = nrows;
    mkde_grid rg{ncols, nrows, (float)gt[0], (float)(gt[3] + (double)nrows * gt[5]), (float)gt[1], nodata};
    CHK(check_grid(&rg));
    const int64_t bpx = (tile_cols + MKDE_DT_BW - 1) / MKDE_DT_BW, bpy = (tile_rows + 63) / 64;
    const int64_t ntx = (ncols + tile_cols - 1) / tile_cols, nty = (nrows + tile_rows - 1) / tile_rows;
    if (bpx * bpy * ntx * nty > INT32_MAX) return fail(MKDE_E_RANGE, "raster has more than 2^31-1 density blocks");
    CHK(mkde_tile_upload(c, &rg, mask));
    c->tl = Tiling{tile_cols, tile_rows, (int)ntx, (int)bpx, (int)(bpx * bpy), gt[0], gt[1], gt[3], gt[5]};
    return MKDE_OK;
}

extern "C" int mkde_density_tile_resident(mkde_ctx *c)
{
    CHK(need_points(c));
    CHK(need_h(c));
    if (!c->tile_set) return fail(MKDE_E_STATE, "no tile uploaded");
    CHK(prep(c));
    PointIndex I = index_of(c);
    const float *tm = c->tile_is_ec_mask ? c->mask.p : c->tile_has_mask ? c->tmask.p : nullptr;
    const int *order = nullptr;
    int n_items = 0, n_split = 0;
    if (c->dt_order_on) {
        const int64_t nblk = mkde_dev::raster_blocks(c->tg, c->tl);
        CHK(c->blk_cost.ensure(nblk));
        CHK(c->blk_parts.ensure(nblk));
        CHK(c->blk_item0.ensure(nblk));
        CHK(c->blk_slot0.ensure(nblk));
        CHK(c->blk_sidx.ensure(nblk));
        CHK(c->blk_sidx2.ensure(nblk));
        CHK(c->blk_chunks.ensure((nblk + 1023) / 1024));
        CHK(c->blk_counts.ensure(3));
        CHK(c->blk_hist.ensure(32));
        CHK(c->blk_total.ensure(1));
        DensityWork W{c->blk_cost.p, c->blk_parts.p,  c->blk_item0.p, c->blk_slot0.p, c->blk_sidx.p,
                      c->blk_sidx2.p, c->blk_chunks.p, c->blk_total.p, c->blk_counts.p, c->blk_hist.p};
        HIPCHK(mkde_dev::density_plan(c->stream, c->tg, tm, I, c->tl, W, c->tune));
        unsigned cnt[3];
        CHK(copy_d2h(c, cnt, c->blk_counts.p, sizeof cnt));
        n_items = (int)cnt[0];
        n_split = (int)cnt[2];
        CHK(c->dt_items.ensure(n_items));
        CHK(c->dt_order.ensure(n_items));
        CHK(c->dt_splits.ensure(n_split));
        CHK(c->dt_partials.ensure((size_t)cnt[1] * MKDE_DT_CELLS));
        HIPCHK(mkde_dev::density_items(c->stream, (int)nblk, W, n_items, c->dt_items.p, c->dt_splits.p,
                                       c->dt_order.p, nullptr));
        order = c->dt_order.p;
    }
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    HIPCHK(mkde_dev::density_tile(c->stream, c->tg, tm, c->tout.p, I, c->dt_stats_on ? c->dt_stats.p : nullptr,
                                  order, c->dt_items.p, c->blk_cost.p, c->dt_partials.p, n_items, c->tl));
    HIPCHK(mkde_dev::density_merge(c->stream, c->tg, c->tl, tm, n_split, c->dt_splits.p, c->dt_partials.p,
                                   c->tout.p));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    return MKDE_OK;
}

extern "C" int mkde_tile_download(mkde_ctx *c, float *out)
{
    if (!c || !out) return fail(MKDE_E_ARG, "null argument");
    if (!c->tile_set) return fail(MKDE_E_STATE, "no tile uploaded");
    CHK(set_device(c));
    CHK(copy_d2h(c, out, c->tout.p, (size_t)(c->tg.ncols * c->tg.nrows) * sizeof(float)));
    HIPCHK(hipEventSynchronize(c->ev1));
    HIPCHK(hipEventElapsedTime(&c->last_density_ms, c->ev0, c->ev1));
    return MKDE_OK;
}

extern "C" int mkde_density_tile(mkde_ctx *c, const mkde_grid *tile, const float *mask, float *out)
{
    CHK(mkde_tile_upload(c, tile, mask));
    CHK(mkde_density_tile_resident(c));
    if (out) return mkde_tile_download(c, out);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventElapsedTime(&c->last_density_ms, c->ev0, c->ev1));
    return MKDE_OK;
}

extern "C" int mkde_density_raster(mkde_ctx *c, int64_t ncols, int64_t nrows, const double *gt, float nodata,
                                   int64_t tile_cols, int64_t tile_rows, const float *mask, float *out)
{
    CHK(mkde_raster_upload(c, ncols, nrows, gt, nodata, tile_cols, tile_rows, mask));
    CHK(mkde_density_tile_resident(c));
    if (out) return mkde_tile_download(c, out);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventElapsedTime(&c->last_density_ms, c->ev0, c->ev1));
    return MKDE_OK;
}

extern "C" int mkde_count_density_pairs(mkde_ctx *c, const mkde_grid *tile, const float *mask, int64_t *n_pairs)
{
    CHK(need_points(c));
    CHK(need_h(c));
    CHK(check_grid(tile));
    if (!n_pairs) return fail(MKDE_E_ARG, "n_pairs is NULL");
    RunsHost R;
    try {
        if (mask) {
            build_runs(mask, tile->ncols, tile->nrows, tile->nodata, R);
        } else {
            R.ptr.resize(tile->nrows + 1);
            R.se.resize(tile->nrows);
            for (int64_t r = 0; r <= tile->nrows; ++r) R.ptr[r] = r;
            for (int64_t r = 0; r < tile->nrows; ++r) R.se[r] = int2{0, (int)(tile->ncols - 1)};
        }
    } catch (const std::bad_alloc &) {
        return fail(MKDE_E_NOMEM, "out of host memory");
    }
    DevBuf<int64_t> rp;
    DevBuf<int2> rs;
    CHK(rp.ensure(R.ptr.size()));
    CHK(rs.ensure(R.se.size()));
    CHK(copy_h2d(c, rp.p, R.ptr.data(), R.ptr.size() * sizeof(int64_t)));
    CHK(copy_h2d(c, rs.p, R.se.data(), R.se.size() * sizeof(int2)));
    CHK(prep(c));
    const int64_t nw = mkde_dev::slot_words();
    CHK(c->ucount.ensure(nw));
    HIPCHK(hipMemsetAsync(c->ucount.p, 0, nw * sizeof(unsigned long long), c->stream));
    MaskRuns MR{rp.p, rs.p};
    PointIndex I = index_of(c);
    GridDesc g = to_desc(tile);
    HIPCHK(mkde_dev::count_pairs(c->stream, g, MR, I, c->ucount.p));
    HIPCHK(mkde_dev::fold_slots(c->stream, c->ucount.p));
    unsigned long long t = 0;
    CHK(copy_d2h(c, &t, c->ucount.p + nw - 1, sizeof t));
    *n_pairs = (int64_t)t;
    return MKDE_OK;
}

extern "C" int mkde_device_memory(mkde_ctx *c, int64_t *free_bytes, int64_t *total_bytes)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(set_device(c));
    size_t fr = 0, to = 0;
    HIPCHK(hipMemGetInfo(&fr, &to));
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)to;
    return MKDE_OK;
}

extern "C" int mkde_density_stats(mkde_ctx *c, int enable, int64_t *counters)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    CHK(set_device(c));
    CHK(c->dt_stats.ensure(8));
    if (enable) {
        HIPCHK(hipMemsetAsync(c->dt_stats.p, 0, 8 * sizeof(unsigned long long), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        c->dt_stats_on = true;
    } else {
        c->dt_stats_on = false;
    }
    if (counters) CHK(copy_d2h(c, counters, c->dt_stats.p, 8 * sizeof(int64_t)));
    return MKDE_OK;
}

extern "C" int mkde_set_option(mkde_ctx *c, const char *name, int64_t value)
{
    if (!c || !name) return fail(MKDE_E_ARG, "null argument");
    if (!strcmp(name, "search_graph")) {
        c->search_graph_on = value != 0;
    } else if (!strcmp(name, "density_order")) {
        c->dt_order_on = value != 0;
    } else if (!strcmp(name, "points_per_bin")) {
        if (value < 1 || value > 4096) return fail(MKDE_E_RANGE, "points_per_bin %lld not in [1, 4096]", (long long)value);
        c->pts_per_bin = (int)value;   // takes effect at the next mkde_points_upload
    } else if (!strcmp(name, "ec_chunk")) {
        if (value != 512 && value != 1024 && value != 2048) return fail(MKDE_E_RANGE, "ec_chunk must be 512, 1024 or 2048");
        c->tune.ec_chunk = (int)value;
    } else if (!strcmp(name, "ec_wide_min")) {
        if (value < 1 || value > mkde_dev::ec_wide_max())
            return fail(MKDE_E_RANGE, "ec_wide_min %lld not in [1, %d]", (long long)value, mkde_dev::ec_wide_max());
        c->tune.ec_wide_min = (int)value;
    } else if (!strcmp(name, "loo_split_target") || !strcmp(name, "pts_split_target") ||
               !strcmp(name, "density_split_target")) {
        if (value < 64 || value > ((int64_t)1 << 24)) return fail(MKDE_E_RANGE, "%s %lld not in [64, 2^24]", name, (long long)value);
        (name[0] == 'l' ? c->tune.loo_target : name[0] == 'p' ? c->tune.pts_target : c->tune.dt_target) =
            (unsigned long long)value;
    } else {
        return fail(MKDE_E_ARG, "unknown option '%s'", name);
    }
    return MKDE_OK;
}

extern "C" int mkde_get_option(mkde_ctx *c, const char *name, int64_t *value)
{
    if (!c || !name || !value) return fail(MKDE_E_ARG, "null argument");
    if (!strcmp(name, "search_graph")) *value = c->search_graph_on ? 1 : 0;
    else if (!strcmp(name, "density_order")) *value = c->dt_order_on ? 1 : 0;
    else if (!strcmp(name, "points_per_bin")) *value = c->pts_per_bin;
    else if (!strcmp(name, "ec_chunk")) *value = c->tune.ec_chunk;
    else if (!strcmp(name, "ec_wide_min")) *value = c->tune.ec_wide_min;
    else if (!strcmp(name, "loo_split_target")) *value = (int64_t)c->tune.loo_target;
    else if (!strcmp(name, "pts_split_target")) *value = (int64_t)c->tune.pts_target;
    else if (!strcmp(name, "density_split_target")) *value = (int64_t)c->tune.dt_target;
    else return fail(MKDE_E_ARG, "unknown option '%s'", name);
    return MKDE_OK;
}

// ------------------------------------------------------------ autotuner --
// Coordinate descent over the launch-shape options of a context, timing a
// resident workload on the device it runs on: each option in turn takes each
// of its candidate values with the others at their current choice; a value
// replaces the current one only when its best-of-reps time is more than 1%
// faster.  The context keeps the chosen values; the choices are recorded per
// job class (log2 buckets of the point count and raster cells, bandwidth
// option, edge correction, optimiser) for mkde_autotune_lookup, and a JSON
// report lists every timing.  Block width is not tuned (a build-time
// constant, MKDE_DT_BW).
namespace {
struct TuneChoice {
    Tuning t;
    int ppb;
};
std::mutex g_tune_mu;
std::map<std::vector<int64_t>, TuneChoice> g_tune_table;

std::vector<int64_t> tune_class(int64_t n, int64_t cells, int hoption, int ec, int optimise)
{
    auto lg = [](int64_t v) {
        int b = 0;
        while (v > 1) {
            v >>= 1;
            ++b;
        }
        return (int64_t)b;
    };
    return {lg(n), lg(cells), hoption, ec ? 1 : 0, optimise ? 1 : 0};
}

struct TuneOpt {
    const char *name;
    std::vector<int64_t> values;
};

int autotune_run(mkde_ctx *c, const std::vector<TuneOpt> &opts, int reps, const std::function<int()> &work,
                 std::string &rep, double *best_out)
{
    auto timed = [&](double *ms) -> int {
        double best = 1e300;
        for (int r = 0; r < reps; ++r) {
            HIPCHK(hipStreamSynchronize(c->stream));
            const auto t0 = std::chrono::steady_clock::now();
            CHK(work());
            HIPCHK(hipStreamSynchronize(c->stream));
            best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        *ms = best;
        return MKDE_OK;
    };
    CHK(work());   // warm-up (code objects, buffers at size)
    double cur = 0.0;
    CHK(timed(&cur));
    char buf[256];
    snprintf(buf, sizeof buf, "\"baseline_ms\": %.4f, \"trials\": [", cur);
    rep += buf;
    bool first = true;
    // an option takes a value; points_per_bin rebuilds the point index
    auto apply = [&](const char *name, int64_t v) -> int {
        CHK(mkde_set_option(c, name, v));
        if (!strcmp(name, "points_per_bin")) CHK(mkde_points_upload(c, c->n, c->hx.data(), c->hy.data(), c->hw.data()));
        return MKDE_OK;
    };
    for (const TuneOpt &o : opts) {
        int64_t keep = 0;
        CHK(mkde_get_option(c, o.name, &keep));
        int64_t best_v = keep;
        double best_ms = cur;
        for (int64_t v : o.values) {
            if (v == keep) continue;
            double ms = 0.0;
            int st = apply(o.name, v);
            if (st == MKDE_OK) st = timed(&ms);
            if (st != MKDE_OK) {
                // the option keeps its value from before the failed trial
                const std::string msg = g_err;
                (void)apply(o.name, keep);
                return fail(st, "%s", msg.c_str());
            }
            snprintf(buf, sizeof buf, "%s{\"option\": \"%s\", \"value\": %lld, \"ms\": %.4f}", first ? "" : ", ", o.name,
                     (long long)v, ms);
            rep += buf;
            first = false;
            if (ms < best_ms * 0.99) {
                best_ms = ms;
                best_v = v;
            }
        }
        CHK(apply(o.name, best_v));
        cur = best_ms;
    }
    rep += "], \"chosen\": {";
    first = true;
    for (const TuneOpt &o : opts) {
        int64_t v = 0;
        CHK(mkde_get_option(c, o.name, &v));
        snprintf(buf, sizeof buf, "%s\"%s\": %lld", first ? "" : ", ", o.name, (long long)v);
        rep += buf;
        first = false;
    }
    snprintf(buf, sizeof buf, "}, \"final_ms\": %.4f", cur);
    rep += buf;
    *best_out = cur;
    return MKDE_OK;
}

int tune_report(const std::string &rep, char *report, int64_t cap)
{
    if (!report || cap <= 0) return MKDE_OK;
    if ((int64_t)rep.size() + 1 > cap) return fail(MKDE_E_RANGE, "report needs %zu bytes", rep.size() + 1);
    memcpy(report, rep.c_str(), rep.size() + 1);
    return MKDE_OK;
}
}  // namespace

extern "C" int mkde_autotune_job(mkde_ctx *c, const mkde_job_opts *o, int reps, char *report, int64_t report_cap)
{
    CHK(need_points(c));
    if (!o || reps < 1 || reps > 100) return fail(MKDE_E_ARG, "mkde_autotune_job: bad arguments");
    if (!c->tile_set) return fail(MKDE_E_STATE, "no density tile uploaded");
    const bool ec = o->edge_correction != 0;
    const bool loo = o->hoption == 1 || (o->hoption == 2 && o->optimise);
    std::vector<TuneOpt> opts;
    opts.push_back({"density_split_target", {2048, 4096, 8192, 16384}});
    if (o->hoption == 2) opts.push_back({"pts_split_target", {4096, 16384, 65536}});
    if (loo) opts.push_back({"loo_split_target", {8192, 32768, 131072}});
    if (ec) {
        opts.push_back({"ec_chunk", {512, 1024, 2048}});
        opts.push_back({"ec_wide_min", {512, 1024, 2048, 4096}});
    }
    if (o->hoption != -1) opts.push_back({"points_per_bin", {32, 64, 128}});
    std::string rep = "{\"workload\": \"job\", ";
    double ms = 0.0;
    CHK(autotune_run(c, opts, reps, [&]() { return mkde_run_job(c, o, nullptr); }, rep, &ms));
    rep += "}";
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_table[tune_class(c->n, c->tg.ncols * c->tg.nrows, o->hoption, ec, o->optimise)] =
            TuneChoice{c->tune, c->pts_per_bin};
    }
    return tune_report(rep, report, report_cap);
}

extern "C" int mkde_autotune_search(mkde_ctx *c, int K, const float *h_cand, int ec_on, int reps, char *report,
                                    int64_t report_cap)
{
    CHK(need_points(c));
    if (K <= 0 || K > 1024 || !h_cand || reps < 1 || reps > 100) return fail(MKDE_E_ARG, "mkde_autotune_search: bad arguments");
    std::vector<double> L(K);
    std::vector<TuneOpt> opts;
    opts.push_back({"loo_split_target", {8192, 32768, 131072}});
    if (ec_on) {
        opts.push_back({"ec_chunk", {512, 1024, 2048}});
        opts.push_back({"ec_wide_min", {512, 1024, 2048, 4096}});
    }
    opts.push_back({"points_per_bin", {32, 64, 128}});
    std::string rep = "{\"workload\": \"search\", ";
    double ms = 0.0;
    CHK(autotune_run(c, opts, reps, [&]() { return mkde_loo_loglik(c, K, h_cand, ec_on, L.data()); }, rep, &ms));
    rep += "}";
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_table[tune_class(c->n, c->mask_set ? c->mg.ncols * c->mg.nrows : 0, MKDE_H_SEARCH, ec_on, 0)] =
            TuneChoice{c->tune, c->pts_per_bin};
    }
    return tune_report(rep, report, report_cap);
}

extern "C" int mkde_autotune_lookup(mkde_ctx *c, int64_t n, int64_t cells, int hoption, int ec, int optimise)
{
    if (!c) return fail(MKDE_E_ARG, "null context");
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_table.find(tune_class(n, cells, hoption, ec, optimise));
    if (it == g_tune_table.end()) return fail(MKDE_E_STATE, "no tuned settings for this job class");
    c->tune = it->second.t;
    c->pts_per_bin = it->second.ppb;
    return MKDE_OK;
}

extern "C" int mkde_last_kernel_ms(mkde_ctx *c, const char *which, double *ms)
{
    if (!c || !which || !ms) return fail(MKDE_E_ARG, "null argument");
    if (!strcmp(which, "density")) *ms = c->last_density_ms;
    else if (!strcmp(which, "ec")) *ms = c->last_ec_ms;
    else if (!strcmp(which, "points")) *ms = c->last_points_ms;
    else return fail(MKDE_E_ARG, "unknown kernel '%s'", which);
    return MKDE_OK;
}

// ------------------------------------------------------------------ jobs --

namespace {
// Phase times of a job: events are recorded around each phase and read once
// at the end of the job, so timing adds no host synchronisation.  The events
// come from the context's pool (created on first use, kept until the context
// is destroyed), so a job creates none and an early error return leaks none.
struct PhaseTimes {
    struct Ev {
        hipEvent_t a, b;
        double *acc;
    };
    mkde_ctx *c;
    std::vector<Ev> ev;
    explicit PhaseTimes(mkde_ctx *ctx) : c(ctx) { c->ev_next = 0; }
    ~PhaseTimes() { c->ev_next = 0; }
    hipEvent_t get()
    {
        if (c->ev_next < c->ev_pool.size()) return c->ev_pool[c->ev_next++];
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->ev_pool.push_back(e);
        c->ev_next = c->ev_pool.size();
        return e;
    }
    void finish()
    {
        for (const Ev &e : ev) {
            float ms = 0.0f;
            if (e.a && e.b && hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess && e.acc) *e.acc += ms;
        }
    }
};
struct PhaseTimer {
    PhaseTimes &T;
    hipStream_t s;
    PhaseTimes::Ev e{nullptr, nullptr, nullptr};
    PhaseTimer(PhaseTimes &times, hipStream_t st, double *accum) : T(times), s(st)
    {
        e.acc = accum;
        e.a = T.get();
        e.b = T.get();
        if (e.a) (void)hipEventRecord(e.a, s);
    }
    ~PhaseTimer()
    {
        if (e.b) (void)hipEventRecord(e.b, s);
        T.ev.push_back(e);
    }
};
}  // namespace

extern "C" int mkde_run_job(mkde_ctx *c, const mkde_job_opts *o, mkde_job_stats *st)
{
    CHK(need_points(c));
    if (!o) return fail(MKDE_E_ARG, "options are NULL");
    if (o->hoption < -1 || o->hoption > 2) return fail(MKDE_E_ARG, "hoption %d not in {-1,0,1,2}", o->hoption);
    if (!c->tile_set) return fail(MKDE_E_STATE, "no density tile uploaded");
    const int ec_on = o->edge_correction ? 1 : 0;
    if (ec_on) CHK(need_mask(c));
    if (o->hoption == -1) CHK(need_h(c));
    mkde_job_stats S;
    memset(&S, 0, sizeof S);
    PhaseTimes times(c);
    hipEvent_t t0 = times.get(), t1 = times.get();
    if (!t0 || !t1) return fail(MKDE_E_HIP, "hipEventCreate failed");
    HIPCHK(hipEventRecord(t0, c->stream));
    if (ec_on) {
        PhaseTimer pt(times, c->stream, &S.ms_boundary);
        if (o->recompute_boundary) {
            c->bflags_ok = false;
            c->nboundary = -1;
            c->bt_ok = false;
            c->dist_set = false;
        }
        CHK(ensure_dist(c));
        // the rule of thumb (sequential float sums on the host, in the
        // reference's order) runs while the device computes the distances
        S.h0 = rule_of_thumb_host(c->hx, c->hy, c->hw);
    } else {
        S.h0 = rule_of_thumb_host(c->hx, c->hy, c->hw);
    }
    if (o->hoption != -1) CHK(mkde_set_bandwidth_fixed(c, S.h0));
    // the final EC pass's pair count and the geometric mean are read after
    // the job's last wait (no extra host synchronisation)
    bool ec_counted = false;
    auto ec_pass = [&](void) -> int {
        PhaseTimer pt(times, c->stream, &S.ms_ec);
        int s = run_ec(c, c->h.p, c->h_fixed, c->ew.p, nullptr, true);
        ec_counted = s == MKDE_OK;
        c->pts_dirty = true;
        return s;
    };
    // EC weights at the job's final bandwidths.  The bandwidth searches
    // (hoption 1 and 2) compute their own EC per candidate and never read the
    // context weights, so those jobs run the final EC pass only.
    auto final_ec = [&](void) -> int {
        if (ec_on) return ec_pass();
        HIPCHK(mkde_dev::fill_f32(c->stream, c->n, 1.0f, c->ew.p));
        c->pts_dirty = true;
        return MKDE_OK;
    };
    if (o->hoption == 0 || o->hoption == -1) CHK(final_ec());
    if (o->hoption == 1) {
        {
            PhaseTimer pt(times, c->stream, &S.ms_bandwidth);
            int it = 0;
            CHK(mkde_mle_fixed_bandwidth(c, S.h0, ec_on, &S.h_opt, &it));
        }
        CHK(mkde_set_bandwidth_fixed(c, S.h_opt));
        CHK(final_ec());
    } else if (o->hoption == 2) {
        float h0 = S.h0, a0 = o->optimise ? -0.5f : o->alpha;
        {
            PhaseTimer pt(times, c->stream, &S.ms_bandwidth);
            if (o->optimise) {
                float out3[3];
                int it = 0;
                CHK(mkde_hooke_jeeves(c, h0, -0.5f, h0 / 10, 0.1f, ec_on, out3, &it));
                h0 = out3[0];
                a0 = out3[1];
            }
            // EC(h0) inside, pilot, g, h_i
            CHK(abramson_core(c, h0, a0, ec_on));
        }
        S.h_opt = h0;
        S.alpha_opt = a0;
        CHK(final_ec());
    }
    {
        PhaseTimer pt(times, c->stream, &S.ms_density);
        CHK(mkde_density_tile_resident(c));
    }
    HIPCHK(hipEventRecord(t1, c->stream));
    HIPCHK(hipEventSynchronize(t1));
    times.finish();
    if (ec_counted) {
        unsigned long long ne = 0;
        CHK(copy_d2h(c, &ne, c->ucount.p + mkde_dev::slot_words() - 1, sizeof ne));
        S.n_ec_pairs = (int64_t)ne;
    }
    if (o->hoption == 2) CHK(copy_d2h(c, &S.g, c->gval.p, sizeof(double)));
    float tot = 0.0f;
    HIPCHK(hipEventElapsedTime(&tot, t0, t1));
    HIPCHK(hipEventElapsedTime(&c->last_density_ms, c->ev0, c->ev1));
    S.ms_total = tot;
    if (st) *st = S;
    return MKDE_OK;
}
