// mkde_device.h -- types shared by the HIP kernels (mkde_kernels.hip) and the
// host library (mkde_capi.cpp, mkde_serve.cpp).  Plain structs only: no torch, no STL.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Utilities.h:8 -- the reference's TWO_PI is a float literal.
#define MKDE_TWO_PI_F 6.2831853071795862f
#define MKDE_LOG2E 1.4426950408889634073599
// Utilities.h:19 -- EPSILONDENSITY (a float literal).
#define MKDE_EPS_DEN 0.00000000000001f
// candidate bandwidths per batched leave-one-out pass
#define MKDE_LM_K 16
// static device scratch of the allocation-free entry point (bytes per device)
// and the device ordinals it is resolved for
#define MKDE_SCRATCH_BYTES (32u << 20)
#define MKDE_MAX_DEVICES 64

// Launch-shape settings of one context: defaults from the environment
// (default_tuning), changed by mkde_set_option or the autotuner.  They set how
// work is cut into pieces, so a change can move FP32 rounding (never the
// cut-off sets or the error bound).
struct Tuning {
    int ec_chunk;                    // 512, 1024 or 2048: columns per prefix chunk of k_ec
    int ec_wide_min;                 // boxes wider than this run in k_ec_wide
    unsigned long long loo_target;   // split target of the batched leave-one-out work list
    unsigned long long pts_target;   // split target of the point-density work list
    unsigned long long dt_target;    // split target of the density-raster work list
};

// Smallest split threshold of the point work lists (candidate sources per part).
#define MKDE_PTS_TMIN 4096ull

// Most items and split parts a point work list of nw 64-query waves can have:
// each wave's cost is at most n sources, the threshold is T = max(tmin,
// total / target), so split parts <= min(target, total / tmin) + the number
// of split waves.
inline void point_work_bounds(int64_t n, unsigned long long target, int64_t *items_max, int64_t *slots_max)
{
    const int64_t nw = (n + 63) / 64;
    const double tot = (double)nw * (double)n;
    const int64_t by_t = (int64_t)((tot + MKDE_PTS_TMIN - 1) / (double)MKDE_PTS_TMIN);
    const int64_t parts = by_t < (int64_t)target ? by_t : (int64_t)target;
    const int64_t split_waves = by_t < nw ? by_t : nw;
    *items_max = nw + parts + 1;
    *slots_max = parts + split_waves + 1;
}

// Raster geometry (AsciiRaster.h:9-33 geometry fields).
struct GridDesc {
    int64_t ncols, nrows;
    float xll, yll, cs, nodata;
};

// Uniform world-space bin grid over the points' bounding box.  Bin (bx, by)
// covers [x0 + bx*size, x0 + (bx+1)*size) x [y0 + by*size, ...).
struct BinGrid {
    double x0, y0, size;
    int nbx, nby;
};

// One sample point in bin order, with the per-job derived constants of the
// Gaussian kernel: coef = w*ew / (2*pi*h2f), a = -log2(e) / (2*h2f),
// thr = 9*h2f, where h2f = (double)fl32(h*h).  A point with h <= 0 has
// thr = -1 (it can never be inside the cut-off) and coef = 0.
struct __attribute__((aligned(32))) PtRec {
    float x, y, coef, pad;
    double a, thr;
};

// Per-bin summary used for culling: bounding box of the bin's points and the
// largest cut-off threshold 9*h2f among them.
struct BinInfo {
    float x0, y0, x1, y1;
};

struct PointIndex {
    const PtRec *pts;       // n, bin order
    const int *bin_start;   // nbins + 1
    const BinInfo *bin_box; // nbins
    const double *bin_thr;  // nbins
    BinGrid bg;
    const double *reach_p;  // device scalar: max over bins of sqrt(bin_thr), world units
    int64_t n;
};

// Density blocks: 64 rows x MKDE_DT_BW columns (64 or 128), one workgroup
// each, one 32 x 32 quadrant per wave (MKDE_DT_BW / 16 waves).
#define MKDE_DT_BW 64
#define MKDE_DT_CELLS (64 * MKDE_DT_BW)

// A raster processed as tiles of tw x th cells, each tile with its own origin
// from the raster geotransform (GDAL order gt0, gt1, gt3, gt5).  Density
// blocks are numbered tile-major: bpx blocks per tile row, bpt per tile, ntx
// tiles per raster row (edge tiles leave some blocks empty).
// tw == 0 means one tile described by its GridDesc.
struct Tiling {
    int64_t tw, th;
    int ntx, bpx, bpt;
    double gt0, gt1, gt3, gt5;
};

// Device buffers of the density kernel's work list (per density block unless
// noted): candidate counts, split parts, scans of parts / split parts / split
// flags, scan scratch, the candidate total and the {items, slots, splits}
// counts read back by the host.
struct DensityWork {
    unsigned *cost, *parts, *item0, *slot0, *sidx, *sidx2;
    unsigned *chunks;              // (nblk + 1023) / 1024
    unsigned long long *total;     // 1
    unsigned *counts;              // 3
    unsigned *hist;                // 32
};

// Boundary-cell hierarchy for exact nearest-boundary distances: leaves of
// 16 x 16 cells hold a 256-bit mask of their boundary cells; every level keeps
// the bounding box {x0, y0, x1, y1} of the boundary cell centres below a node
// (x0 > x1: empty); level l has nx[l] x ny[l] nodes starting at off[l].
#define MKDE_BT_S 16
#define MKDE_BT_MAXL 15
struct BndTree {
    const uint4 *bits;     // 2 per leaf (256 bits, row-major within the leaf)
    const float4 *box;     // all levels
    int nlev;
    int nx[MKDE_BT_MAXL], ny[MKDE_BT_MAXL];
    int64_t off[MKDE_BT_MAXL];
};

// Per-row runs of valid cells of a mask (CSR): runs of row r are
// run_se[run_ptr[r] .. run_ptr[r+1]), each (start, end) inclusive, ascending.
struct MaskRuns {
    const int64_t *run_ptr;
    const int2 *run_se;
};
