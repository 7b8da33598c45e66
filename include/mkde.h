/*
 * mkde.h -- C ABI of the MI355X-native tiled Gaussian KDE library (libmkde.so).
 *
 * The reference (Guiming/mGPUKDE_tiled) has no library API: its GPU path is a
 * set of host functions inside kde_cuda_kdtr_tiled.cu ("tiled" below) that
 * main() calls in sequence, passing SamplePoints / AsciiRaster structs by value
 * and per-GPU device pointers.  Each entry point here replaces one of those
 * internal seams; the comment above it names the reference function (file:line)
 * whose role it takes.  A host caller of the reference's GPU path binds these.
 *
 * Conventions
 *   - Every function returns MKDE_OK (0) or a positive mkde_status code; the
 *     message of the last failure on the calling thread is mkde_last_error().
 *     The library never exits the process (the reference calls exit() on every
 *     CUDA error, e.g. tiled:1725-1729).
 *   - The caller owns every host buffer; the library owns device buffers, held
 *     in a context (one HIP device, one HIP stream).  Host output pointers may
 *     be NULL when the caller wants the result to stay resident on the device.
 *   - Points keep the caller's order in every input and output array.
 *   - Calls on one context must not overlap (one thread at a time per context);
 *     separate contexts are independent.
 *   - Floating point follows the reference CPU path: single-precision
 *     coordinates, the cut-off predicate fl64(dx^2+dy^2) < 9*fl32(h*h) with
 *     dx = fl32(x1-x0) evaluated exactly; densities at points in FP32 with
 *     compensated folds, the density raster in FP32 partial sums of at most
 *     64 terms folded into FP64, edge-correction integrals from FP64 prefix
 *     sums, FP64 for log-likelihood sums.
 */
#ifndef MKDE_H
#define MKDE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MKDE_OK = 0,
    MKDE_E_ARG = 1,     /* invalid argument (null pointer, bad size, bad option) */
    MKDE_E_HIP = 2,     /* HIP runtime failure */
    MKDE_E_NOMEM = 3,   /* device or host allocation failed */
    MKDE_E_STATE = 4,   /* a required input (points, mask, bandwidths) is not set */
    MKDE_E_IO = 5,      /* file read/write/format error */
    MKDE_E_RANGE = 6,   /* a size exceeds a documented maximum */
    MKDE_E_PENDING = 7, /* a served job has not finished yet */
    MKDE_E_CANCELLED = 8 /* a served job was cancelled before it started */
} mkde_status;

/* Raster geometry of a mask, density tile or output raster.  Mirrors the
 * geometry fields of the reference's AsciiRaster (AsciiRaster.h:9-33):
 * row 0 is the north row, cell centres follow COL_TO_XCOORD/ROW_TO_YCOORD
 * (Utilities.h:13-14) evaluated in double and stored to float. */
typedef struct {
    int64_t ncols;
    int64_t nrows;
    float xll;       /* xLLCorner */
    float yll;       /* yLLCorner */
    float cellsize;
    float nodata;    /* noDataValue: a cell is valid iff value != nodata */
} mkde_grid;

typedef struct mkde_ctx mkde_ctx;

const char *mkde_last_error(void);
const char *mkde_version(void);

/* Device context: one HIP device, one HIP stream (default priority).
 * Replaces the reference's per-GPU stream set-up (tiled:367-400). */
int mkde_ctx_create(int device, mkde_ctx **out);
int mkde_ctx_destroy(mkde_ctx *ctx);
int mkde_synchronize(mkde_ctx *ctx);

/* Device memory held by a context's buffers (bytes; buffers grow to the
 * largest job seen and are reused), and a release of all of them (the
 * context keeps its stream and forgets points, mask, tile and bandwidths).
 * No reference counterpart: used by the job server's HBM admission. */
int mkde_ctx_memory(mkde_ctx *ctx, int64_t *bytes);
int mkde_ctx_release(mkde_ctx *ctx);

/* SamplePoints upload (AllocateDeviceSamplePoints + CopyToDeviceSamplePoints,
 * tiled:1713-1805).  Builds the device-resident cell-binned point index that
 * replaces the reference's k-d tree (BuildGPUKDtree tiled:4260;
 * CUDA_KDTree::CreateKDTree CUDA_KDtree.cu:26-102).  w == NULL means unit
 * weights (SAMPLEWEIGHTS = 0, tiled:1694-1696).  Resets bandwidths, boundary
 * distances and edge-correction weights (weights to 1, tiled:889-893). */
int mkde_points_upload(mkde_ctx *ctx, int64_t n, const float *x, const float *y, const float *w);
int mkde_num_points(mkde_ctx *ctx, int64_t *n);

/* Edge-correction mask upload (the Mask AsciiRaster of tiled:412/429, copied to
 * the device by CopyToDeviceAsciiRaster tiled:2316-2373). */
int mkde_mask_upload(mkde_ctx *ctx, const mkde_grid *grid, const float *values);

/* Rule-of-thumb bandwidth h0 = sqrtf(BandWidth2(points)) (tiled:64-68, 481;
 * MeanCenter tiled:1563, StandardDistance2 tiled:1581), with the reference's
 * float/double sums in point order. */
int mkde_rule_of_thumb(mkde_ctx *ctx, float *h0);

/* MarkBoundary (tiled:3956-4069; dMarkBoundary kde_kernel_kdtr.cu:1209-1281):
 * flags[r*ncols+c] = 1 for boundary cells of the edge-correction mask, else 0.
 * flags may be NULL.  *n_boundary (may be NULL) receives the count. */
int mkde_mark_boundary(mkde_ctx *ctx, uint8_t *flags, int64_t *n_boundary);

/* CalcDist2Boundary (tiled:4072-4137, CPU semantics 4091-4131): squared
 * distance (double) from every point to the nearest boundary cell centre,
 * DBL_MAX if there is no boundary cell.  Runs mkde_mark_boundary if needed. */
int mkde_dist_to_boundary(mkde_ctx *ctx, double *dist);
int mkde_set_dist_to_boundary(mkde_ctx *ctx, const double *dist);

/* Bandwidths: hs[] of the reference (AllocateBandwidths tiled:2878-2885,
 * AllocateDeviceBandwidths tiled:2887-2901, CopyToDeviceBandwidths
 * tiled:2903; Hoption -1 reads them from the CSV, tiled:413-420). */
int mkde_set_bandwidth_fixed(mkde_ctx *ctx, float h);
int mkde_set_bandwidths(mkde_ctx *ctx, const float *h);
int mkde_get_bandwidths(mkde_ctx *ctx, float *h);

/* Edge-correction weights with the current bandwidths
 * (EdgeCorrectionWeightsExact tiled:2694-2807; CalcEdgeCorrectionWeights
 * kde_kernel_kdtr.cu:117-265).  Requires the mask; computes boundary distances
 * first if they are not set.  ew may be NULL; *n_eval (may be NULL) receives
 * the number of (point, valid cell) pairs inside the cut-off. */
int mkde_ec_weights(mkde_ctx *ctx, float *ew, int64_t *n_eval);
/* Set weights explicitly (1.0 everywhere when ew == NULL: EC off). */
int mkde_set_ec_weights(mkde_ctx *ctx, const float *ew);
int mkde_get_ec_weights(mkde_ctx *ctx, float *ew);

/* Range-search candidate sets (KDtree::SearchRange KDtree.cu:471-537 as used by
 * LogLikelihood tiled:3309-3334; DensityAtPointsKdtr kde_kernel_kdtr.cu:681-804):
 * for each query point i, all j with fl64(dx^2+dy^2) < 9*fl32(h_i*h_i) using
 * the current bandwidths.  CSR output: offsets[n+1] and idx[capacity], indexes
 * ascending per query.  With idx == NULL only *total is computed. */
int mkde_range_candidates(mkde_ctx *ctx, int64_t *offsets, int32_t *idx, int64_t capacity, int64_t *total);

/* Density at the sample points with the current bandwidths and EC weights,
 * each SOURCE point using its own bandwidth (the sample-point estimator of
 * DensityAtPointsKdtr + dCopyDensityValues, kde_kernel_kdtr.cu:681-981, and
 * the CPU scatter of LogLikelihood tiled:3312-3347, 3615-3647).
 * loo = 0: den0 (self term included); loo = 1: den1 (self term excluded). */
int mkde_density_at_points(mkde_ctx *ctx, int loo, float *den);

/* Allocation-free form of mkde_density_at_points, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and result; den_dev is the caller's DEVICE buffer of n floats
 * (caller's point order).  The work list, the split partial sums and the
 * bin-order densities live in a static 32 MiB array per device (held until the
 * process ends) instead of context buffers; no hipMalloc / hipFree is called.
 * About 5 million points fit; more return MKDE_E_RANGE before any kernel runs.
 * Every call on one device shares that array: calls that overlap on one GPU
 * (two contexts, two threads) race and must be serialised by the caller. */
int mkde_density_at_points_noalloc(mkde_ctx *ctx, int loo, float *den_dev);

/* Log-likelihood of K fixed bandwidths (LogLikelihood fixed-h,
 * tiled:3214-3368), all candidates in one batched pass over point pairs.
 * ec_on != 0 recomputes EC weights per candidate (UPDATEWEIGHTS).  logL[k] is
 * the FP64 sum of log(max(den1, 1e-14)).  K <= 1024.  Leaves the context's
 * bandwidths and EC weights unchanged. */
int mkde_loo_loglik(mkde_ctx *ctx, int K, const float *h_cand, int ec_on, double *logL);

/* Log-likelihood of the adaptive estimator for (h, alpha) pairs
 * (LogLikelihood adaptive-h, tiled:3374-3671): EC(h) -> pilot density at h ->
 * geometric mean -> h_i = h*(den0_i/g)^alpha -> EC(h_i) -> LOO with per-source
 * bandwidths.  Leaves the context's bandwidths and EC weights unchanged. */
int mkde_loo_loglik_adaptive(mkde_ctx *ctx, int K, const float *h, const float *alpha, int ec_on, double *logL);

/* Abramson adaptive bandwidths (Hoption 2 final step, tiled:1081-1127 and
 * CPU tiled:566-578): EC(h0) if ec_on, pilot density den0 at h0,
 * g = exp(mean log den0), h_i = h0 * powf(den0_i / g, alpha).  Sets the
 * context's bandwidths to h_i (and EC weights to those at h0).  h_out, den0_out
 * may be NULL; *g_out may be NULL. */
int mkde_abramson(mkde_ctx *ctx, float h0, float alpha, int ec_on, float *h_out, float *den0_out, double *g_out);

/* MLE_FixedBandWidth (tiled:3044-3093): the reference's bracketing search on
 * [h/20, 2h] with the LOO log-likelihood.  Likelihoods are compared in FP64.
 * *n_iter (may be NULL) receives the iteration count. */
int mkde_mle_fixed_bandwidth(mkde_ctx *ctx, float h, int ec_on, float *hopt, int *n_iter);

/* hooke_jeeves (tiled:3726-3790) over (h, alpha) with the adaptive
 * log-likelihood.  out3 = (h, alpha, Lmax). */
int mkde_hooke_jeeves(mkde_ctx *ctx, float h0, float alpha0, float step_h, float step_a, int ec_on, float *out3,
                      int *n_iter);

/* Density raster of one tile (the tile loop of the GPU path, tiled:1233-1372
 * mode 1 / 1373-1517 mode 0; KernelDesityEstimation_pPoints
 * kde_kernel_kdtr.cu:461-574; CPU semantics tiled:707-733): for every valid
 * cell of the tile (mask value != nodata; mask == NULL: all cells valid) the
 * sum over points of K(h_i, d) * w_i * ew_i with the current bandwidths and EC
 * weights; nodata cells keep the mask value.  Cell centres use the tile's own
 * origin (N2).  out (host) receives nrows*ncols floats; may be NULL (the
 * result stays on the device).  *n_eval (may be NULL) is not computed here. */
int mkde_density_tile(mkde_ctx *ctx, const mkde_grid *tile, const float *mask, float *out);

/* Exact number of (point, valid cell) pairs inside the cut-off for a tile
 * (the algorithmic work unit of the density kernel), computed by a separate
 * counting pass. */
int mkde_count_density_pairs(mkde_ctx *ctx, const mkde_grid *tile, const float *mask, int64_t *n_pairs);

/* Device-resident variants used for throughput measurement: the tile mask is
 * uploaded once (mkde_tile_upload), mkde_density_tile_resident computes into a
 * device buffer owned by the context, mkde_tile_download copies it out. */
int mkde_tile_upload(mkde_ctx *ctx, const mkde_grid *tile, const float *mask);
int mkde_density_tile_resident(mkde_ctx *ctx);
int mkde_tile_download(mkde_ctx *ctx, float *out);

/* Every tile of a GeoTIFF-style raster in one pass (replaces the tile loop of
 * main(), tiled:1235-1309 with GeoTIFFReader::GetRasterBand_NextTile,
 * GeoTIFF.cpp:333-476, and AsciiRasterFromGeoTIFFTile, tiled:2193-2262).
 * The raster has ncols x nrows cells and geotransform gt[6] (GDAL order);
 * tiles are tile_cols x tile_rows cells from the top-left, edge tiles
 * smaller.  Each tile's cells use that tile's origin exactly as the
 * reference builds it: xll = (float)(gt[0] + gt[1]*xoff), yll =
 * (float)((gt[3] + gt[5]*yoff) + rows*gt[5]), cellsize = (float)gt[1], so
 * the result equals mkde_density_tile on each tile.  mask (host, nrows*ncols
 * floats, row-major) may be NULL (all cells valid); nodata cells keep the mask
 * value.  mkde_raster_upload leaves the raster as the context's current
 * tile (for mkde_density_tile_resident / mkde_tile_download / mkde_run_job);
 * mkde_density_raster uploads, computes and, if out != NULL, copies out. */
/* The density tile is the resident edge-correction mask raster (same grid,
 * same values; no second upload).  A later mkde_mask_upload unsets it. */
int mkde_tile_from_mask(mkde_ctx *ctx);

int mkde_raster_upload(mkde_ctx *ctx, int64_t ncols, int64_t nrows, const double *gt, float nodata,
                       int64_t tile_cols, int64_t tile_rows, const float *mask);
int mkde_density_raster(mkde_ctx *ctx, int64_t ncols, int64_t nrows, const double *gt, float nodata,
                        int64_t tile_cols, int64_t tile_rows, const float *mask, float *out);

/* Diagnostic counters of the density kernel (no reference counterpart).
 * enable != 0 zeroes and enables them; counters (8 entries, may be NULL)
 * receives: [0] candidate points classified, [1] FULL and [2] PARTIAL
 * (point, 64x64 block) pairs, [3] FULL (point, 32x32 wave quadrant) pairs
 * (the quadrants of block-FULL points included), [4] PARTIAL and [5] EMPTY
 * (point, quadrant) pairs of block-PARTIAL points, [6] blocks without a valid
 * cell; enable == 0 reads them and switches them off. */
int mkde_density_stats(mkde_ctx *ctx, int enable, int64_t *counters);

/* Per-context options (no reference counterpart; measurement and tuning):
 *   "search_graph"   1 (default): a leave-one-out search runs as one captured
 *                    HIP graph, replayed while the resident inputs and K stay
 *                    the same; 0: the host-sized launches (same results)
 *   "density_order"  1 (default): density blocks visited heaviest first with
 *                    heavy blocks split; 0: raster order, no split
 *   "points_per_bin" target points per bin of the point index (default 64),
 *                    used by the next mkde_points_upload
 *   "ec_chunk"       512, 1024 (default) or 2048 columns per EC prefix chunk
 *   "ec_wide_min"    boxes wider than this (default: ec_chunk) use the wide
 *                    EC kernel
 *   "loo_split_target", "pts_split_target", "density_split_target"
 *                    split targets of the leave-one-out, point-density and
 *                    density-raster work lists (defaults 32768, 16384, 4096)
 * Split targets and chunking group the FP32 partial sums differently, so
 * results can move by a few ulps (never the cut-off sets).  Defaults can be
 * set from the environment (MKDE_EC_CHUNK, MKDE_EC_WIDE_MIN, MKDE_LOO_TARGET,
 * MKDE_PTS_TARGET, MKDE_DENSITY_TARGET, MKDE_PTS_PER_BIN).  Unknown names
 * return MKDE_E_ARG, values out of range MKDE_E_RANGE. */
int mkde_set_option(mkde_ctx *ctx, const char *name, int64_t value);
int mkde_get_option(mkde_ctx *ctx, const char *name, int64_t *value);

/* Kernel timing of the last density launch set on the context's stream
 * (milliseconds, HIP events around the dominant kernel). */
int mkde_last_kernel_ms(mkde_ctx *ctx, const char *which, double *ms);

/* ------------------------------------------------------------------ jobs --
 * One KDE job on the context's resident points, mask and tile: the sequence
 * of the reference's GPU block in main() (tiled:877-1517) without any host
 * staging.  Bandwidth options (tiled:280-284):
 *   MKDE_H_FILE (-1)  use the bandwidths already set on the context
 *   MKDE_H_RULE (0)   rule-of-thumb h0 for every point
 *   MKDE_H_MLE  (1)   MLE_FixedBandWidth from h0
 *   MKDE_H_ADAPTIVE (2) hooke_jeeves from (h0, alpha0 = -0.5) when optimise
 *                     != 0, then Abramson bandwidths from the result; with
 *                     optimise == 0 the Abramson step runs directly at
 *                     (h0, alpha) -- the fixed-alpha adaptive estimator.
 * Edge correction (UPDATEWEIGHTS) recomputes boundary distances when
 * recompute_boundary != 0.  The density of the uploaded tile is left in the
 * context (mkde_tile_download copies it out). */
typedef struct {
    int hoption;
    int edge_correction;
    int optimise;
    float alpha;
    int recompute_boundary;
} mkde_job_opts;

typedef struct {
    float h0;              /* rule-of-thumb bandwidth */
    float h_opt;           /* MLE result (hoption 1) or optimised h0 (hoption 2) */
    float alpha_opt;       /* optimised alpha (hoption 2) */
    double g;              /* geometric mean of the pilot densities (hoption 2) */
    int64_t n_ec_pairs;    /* (point, valid cell) pairs of the last EC pass */
    double ms_boundary;    /* MarkBoundary + CalcDist2Boundary */
    double ms_ec;          /* all EC passes */
    double ms_bandwidth;   /* bandwidth selection (pilot, optimisers) */
    double ms_density;     /* density raster kernel */
    double ms_total;       /* whole job, device time on the context stream */
} mkde_job_stats;

int mkde_run_job(mkde_ctx *ctx, const mkde_job_opts *opts, mkde_job_stats *stats);

/* Free / total device memory of the context's GPU (the reference sizes its
 * raster tiles from cudaMemGetInfo, tiled:1152-1165). */
int mkde_device_memory(mkde_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes);

/* Autotuner (no reference counterpart): coordinate descent over the options
 * above on the resident workload, timed on this device (best of reps runs per
 * value; a value is kept when it is more than 1% faster than the current
 * choice).  mkde_autotune_job times mkde_run_job(opts) (needs points, mask and
 * tile); mkde_autotune_search times mkde_loo_loglik(K, h_cand, ec_on).  The
 * context keeps the chosen options, which are also recorded per job class
 * (log2 buckets of the point count and raster cells, bandwidth option --
 * MKDE_H_SEARCH for searches --, edge correction, optimiser);
 * mkde_autotune_lookup applies a recorded choice to a context (MKDE_E_STATE
 * if the class has none).  report (may be NULL) receives a JSON text of every
 * timing and the choices; MKDE_E_RANGE if report_cap is too small.  Trying
 * "points_per_bin" rebuilds the point index from the uploaded points (as
 * mkde_points_upload: bandwidths, boundary distances and EC weights are
 * reset); jobs with hoption -1 (caller's bandwidths) do not try it.  When a
 * trial fails, its option is set back to its value before the trial and the
 * error is returned. */
int mkde_autotune_job(mkde_ctx *ctx, const mkde_job_opts *opts, int reps, char *report, int64_t report_cap);
int mkde_autotune_search(mkde_ctx *ctx, int K, const float *h_cand, int ec_on, int reps, char *report,
                         int64_t report_cap);
int mkde_autotune_lookup(mkde_ctx *ctx, int64_t n, int64_t cells, int hoption, int edge_correction, int optimise);

/* ------------------------------------------------------------------- I/O --
 * GeoTIFF and CSV I/O of the command line (the reference uses GDAL through
 * GeoTIFFReader / GeoTIFFWriter, GeoTIFF.cpp:20-577, and ReadSamplePoints /
 * WriteSamplePoints, tiled:1623-1711, 1857-1870).  Errors go to
 * mkde_last_error(). */

/* Dimensions, geotransform (GDAL order), nodata and block size of band 1
 * (GeoTIFFReader::GetDimensions / GetGeoTransform / GetNoDataValue /
 * GetBlockXYSize).  Any output pointer may be NULL. */
int mkde_tiff_info(const char *path, int64_t *ncols, int64_t *nrows, double *geotransform, double *nodata,
                   int *has_nodata, int *block_x, int *block_y);
/* A window of band 1 converted to float (GetRasterBand_NextTile). */
int mkde_tiff_read(const char *path, int64_t xoff, int64_t yoff, int64_t ncols, int64_t nrows, float *out);

/* Float32 BigTIFF writer, LZW compressed, with georeferencing, nodata and
 * band statistics (GeoTIFFWriter): open, put windows (WriteGeoTIFF_NextTile),
 * close.  georef_from (may be NULL) names a GeoTIFF whose GeoKey tags
 * (projection) are copied. */
typedef struct mkde_raster_writer mkde_raster_writer;
int mkde_writer_open(const char *path, int64_t ncols, int64_t nrows, const double *geotransform, float nodata,
                     const char *georef_from, mkde_raster_writer **out);
int mkde_writer_put(mkde_raster_writer *w, int64_t xoff, int64_t yoff, int64_t ncols, int64_t nrows,
                    const float *data);
int mkde_writer_close(mkde_raster_writer *w, int compress);

/* ReadSamplePoints: header line then x,y,[h,]w per line.  Call with x == NULL
 * to count (*n receives the count), then with buffers of *n entries. */
int mkde_csv_read_points(const char *path, int with_h, int sample_weights, int64_t *n, float *x, float *y, float *h,
                         float *w);
/* WriteSamplePoints(x, y, h, w) ("x, y, h, w" header, %f fields). */
int mkde_csv_write_points(const char *path, int64_t n, const float *x, const float *y, const float *h,
                          const float *w);

/* TIFF LZW codec (TIFF 6.0 section 13); returns the output length or -1. */
int64_t mkde_lzw_encode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap);
int64_t mkde_lzw_decode(const uint8_t *in, int64_t n, uint8_t *out, int64_t cap);

/* --------------------------------------------------------------- serving --
 * A job server on one device (no reference counterpart: the reference runs
 * one job per process).  `slots` worker threads each own a context (its own
 * stream; device buffers grow to the largest job seen and are reused).  Jobs
 * start in submission order when a slot is free and HBM admission allows it
 * (a later job that fits may start ahead of a head that does not, at most 16
 * times per head):
 * the device bytes held by all slots, plus the footprint bound
 * (mkde_job_footprint) of every running job, plus the new job's bound must
 * not exceed hbm_budget (<= 0: 80% of the device's free memory at creation).
 * A job whose bound alone exceeds the budget is refused at submit with
 * MKDE_E_NOMEM; when the next job does not fit and no job runs, idle slots
 * release their buffers first.  Jobs share nothing on the device, so a job's
 * result is the same alone and in any mix.  The caller's input and output
 * buffers must stay valid until the job has finished (mkde_server_wait). */
typedef struct mkde_server mkde_server;

#define MKDE_H_SEARCH 3   /* hoption: argmax LOO log-likelihood over candidates */

typedef struct {
    int64_t n;
    const float *x, *y;    /* n points */
    const float *w;        /* sample weights, NULL: all 1 */
    const float *h;        /* hoption -1: n bandwidths, else NULL */
    mkde_grid grid;        /* edge-correction mask and density raster geometry */
    const float *mask;     /* nrows*ncols values (nodata: outside); NULL: all cells
                              valid, allowed only without edge correction */
    int hoption;           /* -1, 0, 1, 2 as mkde_run_job, or MKDE_H_SEARCH */
    int edge_correction;
    int optimise;          /* hoption 2: Hooke-Jeeves first */
    float alpha;           /* hoption 2 without optimise */
    int n_candidates;      /* MKDE_H_SEARCH: K <= 1024 bandwidths */
    float cand_lo, cand_hi; /* h_k = (float)(h0*lo*(hi/lo)^(k/(K-1))), h0 = rule of thumb */
    float *out;            /* nrows*ncols density raster (host) */
    int one_candidate_per_pass; /* MKDE_H_SEARCH: 0 = up to 16 candidates per pass over the
                                   point pairs (batched); 1 = one pass per candidate, as the
                                   reference's per-candidate LogLikelihood calls (same result) */
} mkde_serve_job;

typedef struct {
    int status;            /* MKDE_OK, the job's error, or MKDE_E_CANCELLED */
    char message[256];     /* error text of a failed job */
    int slot;              /* worker slot that ran it (-1: never started) */
    double queue_ms, run_ms, latency_ms;   /* submit -> start -> finish */
    float h0;              /* rule-of-thumb bandwidth */
    float h_sel, alpha_sel; /* selected bandwidth (MLE / search / Hooke-Jeeves h0) and alpha */
    int64_t footprint;     /* admission bound used for the job, bytes */
} mkde_serve_result;

typedef struct {
    int slots;
    int64_t budget;        /* bytes */
    int64_t held;          /* device bytes held by all slots now */
    int64_t peak_committed; /* max over time of held + bounds of running jobs */
    int64_t submitted, finished, refused, releases;
    int64_t queued, running;
} mkde_server_info;

/* Admission bound of a job's device memory (bytes) on a fresh context. */
int64_t mkde_job_footprint(int64_t n, int64_t ncols, int64_t nrows, int hoption, int optimise, int n_candidates);
int mkde_server_create(int device, int slots, int64_t hbm_budget, mkde_server **out);
int mkde_server_submit(mkde_server *s, const mkde_serve_job *job, int64_t *ticket);
/* MKDE_OK: the job has finished (res filled, ticket retired); MKDE_E_PENDING:
 * not within timeout_ms (< 0: wait without limit). */
int mkde_server_wait(mkde_server *s, int64_t ticket, double timeout_ms, mkde_serve_result *res);
/* Cancels a job that has not started (MKDE_E_STATE if it has). */
int mkde_server_cancel(mkde_server *s, int64_t ticket);
int mkde_server_info_get(mkde_server *s, mkde_server_info *info);
/* Cancels queued jobs, waits for running ones, frees everything. */
int mkde_server_destroy(mkde_server *s);

#ifdef __cplusplus
}
#endif

#endif /* MKDE_H */
