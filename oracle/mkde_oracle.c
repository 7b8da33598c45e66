/*
 * mkde_oracle.c -- CPU restatement of the reference's OpenMP/sequential KDE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product
 * library (libmkde) never links, loads or calls anything here.
 *
 * Every function restates one piece of the reference CPU path and cites the
 * reference file:line it follows (files are named relative to the reference
 * root: kde_cuda_kdtr_tiled.cu is "tiled", Utilities.h, KDtree.h/.cu).
 *
 * Two accumulation modes where the reference accumulates:
 *   ORC_REF32 (0): the reference's float arithmetic and a fixed loop order
 *                  (the reference's OpenMP scatter loops race; this restates
 *                  them in the order a single thread would run them).
 *   ORC_FP64  (1): every kernel value and sum in double -- the error oracle.
 *
 * Parity pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_col_x / orc_row_y        closed form on exact grids, Redwood geotransform
 *   orc_distance2 / orc_kernel   closed forms; K(0) = 1/(2*pi*h^2); cut-off edge d^2 = 9h^2
 *   orc_rule_of_thumb            closed form on symmetric point sets
 *   orc_mark_boundary            hand-built masks with hand-written flags
 *   orc_dist_to_boundary         numpy brute force; hand-computed small cases
 *   orc_ec_weights               closed form (point far inside: mass of the truncated
 *                                Gaussian), half-plane symmetry, numpy brute force
 *   orc_density_tile             single-point closed form, mass invariant, numpy brute force
 *   orc_range_candidates         the reference KDtree (compiled from its own sources
 *                                into oracle/_ref) on random and adversarial inputs
 *   orc_loo_density              den0 - den1 = w*ew/(2*pi*h2f); symmetry; numpy brute force
 *   orc_pilot_density            numpy brute force; den0 >= self term
 *   orc_pilot_density_queries    orc_pilot_density on the same queries (bit-identical)
 *   orc_point_density_binned     orc_loo_density / orc_pilot_density (same pair counts,
 *                                FP64 sums) on random and integer-lattice inputs
 *   orc_compgml / orc_abramson   closed forms (constant density -> h_i = h0)
 *   orc_golden_section / orc_hooke_jeeves   synthetic concave likelihoods with known maximiser
 *   orc_msvc_rand_points         the published MSVC CRT rand() sequence (41, 18467, 6334, ...)
 *   orc_tile_point_filter        superset property against brute force contributions
 * Parity unpinned at ulp level (no reference-held output exists): the reference's
 * expf/logf/powf (CUDA 11.7 device math and the MSVC CRT), its qsort order after
 * SortSamplePoints (tiled:4152-4189, comparator reads doubles as floats), and its
 * racy OpenMP scatter order (tiled:3319-3346).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_TWO_PI_F 6.2831853071795862f   /* Utilities.h:8 (float literal) */
#define ORC_CUT 9.0                         /* Utilities.h:29 CUT_OFF_FACTOR */
#define ORC_SQRT_CUT 3.0                    /* Utilities.h:30 SQRT_CUT_OFF_FACTOR */
#define ORC_EPS_DEN 0.00000000000001f       /* Utilities.h:19 EPSILONDENSITY */

enum { ORC_REF32 = 0, ORC_FP64 = 1 };

/* Raster geometry, as the reference's AsciiRaster (AsciiRaster.h:9-33). */
typedef struct {
    int64_t ncols, nrows;
    float xll, yll, cs, nodata;
} orc_grid;

int orc_version(void) { return 3; }

/* ---------------------------------------------------------------- N2 --- */
/* COL_TO_XCOORD / ROW_TO_YCOORD (Utilities.h:13-14): evaluated in double,
 * stored to float by the callers (tiled:709-712, 2731-2732). */
float orc_col_x(const orc_grid *g, int64_t col)
{
    return (float)((double)g->xll + ((double)col + 0.5) * (double)g->cs);
}

float orc_row_y(const orc_grid *g, int64_t row)
{
    return (float)((double)g->yll + (((double)(uint64_t)g->nrows - 0.5) - (double)row) * (double)g->cs);
}

/* ---------------------------------------------------------------- a3 --- */
/* Distance2 (tiled:50-55): float subtraction, double squares and sum. */
double orc_distance2(float x0, float y0, float x1, float y1)
{
    double dx = (double)(float)(x1 - x0);
    double dy = (double)(float)(y1 - y0);
    return dx * dx + dy * dy;
}

/* GaussianKernel (tiled:71-76): cut-off d2 >= 9*h2 -> 0; expf of a float
 * argument; the division by h2*TWO_PI in double, result stored to float. */
float orc_kernel(double h2, double d2)
{
    if (d2 >= ORC_CUT * h2) return 0.0f;
    return (float)((double)expf((float)(d2 / (-2.0 * h2))) / (h2 * (double)ORC_TWO_PI_F));
}

double orc_kernel64(double h2, double d2)
{
    if (d2 >= ORC_CUT * h2) return 0.0;
    return exp(d2 / (-2.0 * h2)) / (h2 * (double)ORC_TWO_PI_F);
}

static inline double h2f(float h) { return (double)(float)(h * h); }

/* ---------------------------------------------------------------- N9 --- */
/* MeanCenter (tiled:1563-1578), StandardDistance2 (tiled:1581-1595),
 * BandWidth2 (tiled:64-68), h0 = sqrtf(BandWidth2) (tiled:481). */
float orc_rule_of_thumb(int64_t n, const float *x, const float *y, const float *w)
{
    float sum_x = 0.0f, sum_y = 0.0f, sum_w = 0.0f;
    for (int64_t p = 0; p < n; p++) {
        float wp = w[p];
        sum_w += wp;
        sum_x += wp * x[p];
        sum_y += wp * y[p];
    }
    float mx = sum_x / sum_w, my = sum_y / sum_w;
    double sum2 = 0.0, sw = 0.0;
    for (int64_t p = 0; p < n; p++) {
        float wp = w[p];
        sw += wp;
        sum2 += (double)wp * orc_distance2(mx, my, x[p], y[p]);
    }
    double d2 = sum2 / sw;
    double bw2 = (double)sqrtf(2.0f / (float)(3u * (unsigned)n)) * d2;
    return sqrtf((float)bw2);
}

/* ---------------------------------------------------------------- N5 --- */
/* MarkBoundary, CPU branch (tiled:3993-4059): a valid cell is a boundary cell
 * iff it is in an outer row/column or one of its 8 neighbours is nodata.
 * flags: 1 boundary, 0 otherwise.  Returns the number of boundary cells. */
int64_t orc_mark_boundary(const orc_grid *g, const float *mask, uint8_t *flags)
{
    int64_t nr = g->nrows, nc = g->ncols, cnt = 0;
    float nd = g->nodata;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
    for (int64_t r = 0; r < nr; r++) {
        for (int64_t c = 0; c < nc; c++) {
            int64_t i = r * nc + c;
            uint8_t f = 0;
            if (mask[i] != nd) {
                if (r == 0 || r == nr - 1 || c == 0 || c == nc - 1) {
                    f = 1;
                } else {
                    for (int dr = -1; dr <= 1 && !f; dr++)
                        for (int dc = -1; dc <= 1; dc++)
                            if (mask[(r + dr) * nc + (c + dc)] == nd) { f = 1; break; }
                }
            }
            flags[i] = f;
            cnt += f;
        }
    }
    return cnt;
}

/* ---------------------------------------------------------------- N6 --- */
/* CalcDist2Boundary, CPU branch (tiled:4091-4131): min over boundary cells of
 * Distance2(point, cell centre) in double, DBL_MAX if there is none. */
void orc_dist_to_boundary(int64_t n, const float *x, const float *y, const orc_grid *g,
                          const uint8_t *flags, double *dist)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t p = 0; p < n; p++) {
        double m = DBL_MAX;
        for (int64_t r = 0; r < g->nrows; r++) {
            float cy = orc_row_y(g, r);
            for (int64_t c = 0; c < g->ncols; c++) {
                if (!flags[r * g->ncols + c]) continue;
                double d2 = orc_distance2(x[p], y[p], orc_col_x(g, c), cy);
                if (d2 < m) m = d2;
            }
        }
        dist[p] = m;
    }
}

/* ---------------------------------------------------------------- N4 --- */
/* The NARROW box of EdgeCorrectionWeightsExact (tiled:2716-2727, 2774-2786)
 * with the macro semantics of Utilities.h:16-17,30-31: YCOORD_TO_ROW mixes a
 * size_t row count with an int, MIN/MAX are unparenthesised ternaries, so the
 * comparisons for row_upper/col_upper are unsigned 64-bit. */
void orc_narrow_box(const orc_grid *g, float px, float py, float h, int *rl, int *ru, int *cl, int *cu)
{
    uint64_t nR = (uint64_t)g->nrows, nC = (uint64_t)g->ncols;
    double yll = g->yll, xll = g->xll, cs = g->cs, hh = h;
    int k_top = (int)(((double)py + ORC_SQRT_CUT * hh - yll) / cs);
    int r = (int)(int32_t)(uint32_t)(nR - 1 - (uint64_t)(int64_t)k_top);
    *rl = 0 < r ? r : 0;
    int k_bot = (int)(((double)py - ORC_SQRT_CUT * hh - yll) / cs);
    uint64_t yu = nR - 1 - (uint64_t)(int64_t)k_bot;
    uint64_t ru64 = (nR - 1 < yu) ? nR - 1 : yu;
    *ru = (int)(int32_t)(uint32_t)ru64;
    int k_left = (int)(((double)px - ORC_SQRT_CUT * hh - xll) / cs);
    *cl = 0 < k_left ? k_left : 0;
    int k_right = (int)(((double)px + ORC_SQRT_CUT * hh - xll) / cs);
    uint64_t kr = (uint64_t)(int64_t)k_right;
    uint64_t cu64 = (nC - 1 < kr) ? nC - 1 : kr;
    *cu = (int)(int32_t)(uint32_t)cu64;
}

/* EdgeCorrectionWeightsExact (tiled:2694-2748 fixed h, 2750-2807 per-point h;
 * the fixed-h variant is the per-point one with equal h).  Skip if the point's
 * boundary distance >= 9*h2 (weight 1); otherwise sum K(h2,d2)*cellArea over
 * valid cells of the box in row-major order; weight = (float)(1.0/ew).
 * ew64 (optional) is the FP64 weight; n_eval (optional) counts the (point,
 * valid cell, d2 < 9*h2) pairs evaluated. */
void orc_ec_weights(int64_t n, const float *x, const float *y, const double *dist, const float *h,
                    const orc_grid *g, const float *mask, float *ew, double *ew64, int64_t *n_eval)
{
    float cellArea = g->cs * g->cs;
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t p = 0; p < n; p++) {
        double hh2 = h2f(h[p]);
        if (dist[p] >= ORC_CUT * hh2) {
            ew[p] = 1.0f;
            if (ew64) ew64[p] = 1.0;
            continue;
        }
        int rl, ru, cl, cu;
        orc_narrow_box(g, x[p], y[p], h[p], &rl, &ru, &cl, &cu);
        float s32 = 0.0f;
        double s64 = 0.0;
        int64_t cnt = 0;
        for (int row = rl; row <= ru; row++) {
            float cy = orc_row_y(g, row);
            for (int col = cl; col <= cu; col++) {
                size_t idx = (size_t)row * (size_t)g->ncols + (size_t)col;
                if (mask[idx] == g->nodata) continue;
                double d2 = orc_distance2(x[p], y[p], orc_col_x(g, col), cy);
                s32 += orc_kernel(hh2, d2) * cellArea;
                if (d2 < ORC_CUT * hh2) {
                    s64 += orc_kernel64(hh2, d2) * (double)cellArea;
                    cnt++;
                }
            }
        }
        ew[p] = (float)(1.0 / s32);
        if (ew64) ew64[p] = 1.0 / s64;
        total += cnt;
    }
    if (n_eval) *n_eval = total;
}

/* ---------------------------------------------------------------- N3 --- */
/* The per-cell KDE loop of the CPU path (tiled:707-733 mode 1, 806-831 mode 0):
 * for every valid cell of the tile, sum over ALL points in point order of
 * GaussianKernel(hp*hp, d2) * w * ew in float.  Nodata cells keep the mask
 * value.  tmask == NULL means every cell is valid (mode-0 tiles, tiled:1398).
 * out64 (optional): FP64 densities; n_eval (optional): contributing pairs. */
void orc_density_tile(int64_t n, const float *x, const float *y, const float *w, const float *ew,
                      const float *h, const orc_grid *t, const float *tmask, float *out, double *out64,
                      int64_t *n_eval)
{
    int64_t total = 0;
    double *hh2 = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    for (int64_t p = 0; p < n; p++) hh2[p] = h2f(h[p]);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int64_t row = 0; row < t->nrows; row++) {
        float cy = orc_row_y(t, row);
        for (int64_t col = 0; col < t->ncols; col++) {
            size_t idx = (size_t)row * (size_t)t->ncols + (size_t)col;
            if (tmask && tmask[idx] == t->nodata) {
                out[idx] = tmask[idx];
                if (out64) out64[idx] = tmask[idx];
                continue;
            }
            float cx = orc_col_x(t, col);
            float den = 0.0f;
            double d64 = 0.0;
            int64_t cnt = 0;
            for (int64_t p = 0; p < n; p++) {
                double d2 = orc_distance2(x[p], y[p], cx, cy);
                den += orc_kernel(hh2[p], d2) * w[p] * ew[p];
                if (out64 && d2 < ORC_CUT * hh2[p]) {
                    d64 += orc_kernel64(hh2[p], d2) * (double)w[p] * (double)ew[p];
                }
                if (d2 < ORC_CUT * hh2[p]) cnt++;
            }
            out[idx] = den;
            if (out64) out64[idx] = d64;
            total += cnt;
        }
    }
    free(hh2);
    if (n_eval) *n_eval = total;
}

/* ---------------------------------------------------------------- N1 --- */
/* The range-search candidate predicate (KDtree.h:77-82, KDtree.cu:471-537;
 * range = 9*h2 with h2 = h*h of a float h, tiled:3309-3310, 3616): j is a
 * candidate of query i iff fl64(dx^2+dy^2) < 9*(double)fl32(h_i*h_i), with
 * dx = fl32(x_i - x_j).  Includes i itself and duplicates.  Brute force.
 * Pass offsets == NULL to count only; returns the total count.  Within each
 * query the indexes are ascending. */
int64_t orc_range_candidates(int64_t n, const float *x, const float *y, const float *h, int64_t *offsets,
                             int32_t *idx)
{
    int64_t total = 0;
    int64_t *cnt = (int64_t *)calloc((size_t)(n + 1), sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        double range = ORC_CUT * h2f(h[i]);
        int64_t c = 0;
        for (int64_t j = 0; j < n; j++)
            if (orc_distance2(x[j], y[j], x[i], y[i]) < range) c++;
        cnt[i] = c;
    }
    for (int64_t i = 0; i < n; i++) total += cnt[i];
    if (offsets) {
        offsets[0] = 0;
        for (int64_t i = 0; i < n; i++) offsets[i + 1] = offsets[i] + cnt[i];
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < n; i++) {
            double range = ORC_CUT * h2f(h[i]);
            int64_t o = offsets[i];
            for (int64_t j = 0; j < n; j++)
                if (orc_distance2(x[j], y[j], x[i], y[i]) < range) idx[o++] = (int32_t)j;
        }
    }
    free(cnt);
    return total;
}

/* ---------------------------------------------------------------- N7 --- */
/* Leave-one-out density at points, the CPU branch of LogLikelihood (fixed h
 * tiled:3300-3366; adaptive tiled:3612-3647 with each SOURCE i using its own
 * h_i).  The reference scatters: den[j] starts at -K(h2_j,0)*w_j*ew_j
 * (double product stored to float) and every source i adds
 * fl32(K(h2_i,d2)*w_i*ew_i) to each j in range(i).  Restated race-free as a
 * gather over sources in ascending i, which is the exact float sequence of the
 * single-thread scatter.  FP64: den1_j = sum over i != j in range of K64*w_i*ew_i.
 * den0 = den1 + K(h2_j,0)*w_j*ew_j (tiled:3356-3358, 3654-3657). */
void orc_loo_density(int64_t n, const float *x, const float *y, const float *w, const float *ew,
                     const float *h, float *den1, double *den1_64, float *den0, double *den0_64)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t j = 0; j < n; j++) {
        double h2j = h2f(h[j]);
        float d32 = (float)(-1.0 * (double)orc_kernel(h2j, 0.0) * (double)w[j] * (double)ew[j]);
        double d64 = 0.0;
        for (int64_t i = 0; i < n; i++) {
            double h2i = h2f(h[i]);
            double d2 = orc_distance2(x[j], y[j], x[i], y[i]);
            if (d2 < ORC_CUT * h2i) {
                float g = orc_kernel(h2i, d2) * w[i] * ew[i];
                d32 += g;
                if (i != j) d64 += orc_kernel64(h2i, d2) * (double)w[i] * (double)ew[i];
            }
        }
        if (den1) den1[j] = d32;
        if (den1_64) den1_64[j] = d64;
        if (den0) den0[j] = d32 + orc_kernel(h2j, 0.0) * w[j] * ew[j];
        if (den0_64) den0_64[j] = d64 + orc_kernel64(h2j, 0.0) * (double)w[j] * (double)ew[j];
    }
}

/* logL = sum_j logf(max(den1_j, EPSILONDENSITY)) accumulated in float
 * (tiled:3351-3354, 3649-3652). */
float orc_loglik_ref32(int64_t n, const float *den1)
{
    float s = 0.0f;
    for (int64_t j = 0; j < n; j++) {
        float v = den1[j] < ORC_EPS_DEN ? ORC_EPS_DEN : den1[j];
        s += logf(v);
    }
    return s;
}

double orc_loglik_fp64(int64_t n, const double *den1)
{
    double s = 0.0;
    for (int64_t j = 0; j < n; j++) {
        double v = den1[j] < (double)ORC_EPS_DEN ? (double)ORC_EPS_DEN : den1[j];
        s += log(v);
    }
    return s;
}

/* ---------------------------------------------------------------- N8 --- */
/* ComputeFixedDensityAtPoints (tiled:3096-3134): brute force over all j,
 * den = sum_{j != i} K(h*h, d2)*w_j*ew_j in float (j ascending) plus the self
 * term; den0 = den + den_itself.  FP64: the same sum in double, self included. */
void orc_pilot_density(int64_t n, const float *x, const float *y, const float *w, const float *ew, float h,
                       float *den0, double *den0_64)
{
    double hh2 = h2f(h);
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t i = 0; i < n; i++) {
        float den = 0.0f, self = 0.0f;
        double d64 = 0.0;
        for (int64_t j = 0; j < n; j++) {
            double d2 = orc_distance2(x[i], y[i], x[j], y[j]);
            float k = orc_kernel(hh2, d2) * w[j] * ew[j];
            if (j == i) self += k; else den += k;
            if (d2 < ORC_CUT * hh2) d64 += orc_kernel64(hh2, d2) * (double)w[j] * (double)ew[j];
        }
        if (den0) den0[i] = den + self;
        if (den0_64) den0_64[i] = d64;
    }
}

/* The float loop of ComputeFixedDensityAtPoints (tiled:3096-3134) for a subset
 * of query points: den0[q] for query i = q_idx[q], brute force over all n
 * sources in ascending order, as orc_pilot_density computes it for every
 * point (the CPU baseline times it on a sample of queries). */
void orc_pilot_density_queries(int64_t n, const float *x, const float *y, const float *w, const float *ew, float h,
                               int64_t nq, const int64_t *q_idx, float *den0)
{
    double hh2 = h2f(h);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t q = 0; q < nq; q++) {
        const int64_t i = q_idx[q];
        float den = 0.0f, self = 0.0f;
        for (int64_t j = 0; j < n; j++) {
            double d2 = orc_distance2(x[i], y[i], x[j], y[j]);
            float k = orc_kernel(hh2, d2) * w[j] * ew[j];
            if (j == i) self += k; else den += k;
        }
        den0[q] = den + self;
    }
}

/* compGML (tiled:3807-3814): float sum of log(den0_i) (the float overload),
 * then expf(sum / n). */
float orc_compgml(int64_t n, const float *den0)
{
    float g = 0.0f;
    for (int64_t i = 0; i < n; i++) g = g + logf(den0[i]);
    return expf(g / (float)n);
}

double orc_compgml64(int64_t n, const double *den0)
{
    double g = 0.0;
    for (int64_t i = 0; i < n; i++) g += log(den0[i]);
    return exp(g / (double)n);
}

/* Abramson update hs[i] = h0 * powf(den0[i] / gml, alpha) (tiled:571-573,
 * 3606-3608). */
void orc_abramson(int64_t n, float h0, float alpha, float gml, const float *den0, float *h_out)
{
    for (int64_t i = 0; i < n; i++) h_out[i] = h0 * powf(den0[i] / gml, alpha);
}

void orc_abramson64(int64_t n, double h0, double alpha, double gml, const double *den0, double *h_out)
{
    for (int64_t i = 0; i < n; i++) h_out[i] = h0 * pow(den0[i] / gml, alpha);
}

/* --------------------------------------------------------------- N10 --- */
typedef double (*orc_loglik_fixed_fn)(void *user, float h);
typedef double (*orc_loglik_adapt_fn)(void *user, float h, float alpha);

/* MLE_FixedBandWidth (tiled:3044-3093): float state, factor = 1 + sqrt(5),
 * two likelihood calls per iteration, strict '>', at most 30 iterations.
 * trace (optional, 2 doubles per iteration: hB, hC) records the probes. */
float orc_golden_section(float h, orc_loglik_fixed_fn f, void *user, int *n_iter, double *trace)
{
    float hA = h / 20;
    float hD = 2 * h;
    float width = hD - hA;
    float epsilon = width / 100;
    float factor = 1 + sqrtf(5.0f);
    int iteration = 0;
    while (width > epsilon && iteration < 30) {
        float hB = hA + width / factor;
        float hC = hD - width / factor;
        double LB = f(user, hB);
        double LC = f(user, hC);
        if (trace) { trace[2 * iteration] = hB; trace[2 * iteration + 1] = hC; }
        if (LB > LC) hD = hC; else hA = hB;
        width = hD - hA;
        iteration += 1;
    }
    if (n_iter) *n_iter = iteration;
    return (hA + hD) / 2;
}

/* hj_likelihood (tiled:3677-3722): the five probes around (h0, alpha0), some
 * skipped depending on the previous move. */
static void orc_hj_probe(orc_loglik_adapt_fn f, void *user, float h0, float a0, float sH, float sA, int last,
                         float *Ls)
{
    if (last == -1) Ls[0] = (float)f(user, h0, a0);
    if (last != 2) Ls[1] = (float)f(user, h0 - sH, a0);
    if (last != 1) Ls[2] = (float)f(user, h0 + sH, a0);
    if (last != 4) Ls[3] = (float)f(user, h0, a0 + sA);
    if (last != 3) Ls[4] = (float)f(user, h0, a0 - sA);
}

/* hooke_jeeves (tiled:3726-3790), restated with its bookkeeping as written
 * (dmax 3 and 4 both shift Ls[4] into Ls[0]; Lmax is a running maximum).
 * The likelihood values are kept as float, as the reference stores them.
 * out[0..2] = (h, alpha, Lmax); returns the iteration count. */
int orc_hooke_jeeves(float h0, float alpha0, float stepH, float stepA, orc_loglik_adapt_fn f, void *user,
                     float *out)
{
    float Ls[5] = {0, 0, 0, 0, 0};
    orc_hj_probe(f, user, h0, alpha0, stepH, stepA, -1, Ls);
    float Lmax = Ls[0];
    float s = stepH / 10;
    float a = stepA / 10;
    int iteration = 0;
    while ((stepH > s || stepA > a) && (h0 - stepH > 0) && iteration <= 30) {
        int dmax = 0;
        for (int i = 0; i < 5; i++) {
            if (Ls[i] > Lmax) { Lmax = Ls[i]; dmax = i; }
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
        orc_hj_probe(f, user, h0, alpha0, stepH, stepA, dmax, Ls);
        iteration++;
    }
    out[0] = h0;
    out[1] = alpha0;
    out[2] = Lmax;
    return iteration;
}

/* ----------------------------------------------------- N7/N8, binned --- */
/* The FP64 sums of orc_loo_density / orc_pilot_density for a subset of query
 * points, with the sources found through a uniform grid of bins instead of a
 * scan over all of them (the reference finds them with its k-d tree,
 * KDtree::SearchRange KDtree.cu:471-537, as called from LogLikelihood
 * tiled:3309-3334):
 *   den64[q] = sum over sources i with orc_distance2(x_j, y_j, x_i, y_i) <
 *              9*h2f(h_i) of K64(h2f(h_i), d2) * w_i * ew_i,   j = q_idx[q],
 * excluding i == j when loo != 0 (LOO den1, tiled:3312-3347) and including it
 * otherwise (den0 of the pilot, tiled:3096-3134).  The bins only skip sources
 * whose bin lies farther from the query than 3*max(h)*(1 + 2^-16) + 2^-16
 * (fl32 rounding of a difference moves it by at most one part in 2^24), so the
 * contributing set is exactly the brute-force one; only the order of the FP64
 * additions differs.  *n_pairs (optional) gets the number of contributing
 * (query, source) pairs.  Pinned against orc_loo_density / orc_pilot_density
 * (same pair counts, FP64 sums equal to 1e-12). */
void orc_point_density_binned(int64_t n, const float *x, const float *y, const float *w, const float *ew,
                              const float *h, int loo, int64_t nq, const int64_t *q_idx, double *den64,
                              int64_t *n_pairs)
{
    if (n <= 0 || nq <= 0) {
        if (n_pairs) *n_pairs = 0;
        return;
    }
    double x0 = DBL_MAX, x1 = -DBL_MAX, y0 = DBL_MAX, y1 = -DBL_MAX, hmax = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (x[i] < x0) x0 = x[i];
        if (x[i] > x1) x1 = x[i];
        if (y[i] < y0) y0 = y[i];
        if (y[i] > y1) y1 = y[i];
        if (h[i] > hmax) hmax = h[i];
    }
    const double R = ORC_SQRT_CUT * hmax * (1.0 + 1.0 / 65536.0) + 1.0 / 65536.0;
    double size = R / 3.0;
    const double ext = (x1 - x0 > y1 - y0 ? x1 - x0 : y1 - y0);
    if (size < ext / 4096.0) size = ext / 4096.0;
    if (size <= 0.0) size = 1.0;
    const int nbx = (int)floor((x1 - x0) / size) + 1, nby = (int)floor((y1 - y0) / size) + 1;
    const int64_t nb = (int64_t)nbx * nby;
    int64_t *start = (int64_t *)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t *key = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    int64_t *order = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    double *hh2 = (double *)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        int bx = (int)floor((x[i] - x0) / size), by = (int)floor((y[i] - y0) / size);
        bx = bx < 0 ? 0 : (bx >= nbx ? nbx - 1 : bx);
        by = by < 0 ? 0 : (by >= nby ? nby - 1 : by);
        key[i] = (int64_t)by * nbx + bx;
        start[key[i] + 1]++;
        hh2[i] = h2f(h[i]);
    }
    for (int64_t b = 0; b < nb; b++) start[b + 1] += start[b];
    {
        int64_t *pos = (int64_t *)malloc((size_t)nb * sizeof(int64_t));
        memcpy(pos, start, (size_t)nb * sizeof(int64_t));
        for (int64_t i = 0; i < n; i++) order[pos[key[i]]++] = i;
        free(pos);
    }
    const int k = (int)floor(R / size) + 1;
    int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int64_t q = 0; q < nq; q++) {
        const int64_t j = q_idx[q];
        int bx = (int)floor((x[j] - x0) / size), by = (int)floor((y[j] - y0) / size);
        bx = bx < 0 ? 0 : (bx >= nbx ? nbx - 1 : bx);
        by = by < 0 ? 0 : (by >= nby ? nby - 1 : by);
        double d64 = 0.0;
        int64_t cnt = 0;
        for (int cy = by - k; cy <= by + k; cy++) {
            if (cy < 0 || cy >= nby) continue;
            for (int cx = bx - k; cx <= bx + k; cx++) {
                if (cx < 0 || cx >= nbx) continue;
                const int64_t b = (int64_t)cy * nbx + cx;
                for (int64_t t = start[b]; t < start[b + 1]; t++) {
                    const int64_t i = order[t];
                    if (loo && i == j) continue;
                    const double d2 = orc_distance2(x[j], y[j], x[i], y[i]);
                    if (d2 < ORC_CUT * hh2[i]) {
                        d64 += orc_kernel64(hh2[i], d2) * (double)w[i] * (double)ew[i];
                        cnt++;
                    }
                }
            }
        }
        den64[q] = d64;
        total += cnt;
    }
    free(start);
    free(key);
    free(order);
    free(hh2);
    if (n_pairs) *n_pairs = total;
}

/* --------------------------------------------------------------- N11 --- */
/* ExtractSamplePointsCurTile (tiled:4672-4723): keep point i iff its squared
 * distance to the tile bbox [xll, xll+cs*ncols] x [yll, yll+cs*nrows] is
 * < 9*(double)h*(double)h.  Branch structure as written (edges on the else
 * branch).  Returns the count; index_out (optional) gets ascending indexes. */
int64_t orc_tile_point_filter(int64_t n, const float *x, const float *y, const float *h, const orc_grid *t,
                              int32_t *index_out)
{
    float xll = t->xll, yll = t->yll;
    float xur = xll + t->cs * (float)t->ncols;
    float yur = yll + t->cs * (float)t->nrows;
    int64_t k = 0;
    for (int64_t i = 0; i < n; i++) {
        float px = x[i], py = y[i];
        double ph2 = (double)h[i] * (double)h[i];
        double d = 0;
        if (xll <= px && px <= xur && yll <= py && py <= yur) d = 0;
        else if (px < xll && py > yur) d = orc_distance2(px, py, xll, yur);
        else if (xll < px && px < xur && py > yur) { double df = (double)(float)(py - yur); d = df * df; }
        else if (px > xur && py > yur) d = orc_distance2(px, py, xur, yur);
        else if (yll < py && py < yur && px > xur) { double df = (double)(float)(px - xur); d = df * df; }
        else if (xur < px && py < yll) d = orc_distance2(px, py, xur, yll);
        else if (xll < px && px < xur && py < yll) { double df = (double)(float)(py - yll); d = df * df; }
        else if (px < xll && py < yll) d = orc_distance2(px, py, xll, yll);
        else { double df = (double)(float)(px - xll); d = df * df; }
        if (d < ORC_CUT * ph2) {
            if (index_out) index_out[k] = (int32_t)i;
            k++;
        }
    }
    return k;
}

/* --------------------------------------------------------------- N13 --- */
/* Mode-0 points: srand(100) then x = rand()*100.0f/RAND_MAX, y likewise
 * (tiled:407, 1611-1618), with the MSVC CRT generator the reference was built
 * against (state = state*214013 + 2531011, output (state >> 16) & 0x7fff,
 * RAND_MAX = 32767; the published Microsoft C runtime algorithm). */
void orc_msvc_rand_points(int64_t n, uint32_t seed, float *x, float *y, int32_t *raw)
{
    uint32_t s = seed;
    for (int64_t i = 0; i < n; i++) {
        s = s * 214013u + 2531011u;
        int r1 = (int)((s >> 16) & 0x7fff);
        s = s * 214013u + 2531011u;
        int r2 = (int)((s >> 16) & 0x7fff);
        if (raw) { raw[2 * i] = r1; raw[2 * i + 1] = r2; }
        if (x) x[i] = (float)r1 * 100.0f / (float)32767;
        if (y) y[i] = (float)r2 * 100.0f / (float)32767;
    }
}

/* -------------------------------------------------------------- misc --- */
int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
