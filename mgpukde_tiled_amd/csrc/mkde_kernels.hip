// mkde_kernels.hip -- hand-written HIP kernels for gfx950 (CDNA4, wave64).
//
// Hot path of the reference (Guiming/mGPUKDE_tiled) re-designed for MI355X:
//   k_density_tile    per-cell gather density over 64x64-cell blocks
//                     (replaces KernelDesityEstimation_pPoints,
//                     kde_kernel_kdtr.cu:461-574, and the CPU loop tiled:707-733)
//   k_ec              per-point edge-correction integral, O(R) per point
//                     (replaces CalcEdgeCorrectionWeights kde_kernel_kdtr.cu:117-265
//                     and EdgeCorrectionWeightsExact tiled:2694-2807)
//   k_point_density_w density at points over the cell-binned index, pilot and
//                     leave-one-out forms (replaces DensityAtPointsKdtr
//                     kde_kernel_kdtr.cu:681-953 and the CPU scatter tiled:3312-3347)
//   k_loo_multi_w     leave-one-out densities of up to 16 fixed bandwidths in one
//                     pass over point pairs (batched LogLikelihood, tiled:3214-3368)
//   k_boundary, k_dist_boundary   MarkBoundary / CalcDist2Boundary
//                     (kde_kernel_kdtr.cu:1209-1353, tiled:3993-4131)
//   reductions        FP64 sums of logs (replaces ReductionSum kde_kernel_kdtr.cu:1045-1111)
//
// Numerics shared by every kernel (the reference CPU path's rules):
//   cell centre  cx = fl32(xll + (c+0.5)*cs), cy = fl32(yll + (nrows-0.5-r)*cs)
//                evaluated in double (Utilities.h:13-14)
//   cut-off      fl64(dx^2 + dy^2) < 9*h2f with dx = fl32(cx - px), h2f = fl32(h*h)
//                -- evaluated exactly (both squares are exact in double)
//   kernel       K = exp(-d^2/(2 h2f)) / (2*pi*h2f), computed as
//                exp2(fl32(d^2 * a)) with a = -log2(e)/(2 h2f); the density and
//                edge-correction integrands use the separable form
//                exp2(dx^2 a) * exp2(dy^2 a) over raster cells.
//   sums         short FP32 partial sums folded with compensation (Kahan) or
//                into FP64; FP64 prefix sums (edge correction) and FP64
//                log-likelihood sums.
#include <hip/hip_runtime.h>
#include <float.h>
#include <limits.h>
#include <algorithm>
#include <mutex>
#include <stdint.h>
#include <stdlib.h>

#include "mkde_device.h"

#define WAVE 64

// ------------------------------------------------------------------ helpers --

__device__ __forceinline__ float cell_x(double xll, double cs, int64_t c)
{
    return (float)__dadd_rn(xll, __dmul_rn((double)c + 0.5, cs));
}

// nrm05 = (double)nrows - 0.5 (ROW_TO_YCOORD evaluates (NROWS - 0.5 - ROW)).
__device__ __forceinline__ float cell_y(double yll, double cs, double nrm05, int64_t r)
{
    return (float)__dadd_rn(yll, __dmul_rn(nrm05 - (double)r, cs));
}

// fl64(a^2 + b^2) for float a, b: both squares are exact in double.
__device__ __forceinline__ double sq2(float a, float b)
{
    double da = a, db = b;
    return __dadd_rn(__dmul_rn(da, da), __dmul_rn(db, db));
}

__device__ __forceinline__ double sqd(float a)
{
    double da = a;
    return __dmul_rn(da, da);
}

__device__ __forceinline__ float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }

// Distance from v to the interval [lo, hi] along one axis as fl32(lo - v) or
// fl32(v - hi); 0 inside.  Because fl32 subtraction is monotone, for every u in
// [lo, hi] |fl32(u - v)| >= this value: it is an exact lower bound of the
// reference's |dx| over the interval.
__device__ __forceinline__ float gap1(float v, float lo, float hi)
{
    return v < lo ? (lo - v) : (v > hi ? (v - hi) : 0.0f);
}

// Largest |fl32(u - v)| over u in [lo, hi]: attained at an end point.
__device__ __forceinline__ float far1(float v, float lo, float hi)
{
    float a = fabsf(lo - v), b = fabsf(hi - v);
    return a > b ? a : b;
}

// Launch-wide counters spread over MKDE_CNT_SLOTS cache lines: atomics of
// many workgroups to one address serialise (measured: 400k one-per-workgroup
// adds to a single counter cost ~2.6 ms), so workgroup b adds to slot
// b % MKDE_CNT_SLOTS and k_fold_slots sums the slots into the word after them.
#define MKDE_CNT_SLOTS 512
#define MKDE_CNT_STRIDE 16   // 64-bit words between slots (128 B)
__device__ __forceinline__ void slot_add(unsigned long long *slots, unsigned long long v)
{
    if (v) atomicAdd(&slots[(blockIdx.x % MKDE_CNT_SLOTS) * MKDE_CNT_STRIDE], v);
}

// Kahan fold of a short FP32 partial sum into (sum, comp).
__device__ __forceinline__ void kahan_fold(float part, float &sum, float &comp)
{
    float y = part - comp;
    float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// Order-preserving compaction across a workgroup of NW waves.  Returns this
// thread's slot among the flagged threads; *total gets the flagged count.
template <int NW>
__device__ __forceinline__ int block_compact(bool flag, int *s_wcnt, int *total)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    uint64_t m = __ballot(flag);
    int rank = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        off += (w < wave) ? s_wcnt[w] : 0;
        tot += s_wcnt[w];
    }
    *total = tot;
    __syncthreads();
    return off + rank;
}
__device__ __forceinline__ int block_compact256(bool flag, int *s_wcnt, int *total)
{
    return block_compact<4>(flag, s_wcnt, total);
}

__device__ __forceinline__ int clampi(int64_t v, int64_t lo, int64_t hi)
{
    return (int)(v < lo ? lo : (v > hi ? hi : v));
}

// Candidate-bin range of a query rectangle [qx0,qx1]x[qy0,qy1] grown by reach.
__device__ __forceinline__ void bin_range(const BinGrid &bg, double qx0, double qx1, double qy0, double qy1,
                                          double reach, int &bxa, int &bxb, int &bya, int &byb)
{
    bxa = clampi((int64_t)floor((qx0 - reach - bg.x0) / bg.size), 0, bg.nbx - 1);
    bxb = clampi((int64_t)floor((qx1 + reach - bg.x0) / bg.size), 0, bg.nbx - 1);
    bya = clampi((int64_t)floor((qy0 - reach - bg.y0) / bg.size), 0, bg.nby - 1);
    byb = clampi((int64_t)floor((qy1 + reach - bg.y0) / bg.size), 0, bg.nby - 1);
}

// Exact lower bound of the cut-off distance between any point of bin box B and
// any position in rectangle Q (both in float coordinates).
__device__ __forceinline__ double box_gap2(const BinInfo &B, float qx0, float qx1, float qy0, float qy1)
{
    float gx = B.x1 < qx0 ? (qx0 - B.x1) : (B.x0 > qx1 ? (B.x0 - qx1) : 0.0f);
    float gy = B.y1 < qy0 ? (qy0 - B.y1) : (B.y0 > qy1 ? (B.y0 - qy1) : 0.0f);
    return sq2(gx, gy);
}

// ============================================================ point index ====

// Per-job point records in bin order: gather h and ew through the permutation
// and derive the kernel constants.
__global__ void k_prep_points(int64_t n, const int *perm, const float *xs, const float *ys, const float *w_orig,
                              const float *h_orig, const float *ew_orig, float h_fixed, PtRec *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int o = perm[i];
    float h = h_fixed > 0.0f ? h_fixed : (h_orig ? h_orig[o] : 0.0f);
    float w = w_orig[o];
    float ew = ew_orig[o];
    double h2 = (double)(float)(h * h);
    PtRec r;
    r.x = xs[i];
    r.y = ys[i];
    r.pad = 0.0f;
    if (h2 > 0.0 && h2 < 1e300) {
        r.coef = (float)((double)w * (double)ew / (h2 * (double)MKDE_TWO_PI_F));
        r.a = -MKDE_LOG2E / (2.0 * h2);
        r.thr = 9.0 * h2;
    } else {
        r.coef = 0.0f;
        r.a = 0.0;
        r.thr = -1.0;
    }
    out[i] = r;
}

// Per-bin bounding box of the points and largest threshold.
// One wave per bin (bins in dense clusters hold thousands of points):
// lanes stride over the bin's points, then wave min / max reductions.
__global__ __launch_bounds__(256) void k_bin_stats(int nbins, const int *bin_start, const PtRec *pts, BinInfo *box,
                                                   double *thr)
{
    const int lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (b >= nbins) return;
    const int s = bin_start[b], e = bin_start[b + 1];
    float x0 = FLT_MAX, y0 = FLT_MAX, x1 = -FLT_MAX, y1 = -FLT_MAX;
    double t = -1.0;
    for (int i = s + lane; i < e; i += 64) {
        const PtRec p = pts[i];
        x0 = fminf(x0, p.x);
        x1 = fmaxf(x1, p.x);
        y0 = fminf(y0, p.y);
        y1 = fmaxf(y1, p.y);
        t = fmax(t, p.thr);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        x0 = fminf(x0, __shfl_xor(x0, d, WAVE));
        x1 = fmaxf(x1, __shfl_xor(x1, d, WAVE));
        y0 = fminf(y0, __shfl_xor(y0, d, WAVE));
        y1 = fmaxf(y1, __shfl_xor(y1, d, WAVE));
        t = fmax(t, __shfl_xor(t, d, WAVE));
    }
    if (lane == 0) {
        box[b] = BinInfo{x0, y0, x1, y1};
        thr[b] = t;
    }
}

// Wait states between the matrix cores and the inline-assembly FMAs on the
// same accumulators.  The compiler inserts them between its own instructions
// but not around inline assembly, so both directions are pinned here: the
// asm statements name the accumulators as operands, which keeps them between
// the producer and the consumer.  A VALU write of a VGPR needs 2 wait states
// before an MFMA reads it as SrcC (5 here); a 16-pass MFMA result needs 18
// before a VALU instruction reads it (24 here).
typedef float f32x16 __attribute__((ext_vector_type(16)));
#define MKDE_ACC16(a)                                                                                                 \
    "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]), "+v"(a[8]),     \
        "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
__device__ __forceinline__ void wait_valu_to_mfma(float (&a)[16]) { asm volatile("s_nop 4" : MKDE_ACC16(a)); }
__device__ __forceinline__ void wait_mfma_to_valu(float (&a)[16]) { asm volatile("s_nop 15\ns_nop 7" : MKDE_ACC16(a)); }

// Reach of one raster row for one point: the smallest float f >= 0 with
// fl64(f^2 + dy2) >= thr.  f^2 is exact in FP64 and the sum is monotone in f,
// so a cell of the row is inside the cut-off fl64(dx2 + dy2) < thr exactly
// when |fl32(cx - px)| < f; f = 0 when no cell of the row is inside.  The
// hardware FP64 square root (within about one FP32 ulp) only seeds the search.
// The answer is almost always the seed or the float after it, which two
// evaluations of the exact predicate settle without a loop: inside(seed)
// says which of the two it can be, and the predicate at the neighbour on the
// other side confirms it.  Anything else walks by single ulps.  v^2 is exact
// in FP64 for a float v, so fma(v, v, dy2) rounds once, like fl64(v^2 + dy2).
__device__ __forceinline__ float row_reach(double dy2, double thr)
{
    auto inside = [&](float v) { return __fma_rn((double)v, (double)v, dy2) < thr; };
    auto up = [](float v) { return __uint_as_float(__float_as_uint(v) + 1u); };
    auto down = [](float v) { return __uint_as_float(__float_as_uint(v) - 1u); };
    const bool any = dy2 < thr;
    float f = (float)__builtin_amdgcn_sqrt(fmax(thr - dy2, 0.0));
    // c0: the answer is above f; then f' = up(f) must be outside, else
    // f' = f must have down(f) inside (f > 0: inside(0) holds when any)
    const bool c0 = inside(f);
    const float t = c0 ? up(f) : down(f);
    const bool ok = f > 0.0f && (inside(t) != c0);
    if (__builtin_expect(!any || ok, 1)) return any ? (c0 ? t : f) : 0.0f;
    for (int k = 0; k < 4096 && inside(f); ++k) f = up(f);
    for (int k = 0; k < 4096 && f > 0.0f && !inside(down(f)); ++k) f = down(f);
    return f;
}

// Accumulators R..R+3 of a PARTIAL sub-block: m_i = |dx| < reach_i ? ex : 0
// (the exact cut-off, row_reach), then acc[R+i] += Ey[R+i] * m_i.  Ey of a
// wave's 16 accumulator rows lives in one VGPR (lane l holds the row of
// accumulator l & 15, the same 16 rows in every 16-lane row of the wave), and
// DPP row_newbcast:j hands accumulator j's value to the lanes as the FMA's
// first operand instead of a 64-lane LDS broadcast read per value.  The four
// lane masks go to four SGPR pairs and the four m_i are formed before the
// four DPP FMAs, so each m_i is at least 2 VALU instructions old when it is
// read (the wait states a DPP read needs after a VALU write).
#define MKDE_XCMP(I) "v_cmp_gt_f32_e64 %[s" #I "], %[x" #I "], |%[dx]|\n"
#define MKDE_XSEL(I) "v_cndmask_b32_e64 %[m" #I "], 0, %[ex], %[s" #I "]\n"
#define MKDE_XFMA(I, J) "v_fmac_f32_dpp %[a" #I "], %[ey], %[m" #I "] row_newbcast:" #J " row_mask:0xf bank_mask:0xf\n"
#define MKDE_PART4(R0, R1, R2, R3)                                                                                    \
    __device__ __forceinline__ void part4_##R0(float &a0, float &a1, float &a2, float &a3, float4 x, float ex,         \
                                               float eyv, float dx)                                                   \
    {                                                                                                                 \
        float m0, m1, m2, m3;                                                                                         \
        uint64_t s0, s1, s2, s3;                                                                                      \
        asm(MKDE_XCMP(0) MKDE_XCMP(1) MKDE_XCMP(2) MKDE_XCMP(3) MKDE_XSEL(0) MKDE_XSEL(1) MKDE_XSEL(2) MKDE_XSEL(3)   \
                MKDE_XFMA(0, R0) MKDE_XFMA(1, R1) MKDE_XFMA(2, R2) MKDE_XFMA(3, R3)                                   \
            : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [m0] "=&v"(m0), [m1] "=&v"(m1),             \
              [m2] "=&v"(m2), [m3] "=&v"(m3), [s0] "=&s"(s0), [s1] "=&s"(s1), [s2] "=&s"(s2), [s3] "=&s"(s3)          \
            : [x0] "v"(x.x), [x1] "v"(x.y), [x2] "v"(x.z), [x3] "v"(x.w), [ex] "v"(ex), [ey] "v"(eyv), [dx] "v"(dx));   \
    }
MKDE_PART4(0, 1, 2, 3)
MKDE_PART4(4, 5, 6, 7)
MKDE_PART4(8, 9, 10, 11)
MKDE_PART4(12, 13, 14, 15)

// ======================================================== density raster ====
//
// One workgroup computes a block of 64 x MKDE_DT_BW cells of a tile (64 x 64,
// 256 threads).  Each wave owns a 32x32 quadrant (lane layout below), each
// lane one column of it and 16 of its rows (16 accumulators).
// Blocks without a valid cell only copy the mask's nodata values.
// The workgroup walks the candidate bins, classifies every point of those
// bins against the block exactly (EMPTY: no cell inside the cut-off; FULL:
// every cell inside; PARTIAL), queues the non-empty ones in bin order, and
// consumes them in sub-batches of DT_SB points:
//   1. cooperatively, in LDS: Ex[p][col] = exp2(dx2 a_p) for the 64 columns,
//      Ey[p][row] = coef_p exp2(dy2 a_p) for the 64 rows, and for PARTIAL
//      points the exact reach of each row (row_reach) and the class of each
//      wave's sub-block (the same exact EMPTY / FULL test on the
//      smaller rectangle);
//   2. each wave adds Ex * Ey[p][row] to its 16 cells with one FMA each
//      (FULL), or with Ex replaced by 0 on the lanes with |dx| >= reach[p][row]
//      (PARTIAL: the exact cut-off fl64(dx2 + dy2) < thr), and skips EMPTY
//      sub-blocks;
//   3. the FP32 partial sums are added to FP64 cell sums after every DT_FOLD
//      sub-batches.
// The order of points is fixed (bins in row-major order, points in bin order)
// so results are deterministic.

#define DT_SB 32
#define DT_EYP 68   // floats per point row of the Ey and reach tables (64 + pad)
#define DT_Q (DT_SB + DT_NT)
// Sub-batches summed in FP32 before the fold into FP64: at most 64 positive
// terms, a relative error below 64 * 2^-24 = 3.8e-6 per fold
#define DT_FOLD 2

// stats[0] candidate points classified, [1] FULL blocks, [2] PARTIAL blocks,
// [3] FULL / [4] PARTIAL / [5] EMPTY wave sub-blocks of PARTIAL points,
// [6] blocks skipped (no valid cell)
#define DT_NSTATS 8

// Lane -> cells of a wave: the block's 64 x MKDE_DT_BW cells are DT_QX x 2
// quadrants of 32 x 32, wave w covering the one at column 32 (w % DT_QX), row
// 32 (w / DT_QX); lane l holds column 32 (w % DT_QX) + (l & 31) and, in
// accumulator j < 16, row 32 (w / DT_QX) + 8 (j >> 2) + 4 (l >> 5) + (j & 3):
// the accumulator layout of v_mfma_f32_32x32x2_f32.  Accumulators 4g .. 4g+3
// of a lane are four consecutive rows, and the lanes of one 16-lane DPP row
// all hold the same 16 rows.
#define DT_QX (MKDE_DT_BW / 32)
#define DT_NW (2 * DT_QX)   // waves per workgroup
#define DT_NT (64 * DT_NW)  // threads per workgroup
#define DT_WX0(w) (32 * ((w) % DT_QX))
#define DT_WY0(w) (32 * ((w) / DT_QX))
#define DT_WC(w, l) (DT_WX0(w) + ((l) & 31))
#define DT_ROW(w, l, j) (DT_WY0(w) + 8 * ((j) >> 2) + 4 * ((l) >> 5) + ((j) & 3))
#define DT_WNX 32
#define DT_WNY 32

// Raster of one or more tiles.  Single tile: g holds the tile geometry.
// Tiled (tl.tw > 0): g.ncols / g.nrows are the raster size and each tile takes
// its origin from the raster geotransform as AsciiRasterFromGeoTIFFTile does
// for a GeoTIFF tile (tiled:2218-2220, GeoTIFF.cpp:190-195):
//   xll = fl32(gt0 + gt1*xoff), yll = fl32((gt3 + gt5*yoff) + rows*gt5), cs = fl32(gt1).
// A work item is one part of a 64x64 block: {block, part j, parts K, partial
// slot (K > 1) or -1}.  A block whose candidate count exceeds the split
// threshold is cut into K parts by candidate ordinal (the position of a point
// in the block's bin-order candidate sequence): part j takes ordinals
// [C j / K, C (j+1) / K) of the C candidates and writes FP64 partial sums to
// its slot; k_density_merge adds the K partials in part order.  Items are
// visited heaviest first (order[]).
struct TileArgs {
    GridDesc g;
    const float *mask;   // raster values (nullable: all valid)
    float *out;
    PointIndex idx;
    unsigned long long *stats;   // nullable
    const int *order;            // nullable: items in raster order, one per block
    const int4 *items;           // with order: work items
    const unsigned *cost;        // with order: candidate count of each block
    double *partials;            // MKDE_DT_CELLS per slot
    int nbx;                     // density blocks per raster row
    Tiling tl;
};

// Geometry of one 64x64 block: global cell range [c0,c1]x[r0,r1], and the
// origin of the tile it belongs to (block-local cell index = global - off).
struct BlockGeom {
    int64_t c0, r0, c1, r1, xoff, yoff;
    double xll, yll, cs, nrm05;
};

__device__ __forceinline__ BlockGeom block_geom(const GridDesc &g, const Tiling &tl, int nbx, int blk)
{
    BlockGeom b;
    if (tl.tw > 0) {
        const int t = blk / tl.bpt, lb = blk - t * tl.bpt;
        b.xoff = (int64_t)(t % tl.ntx) * tl.tw;
        b.yoff = (int64_t)(t / tl.ntx) * tl.th;
        const int64_t tcols = (tl.tw < g.ncols - b.xoff) ? tl.tw : g.ncols - b.xoff;
        const int64_t trows = (tl.th < g.nrows - b.yoff) ? tl.th : g.nrows - b.yoff;
        b.c0 = b.xoff + (int64_t)(lb % tl.bpx) * MKDE_DT_BW;
        b.r0 = b.yoff + (int64_t)(lb / tl.bpx) * 64;
        b.c1 = (b.c0 + MKDE_DT_BW - 1 < b.xoff + tcols - 1) ? b.c0 + MKDE_DT_BW - 1 : b.xoff + tcols - 1;
        b.r1 = (b.r0 + 63 < b.yoff + trows - 1) ? b.r0 + 63 : b.yoff + trows - 1;
        const double g0 = tl.gt0 + tl.gt1 * (double)b.xoff;
        const double g3 = tl.gt3 + tl.gt5 * (double)b.yoff;
        b.xll = (double)(float)g0;
        b.yll = (double)(float)(g3 + (double)trows * tl.gt5);
        b.cs = (double)(float)tl.gt1;
        b.nrm05 = (double)trows - 0.5;
    } else {
        b.c0 = (int64_t)(blk % nbx) * MKDE_DT_BW;
        b.r0 = (int64_t)(blk / nbx) * 64;
        b.c1 = (b.c0 + MKDE_DT_BW - 1 < g.ncols - 1) ? b.c0 + MKDE_DT_BW - 1 : g.ncols - 1;
        b.r1 = (b.r0 + 63 < g.nrows - 1) ? b.r0 + 63 : g.nrows - 1;
        b.xoff = b.yoff = 0;
        b.xll = g.xll;
        b.yll = g.yll;
        b.cs = g.cs;
        b.nrm05 = (double)g.nrows - 0.5;
    }
    return b;
}

// 4 waves per SIMD: at most 128 VGPRs (no spills).  64 x 64 blocks: 37.7 KB
// of LDS per workgroup, 4 workgroups per CU, so each SIMD's 4 waves belong to
// 4 workgroups and a barrier wait in one leaves 3 to issue.  (64 x 128
// blocks share Ey and the reach over twice the columns, but at 55 KB and 8
// waves only 2 workgroups fit per CU: measured 31.6 ms against 24.3 ms on
// the adaptive 1M job.)
template <bool STATS>
__global__ __launch_bounds__(DT_NT) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_density_tile(TileArgs A)
{
    __shared__ int s_bins[DT_NT];
    __shared__ int s_wcnt[DT_NW];
    __shared__ unsigned s_queue[DT_Q];   // point index | 0x80000000 for block-PARTIAL
    // records of the queued points, stored when they are classified, so a
    // sub-batch reads its parameters from LDS instead of global memory
    __shared__ float s_qx[DT_Q], s_qy[DT_Q], s_qc[DT_Q];
    __shared__ double s_qa[DT_Q], s_qt[DT_Q];
    __shared__ float s_px[DT_SB];
    __shared__ float s_py[DT_SB];
    __shared__ float s_coef[DT_SB];
    __shared__ double s_a[DT_SB];
    __shared__ double s_thr[DT_SB];
    __shared__ int s_wcls[DT_SB];   // 2 bits per wave: 0 skip, 1 full, 2 partial
    __shared__ int s_nfull;         // block-FULL points at the front of the sub-batch
    __shared__ __attribute__((aligned(16))) float s_ex[DT_SB * MKDE_DT_BW];
    __shared__ __attribute__((aligned(16))) float s_ey[DT_SB * DT_EYP];
    __shared__ __attribute__((aligned(16))) float s_xr[DT_SB * DT_EYP];   // row reach, PARTIAL points
    __shared__ float s_ccx[MKDE_DT_BW], s_ccy[64];   // cell-centre x of the block's columns, y of its rows
    __shared__ float s_wyb[2 * DT_NW];               // y of the first / last row of each wave's sub-block
    __shared__ float s_wxb[2 * DT_NW];               // x of the first / last column of each wave's sub-block

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GridDesc g = A.g;
    int4 item = make_int4((int)blockIdx.x, 0, 1, -1);
    if (A.order) item = A.items[A.order[blockIdx.x]];
    const int blk = item.x;
    // candidate-ordinal window of this part
    int64_t win0 = 0, win1 = INT64_MAX;
    if (item.z > 1) {
        // the first and last parts are open-ended, so the parts always cover
        // the whole candidate sequence
        const int64_t C = A.cost[blk];
        if (item.y > 0) win0 = C * item.y / item.z;
        if (item.y < item.z - 1) win1 = C * (item.y + 1) / item.z;
    }
    const BlockGeom B = block_geom(g, A.tl, A.nbx, blk);
    const double xll = B.xll, yll = B.yll, cs = B.cs, nrm05 = B.nrm05;
    const int64_t c0 = B.c0, r0 = B.r0, c1 = B.c1, r1 = B.r1;
    // tile-local indices for cell centres
    const int64_t lc0 = c0 - B.xoff, lr0 = r0 - B.yoff, lc1 = c1 - B.xoff, lr1 = r1 - B.yoff;
    const int wc = DT_WC(wave, lane);   // this lane's column
    const int64_t col = c0 + wc;

    // validity of this thread's cells; blocks without a valid cell are skipped
    unsigned vbits = 0;
    if (col <= c1) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int64_t row = r0 + DT_ROW(wave, lane, j);
            if (row <= r1) {
                bool v = true;
                if (A.mask) v = A.mask[row * g.ncols + col] != g.nodata;
                vbits |= (v ? 1u : 0u) << j;
            }
        }
    }
    if (!__syncthreads_or(vbits != 0)) {
        if (col <= c1) {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int64_t row = r0 + DT_ROW(wave, lane, j);
                if (row <= r1) A.out[row * g.ncols + col] = A.mask[row * g.ncols + col];
            }
        }
        if (STATS && tid == 0) atomicAdd(&A.stats[6], 1ull);
        return;
    }

    // Block extent in cell-centre coordinates (exact N2 floats).
    const float bx0 = cell_x(xll, cs, lc0), bx1 = cell_x(xll, cs, lc1);
    const float by1 = cell_y(yll, cs, nrm05, lr0), by0 = cell_y(yll, cs, nrm05, lr1);
    const float cx = cell_x(xll, cs, lc0 + wc);
    {
        // thread ranges (even starts): columns, rows, then per wave [0] first /
        // [1] last row and [0] first / [1] last column of its quadrant
        constexpr int T1 = MKDE_DT_BW, T2 = T1 + 64, T3 = T2 + 2 * DT_NW, T4 = T3 + 2 * DT_NW;
        static_assert(T4 <= DT_NT, "too few threads for the block tables");
        if (tid < T1) s_ccx[tid] = cell_x(xll, cs, lc0 + tid);
        else if (tid < T2) s_ccy[tid - T1] = cell_y(yll, cs, nrm05, lr0 + (tid - T1));
        else if (tid < T3) {
            const int w = (tid - T2) >> 1;
            const int64_t wr0 = lr0 + DT_WY0(w);
            const int64_t wr1 = (wr0 + DT_WNY - 1 < lr1) ? wr0 + DT_WNY - 1 : lr1;
            s_wyb[tid - T2] = (tid & 1) ? cell_y(yll, cs, nrm05, wr1) : cell_y(yll, cs, nrm05, wr0);
        } else if (tid < T4) {
            const int w = (tid - T3) >> 1;
            const int64_t wc0 = lc0 + DT_WX0(w);
            const int64_t wc1 = (wc0 + DT_WNX - 1 < lc1) ? wc0 + DT_WNX - 1 : lc1;
            s_wxb[tid - T3] = (tid & 1) ? cell_x(xll, cs, wc1) : cell_x(xll, cs, wc0);
        }
    }

    // FP32 partial sums (at most DT_FOLD sub-batches of terms) folded into
    // FP64 cell sums
    float acc[16];
    double sum[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { acc[j] = 0.0f; sum[j] = 0.0; }
    unsigned st_cand = 0, st_full = 0, st_part = 0, st_wf = 0, st_wp = 0, st_we = 0;

    const PointIndex &I = A.idx;
    int bxa, bxb, bya, byb;
    bin_range(I.bg, bx0, bx1, by0, by1, *I.reach_p, bxa, bxb, bya, byb);
    const int nbx_r = bxb - bxa + 1;
    const int nb_total = nbx_r * (byb - bya + 1);

    int qn = 0;       // queue length (uniform)
    int npend = 0;    // sub-batches in acc since the last fold (uniform)
    auto fold = [&]() {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            sum[j] += (double)acc[j];
            acc[j] = 0.0f;
        }
        npend = 0;
    };
    int64_t ord = 0;  // candidate ordinal at the start of the next candidate bin (uniform)

    auto process = [&](int cnt, int head) {
        // 1a. point parameters of queue entries [head, head + cnt), block-FULL
        //     points first (stable partition, so the order stays
        //     deterministic); cnt <= DT_SB <= 64: one wave
        if (tid < 64) {
            const int e = head + tid;
            const unsigned q = tid < cnt ? s_queue[e] : 0u;
            const bool isfull = tid < cnt && !(q >> 31);
            const uint64_t fm = __ballot(isfull);
            const uint64_t below = (1ull << tid) - 1ull;
            const int nf = __popcll(fm);
            const int slot = isfull ? __popcll(fm & below) : nf + (tid - __popcll(fm & below));
            if (tid == 0) s_nfull = nf;
            if (tid < cnt) {
                const float px = s_qx[e], py = s_qy[e];
                const double thr = s_qt[e];
                const int t = slot;
                s_px[t] = px;
                s_py[t] = py;
                s_coef[t] = s_qc[e];
                s_a[t] = s_qa[e];
                s_thr[t] = thr;
                // wave sub-block classes
                // 2 bits per wave: 0 skip, 1 full, 2 partial
                int cl4 = (int)(0x55555555u >> (32 - 2 * DT_NW));
                if (q >> 31) {
                    cl4 = 0;
#pragma unroll
                    for (int w = 0; w < DT_NW; ++w) {
                        int cw = 0;
                        if (lr0 + DT_WY0(w) <= lr1 && lc0 + DT_WX0(w) <= lc1) {
                            const float wy1 = s_wyb[2 * w], wy0 = s_wyb[2 * w + 1];
                            const float wx0 = s_wxb[2 * w], wx1 = s_wxb[2 * w + 1];
                            const float fx = far1(px, wx0, wx1), nx = gap1(px, wx0, wx1);
                            if (sq2(nx, gap1(py, wy0, wy1)) < thr) cw = sq2(fx, far1(py, wy0, wy1)) < thr ? 1 : 2;
                        }
                        cl4 |= cw << (2 * w);
                    }
                }
                s_wcls[t] = cl4;
            }
        }
        __syncthreads();
        // 1b. tables: lane -> row / column slot, wave -> DT_SB / DT_NW points;
        //     consecutive lanes write consecutive LDS words (no bank
        //     conflicts).  Ey and the reach of a row serve every column of
        //     the block, so a wider block shares them.
        const int nfull = s_nfull;
        {
            const float cyr = s_ccy[lane];
            float cxr[MKDE_DT_BW / 64];
#pragma unroll
            for (int cb = 0; cb < MKDE_DT_BW / 64; ++cb) cxr[cb] = s_ccx[lane + 64 * cb];
            // the point index is wave-uniform: made explicit (scalar index
            // and bound test, fewer VALU address operations)
            const int pw = __builtin_amdgcn_readfirstlane(wave) * (DT_SB / DT_NW);
#pragma unroll 2
            for (int q = 0; q < DT_SB / DT_NW; ++q) {
                const int p = pw + q;
                if (p < cnt) {
                    const float py = s_py[p], px = s_px[p], coef = s_coef[p];
                    const double a = s_a[p];
                    const double dy2 = sqd(cyr - py);
                    s_ey[p * DT_EYP + lane] = fast_exp2((float)(dy2 * a)) * coef;
#pragma unroll
                    for (int cb = 0; cb < MKDE_DT_BW / 64; ++cb)
                        s_ex[p * MKDE_DT_BW + lane + 64 * cb] = fast_exp2((float)(sqd(cxr[cb] - px) * a));
                    if (p >= nfull) s_xr[p * DT_EYP + lane] = row_reach(dy2, s_thr[p]);
                }
            }
        }
        __syncthreads();
        // 2a. this wave's FULL points on the matrix cores: the quadrant's
        //     cells += Ey (32 rows x 2 points) * Ex (2 points x 32 columns),
        //     one v_mfma_f32_32x32x2_f32 per two points, lane l feeding row /
        //     column l & 31 of the pair's point l >> 5 (0 for a missing
        //     second point).  FULL are the block-FULL points [0, nfull) and
        //     the block-PARTIAL points whose class for this wave is FULL;
        //     the classes come as two wave-uniform bit masks over the
        //     sub-batch's PARTIAL points.
        const int np = cnt - nfull;   // <= DT_SB
        const int wcl = lane < np ? (s_wcls[nfull + lane] >> (2 * wave)) & 3 : 0;
        const uint64_t mf = __ballot(wcl == 1);
        uint64_t mp = __ballot(wcl == 2);
        if (STATS && lane == 0) {
            st_wf += nfull + __popcll(mf);
            st_wp += __popcll(mp);
            st_we += np - __popcll(mf) - __popcll(mp);
        }
        if (nfull > 0 || mf) {
            wait_valu_to_mfma(acc);
            f32x16 c;
#pragma unroll
            for (int j = 0; j < 16; ++j) c[j] = acc[j];
            const int k = lane >> 5, li = lane & 31;
            const float *eyp = s_ey + DT_WY0(wave) + li, *exp = s_ex + DT_WX0(wave) + li;
            // block-FULL points eight, then four per step (the later pairs'
            // operands arrive while the first pair's MFMA runs), then the
            // rest in pairs
            int p = 0;
            for (; p + 8 <= nfull; p += 8) {
                const float ey0 = eyp[(p + k) * DT_EYP], ex0 = exp[(p + k) * MKDE_DT_BW];
                const float ey1 = eyp[(p + 2 + k) * DT_EYP], ex1 = exp[(p + 2 + k) * MKDE_DT_BW];
                const float ey2 = eyp[(p + 4 + k) * DT_EYP], ex2 = exp[(p + 4 + k) * MKDE_DT_BW];
                const float ey3 = eyp[(p + 6 + k) * DT_EYP], ex3 = exp[(p + 6 + k) * MKDE_DT_BW];
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey0, ex0, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey1, ex1, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey2, ex2, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey3, ex3, c, 0, 0, 0);
            }
            for (; p + 4 <= nfull; p += 4) {
                const float ey0 = eyp[(p + k) * DT_EYP], ex0 = exp[(p + k) * MKDE_DT_BW];
                const float ey1 = eyp[(p + 2 + k) * DT_EYP], ex1 = exp[(p + 2 + k) * MKDE_DT_BW];
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey0, ex0, c, 0, 0, 0);
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(ey1, ex1, c, 0, 0, 0);
            }
            uint64_t rest = (((1ull << (nfull - p)) - 1ull) << p) | (mf << nfull);
            while (rest) {
                const int p0 = __builtin_ctzll(rest);
                rest &= rest - 1ull;
                const int p1 = rest ? __builtin_ctzll(rest) : -1;
                rest &= rest - 1ull;
                const int q = k ? p1 : p0, qs = q < 0 ? 0 : q;
                const float ey = eyp[qs * DT_EYP], ex = exp[qs * MKDE_DT_BW];
                c = __builtin_amdgcn_mfma_f32_32x32x2f32(q >= 0 ? ey : 0.0f, q >= 0 ? ex : 0.0f, c, 0, 0, 0);
            }
#pragma unroll
            for (int j = 0; j < 16; ++j) acc[j] = c[j];
            wait_mfma_to_valu(acc);
        }
        // 2b. this wave's PARTIAL points: Ey of the wave's rows as one VGPR
        //     (lane l holds the row of accumulator l & 15, broadcast to its
        //     16-lane DPP row), the FMAs masked by the row reach
        const int eyo = DT_ROW(wave, lane, lane & 15);
        while (mp) {
            const int p = nfull + __builtin_ctzll(mp);
            mp &= mp - 1ull;
            const float ex = s_ex[p * MKDE_DT_BW + wc];
            const float eyv = s_ey[p * DT_EYP + eyo];
            const float dx = cx - s_px[p];
            // rows of accumulators 4g .. 4g+3: a float4 every 8 rows
            const float4 *xr = (const float4 *)&s_xr[p * DT_EYP + DT_ROW(wave, lane, 0)];
            part4_0(acc[0], acc[1], acc[2], acc[3], xr[0], ex, eyv, dx);
            part4_4(acc[4], acc[5], acc[6], acc[7], xr[2], ex, eyv, dx);
            part4_8(acc[8], acc[9], acc[10], acc[11], xr[4], ex, eyv, dx);
            part4_12(acc[12], acc[13], acc[14], acc[15], xr[6], ex, eyv, dx);
        }
        if (++npend == DT_FOLD) fold();
        __syncthreads();
    };

    for (int base = 0; base < nb_total; base += DT_NT) {
        // candidate bins of this chunk
        const int k = base + tid;
        bool cand = false;
        int b = 0;
        if (k < nb_total) {
            b = (bya + k / nbx_r) * I.bg.nbx + (bxa + k % nbx_r);
            if (I.bin_start[b + 1] > I.bin_start[b]) {
                double g2 = box_gap2(I.bin_box[b], bx0, bx1, by0, by1);
                cand = g2 < I.bin_thr[b];
            }
        }
        int nb;
        int slot = block_compact<DT_NW>(cand, s_wcnt, &nb);
        if (cand) s_bins[slot] = b;
        __syncthreads();
        for (int kb = 0; kb < nb; ++kb) {
            const int bb = s_bins[kb];
            const int bst = I.bin_start[bb], ben = I.bin_start[bb + 1];
            // clip the bin's points to this part's ordinal window
            const int64_t o0 = ord, o1 = ord + (ben - bst);
            ord = o1;
            if (o1 <= win0 || o0 >= win1) continue;
            const int st = bst + (int)(win0 > o0 ? win0 - o0 : 0);
            const int en = ben - (int)(o1 > win1 ? o1 - win1 : 0);
            for (int off = st; off < en; off += DT_NT) {
                const int i = off + tid;
                int cls = 0;   // 0 empty, 1 full, 2 partial
                PtRec p;
                if (i < en) {
                    p = I.pts[i];
                    if (STATS) ++st_cand;
                    if (p.thr > 0.0) {
                        double near2 = sq2(gap1(p.x, bx0, bx1), gap1(p.y, by0, by1));
                        if (near2 < p.thr) {
                            double far2 = sq2(far1(p.x, bx0, bx1), far1(p.y, by0, by1));
                            cls = far2 < p.thr ? 1 : 2;
                        }
                    }
                    if (STATS && cls == 1) ++st_full;
                    if (STATS && cls == 2) ++st_part;
                }
                int na;
                int pos = block_compact<DT_NW>(cls != 0, s_wcnt, &na);
                if (cls != 0) {
                    const int e = qn + pos;
                    s_queue[e] = (unsigned)i | (cls == 2 ? 0x80000000u : 0u);
                    s_qx[e] = p.x;
                    s_qy[e] = p.y;
                    s_qc[e] = p.coef;
                    s_qa[e] = p.a;
                    s_qt[e] = p.thr;
                }
                __syncthreads();
                qn += na;
                if (qn >= DT_SB) {
                    int head = 0;
                    while (qn - head >= DT_SB) {
                        process(DT_SB, head);
                        head += DT_SB;
                    }
                    // keep the remainder at the front
                    const int rem = qn - head;
                    unsigned v = 0;
                    float vx = 0.0f, vy = 0.0f, vc = 0.0f;
                    double va = 0.0, vt = 0.0;
                    if (tid < rem) {
                        const int e = head + tid;
                        v = s_queue[e];
                        vx = s_qx[e];
                        vy = s_qy[e];
                        vc = s_qc[e];
                        va = s_qa[e];
                        vt = s_qt[e];
                    }
                    __syncthreads();
                    if (tid < rem) {
                        s_queue[tid] = v;
                        s_qx[tid] = vx;
                        s_qy[tid] = vy;
                        s_qc[tid] = vc;
                        s_qa[tid] = va;
                        s_qt[tid] = vt;
                    }
                    __syncthreads();
                    qn = rem;
                }
            }
        }
        __syncthreads();
    }
    if (qn > 0) process(qn, 0);
    if (npend > 0) fold();

    // write back: nodata cells keep the mask value; parts of a split block
    // leave FP64 partial sums for k_density_merge
    if (item.z > 1) {
        double *pp = A.partials + (int64_t)item.w * MKDE_DT_CELLS;
#pragma unroll
        for (int j = 0; j < 16; ++j) pp[DT_ROW(wave, lane, j) * MKDE_DT_BW + wc] = sum[j];
    } else if (col <= c1) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const int64_t row = r0 + DT_ROW(wave, lane, j);
            if (row <= r1) {
                const int64_t ix = row * g.ncols + col;
                A.out[ix] = ((vbits >> j) & 1u) ? (float)sum[j] : A.mask[ix];
            }
        }
    }
    if (STATS) {
        auto put = [&](int slot, unsigned v) {
            unsigned long long t = v;
            for (int d = 32; d >= 1; d >>= 1) t += __shfl_xor(t, d, WAVE);
            if (lane == 0 && t) atomicAdd(&A.stats[slot], t);
        };
        put(0, st_cand);
        put(1, st_full);
        put(2, st_part);
        put(3, st_wf);
        put(4, st_wp);
        put(5, st_we);
    }
}

// Work estimate of each density block: points in the candidate bins (0 for
// blocks without a valid cell), then a heaviest-first visiting order by
// log2 buckets, so the long blocks start early and the launch drains evenly.
__global__ __launch_bounds__(256) void k_block_cost(GridDesc g, const float *mask, PointIndex I, int nbx,
                                                    Tiling tl, unsigned *cost)
{
    __shared__ unsigned s_sum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int blk = blockIdx.x;
    const BlockGeom B = block_geom(g, tl, nbx, blk);
    bool v = false;
    if (mask) {
        // each thread stops at its first valid cell (most blocks are valid
        // from their first row on)
        for (int k = tid; k < MKDE_DT_CELLS && !v; k += 256) {
            const int64_t r = B.r0 + k / MKDE_DT_BW, c = B.c0 + k % MKDE_DT_BW;
            if (r <= B.r1 && c <= B.c1 && mask[r * g.ncols + c] != g.nodata) v = true;
        }
    } else {
        v = B.c0 <= B.c1 && B.r0 <= B.r1;   // edge tiles leave some blocks empty
    }
    const int any = __syncthreads_or(v);
    const float bx0 = cell_x(B.xll, B.cs, B.c0 - B.xoff), bx1 = cell_x(B.xll, B.cs, B.c1 - B.xoff);
    const float by1 = cell_y(B.yll, B.cs, B.nrm05, B.r0 - B.yoff), by0 = cell_y(B.yll, B.cs, B.nrm05, B.r1 - B.yoff);
    unsigned acc = 0;
    if (any) {
        int bxa, bxb, bya, byb;
        bin_range(I.bg, bx0, bx1, by0, by1, *I.reach_p, bxa, bxb, bya, byb);
        const int nbx_r = bxb - bxa + 1;
        const int nb_total = nbx_r * (byb - bya + 1);
        for (int k = tid; k < nb_total; k += 256) {
            const int b = (bya + k / nbx_r) * I.bg.nbx + (bxa + k % nbx_r);
            const int cntb = I.bin_start[b + 1] - I.bin_start[b];
            if (cntb > 0 && box_gap2(I.bin_box[b], bx0, bx1, by0, by1) < I.bin_thr[b]) acc += (unsigned)cntb;
        }
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, WAVE);
    if (lane == 0) s_sum[wave] = acc;
    __syncthreads();
    if (tid == 0) cost[blk] = s_sum[0] + s_sum[1] + s_sum[2] + s_sum[3];
}

__device__ __forceinline__ int cost_bucket(unsigned c)
{
    // 32 buckets, heaviest first: bucket 0 holds the largest costs
    return c == 0 ? 31 : (__clz(c) < 31 ? __clz(c) : 30);
}

__global__ void k_bucket_scan(unsigned *hist)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        unsigned run = 0;
        for (int k = 0; k < 32; ++k) {
            unsigned c = hist[k];
            hist[k] = run;
            run += c;
        }
    }
}

// Sum of the K partials of each split block in part order (FP64), rounded
// once to FP32; nodata cells keep the mask value.  splits[k] = {block, slot0, K}.
__global__ __launch_bounds__(256) void k_density_merge(GridDesc g, Tiling tl, int nbx, const float *mask,
                                                       const int4 *splits, const double *partials, float *out)
{
    const int4 sp = splits[blockIdx.x];
    const BlockGeom B = block_geom(g, tl, nbx, sp.x);
    // 256 threads over the block's 64 x MKDE_DT_BW cells, row-major
    for (int k = threadIdx.x; k < MKDE_DT_CELLS; k += 256) {
        const int r = k / MKDE_DT_BW, cc = k % MKDE_DT_BW;
        const int64_t row = B.r0 + r, col = B.c0 + cc;
        if (row > B.r1 || col > B.c1) continue;
        const int64_t ix = row * g.ncols + col;
        const double *pp = partials + (int64_t)sp.y * MKDE_DT_CELLS + k;
        double t = 0.0;
        for (int q = 0; q < sp.z; ++q) t += pp[(int64_t)q * MKDE_DT_CELLS];
        const bool valid = !mask || mask[ix] != g.nodata;
        out[ix] = valid ? (float)t : mask[ix];
    }
}

// ---- work list of the density kernel
//
// cost[b] (candidates of block b) -> split threshold T = max(DT_TMIN,
// ceil(total / DT_TARGET)), parts[b] = cost > T ? ceil(cost / T) : 1, exclusive
// scans for item and slot offsets, items[], splits[], and a heaviest-first
// order of the items by cost / parts.
#define DT_TMIN 2048u
#define DT_TARGET 4096ull

__global__ void k_cost_total(int n, const unsigned *cost, unsigned long long *total)
{
    __shared__ unsigned long long s_w[4];
    unsigned long long acc = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) acc += cost[i];
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, WAVE);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(total, s_w[0] + s_w[1] + s_w[2] + s_w[3]);
}

__device__ __forceinline__ unsigned split_threshold(unsigned long long total, unsigned tmin, unsigned long long target)
{
    const unsigned long long t = (total + target - 1) / target;
    return t > tmin ? (unsigned)(t < 0xffffffffull ? t : 0xffffffffull) : tmin;
}

// Split parts of each unit: T = max(tmin, ceil(total / target)), parts =
// cost > T ? ceil(cost / T) : 1.
__global__ void k_parts(int n, const unsigned *cost, const unsigned long long *total, unsigned tmin,
                        unsigned long long target, unsigned *parts, unsigned *sparts, unsigned *sflag)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const unsigned T = split_threshold(*total, tmin, target);
    const unsigned c = cost[b];
    const unsigned k = c > T ? (unsigned)(((unsigned long long)c + T - 1) / T) : 1u;
    parts[b] = k;
    sparts[b] = k > 1 ? k : 0u;
    sflag[b] = k > 1 ? 1u : 0u;
}

__global__ void k_items_fill(int n, const unsigned *parts, const unsigned *item0, const unsigned *slot0,
                             const unsigned *sidx, int4 *items, int4 *splits)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const int k = (int)parts[b];
    const int i0 = (int)item0[b];
    for (int j = 0; j < k; ++j) items[i0 + j] = make_int4(b, j, k, k > 1 ? (int)slot0[b] + j : -1);
    if (k > 1) splits[sidx[b]] = make_int4(b, (int)slot0[b], k, 0);
}

// Heaviest-first order of the items: bucket counts and slots are reserved
// once per workgroup (LDS counters, one global atomic per bucket), so the
// few large buckets do not serialise every item on one address.
__global__ __launch_bounds__(256) void k_item_hist(int n, const unsigned *n_dev, const int4 *items, const unsigned *cost,
                                                   unsigned *hist)
{
    if (n_dev) n = (int)*n_dev;   // the item count on the device (grid sized for the most items)
    __shared__ unsigned s_h[32];
    if (threadIdx.x < 32) s_h[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int4 it = items[i];
        atomicAdd(&s_h[cost_bucket(cost[it.x] / (unsigned)it.z)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 32 && s_h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], s_h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_item_order(int n, const unsigned *n_dev, const int4 *items, const unsigned *cost,
                                                    unsigned *offs,
                                                    int *order)
{
    if (n_dev) n = (int)*n_dev;
    __shared__ unsigned s_h[32], s_base[32];
    if (threadIdx.x < 32) s_h[threadIdx.x] = 0;
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int bk = 0;
    unsigned r = 0;
    if (i < n) {
        const int4 it = items[i];
        bk = cost_bucket(cost[it.x] / (unsigned)it.z);
        r = atomicAdd(&s_h[bk], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 32 && s_h[threadIdx.x]) s_base[threadIdx.x] = atomicAdd(&offs[threadIdx.x], s_h[threadIdx.x]);
    __syncthreads();
    if (i < n) order[s_base[bk] + r] = i;
}

// ---- exclusive scan of unsigned values (1024 per workgroup, two levels)
__global__ __launch_bounds__(256) void k_scan_local(int n, const unsigned *in, unsigned *out, unsigned *chunk_sums)
{
    __shared__ unsigned s_w[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * 1024 + tid * 4;
    unsigned v[4], t = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        t += v[k];
    }
    unsigned x = t;   // inclusive wave scan of thread totals
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned y = __shfl_up(x, d, WAVE);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_w[wave] = x;
    __syncthreads();
    unsigned woff = 0;
    for (int w = 0; w < wave; ++w) woff += s_w[w];
    unsigned run = woff + x - t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (tid == 255) chunk_sums[blockIdx.x] = woff + x;
}

__global__ __launch_bounds__(256) void k_scan_chunks(int nchunks, unsigned *chunk_sums, unsigned *total)
{
    __shared__ unsigned s_w[4];
    __shared__ unsigned s_carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < nchunks; base += 256) {
        const unsigned t = base + tid < nchunks ? chunk_sums[base + tid] : 0u;
        unsigned x = t;
        for (int d = 1; d < 64; d <<= 1) {
            const unsigned y = __shfl_up(x, d, WAVE);
            if (lane >= d) x += y;
        }
        if (lane == 63) s_w[wave] = x;
        __syncthreads();
        unsigned woff = s_carry;
        for (int w = 0; w < wave; ++w) woff += s_w[w];
        if (base + tid < nchunks) chunk_sums[base + tid] = woff + x - t;
        __syncthreads();
        if (tid == 255) s_carry = woff + x;
        __syncthreads();
    }
    if (tid == 0) *total = s_carry;
}

__global__ void k_scan_add(int n, unsigned *out, const unsigned *chunk_sums)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += chunk_sums[i / 1024];
}

// ==================================================== density at points ====
//
// One wave owns 64 consecutive query points in bin order (lane = query) and
// walks the candidate bins of its bounding box; the source records are
// wave-uniform and come in by scalar loads.  Each source uses its own
// bandwidth (sample-point estimator); FIXED specialises the common case of
// one bandwidth.  LOO excludes the query itself (same index); duplicates at
// distance 0 are kept.  Waves are launched heaviest first (k_wave_cost).
//
// Cut-off test: d2f = fl32(dx*dx + dy*dy) decides unless it lies within a
// relative 2^-20 band around the threshold (its error is below 2^-22
// relative); for a wave with any lane in the band the exact FP64 test
// fl64(dx^2 + dy^2) < thr is evaluated instead, so the result is exact.

struct PtsArgs {
    PointIndex idx;
    float *den;   // output, bin order
    int loo;
    double thr_fixed, a_fixed;   // FIXED kernels only
    // work items (nullable order: one item per wave): {wave, part j, parts K,
    // partial slot}; part j of a wave takes candidate ordinals
    // [C j / K, C (j+1) / K) of its C candidate sources (cost[wave])
    const int *order;
    const int4 *items;
    const unsigned *cost;
    double *partials;            // 64 per slot
    int n_items;
};

#define PD_BAND (1.0 / 1048576.0)

// Per-wave variant: each wave owns 64 consecutive queries in bin order and
// walks the candidate bins of its own bounding box; a source's record is
// wave-uniform, so it is read with scalar loads into SGPRs (no LDS, and the
// VALU reads SGPR operands for free).  FIXED: FP32 screen with the FP64
// re-check of a 32-source chunk when any lane is inside the 2^-20 band;
// otherwise (per-source bandwidths) the exact FP64 test for every pair.
__device__ __forceinline__ float wave_min(float v)
{
    for (int d = 32; d >= 1; d >>= 1) v = fminf(v, __shfl_xor(v, d, WAVE));
    return v;
}
__device__ __forceinline__ float wave_max(float v)
{
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, WAVE));
    return v;
}
__device__ __forceinline__ float uni(float v) { return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v))); }
__device__ __forceinline__ int4 uniform_item(int4 v)
{
    return make_int4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                     __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
}

// Work estimate of each 64-query wave of the point kernels: the points in
// its candidate bins (same culling as the kernels: bins within the reach of
// the wave's bounding box whose box gap is below the threshold -- lim_fixed
// when > 0, else lim_dev[0] when given, else each bin's own threshold).
// Waves then run heaviest first.
__global__ __launch_bounds__(256) void k_wave_cost(PointIndex I, double lim_fixed, const double *lim_dev,
                                                   unsigned *cost)
{
    const int lane = threadIdx.x & 63;
    const int64_t nw = (I.n + 63) / 64;
    const int64_t wv = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wv >= nw) return;
    const int64_t q = wv * 64 + lane;
    const bool qv = q < I.n;
    float qx = 0.0f, qy = 0.0f;
    if (qv) {
        const PtRec p = I.pts[q];
        qx = p.x;
        qy = p.y;
    }
    const float mnx = wave_min(qv ? qx : FLT_MAX), mxx = wave_max(qv ? qx : -FLT_MAX);
    const float mny = wave_min(qv ? qy : FLT_MAX), mxy = wave_max(qv ? qy : -FLT_MAX);
    const bool uni_lim = lim_fixed > 0.0 || lim_dev;
    const double lim = lim_fixed > 0.0 ? lim_fixed : (lim_dev ? *lim_dev : 0.0);
    const double reach = uni_lim ? sqrt(fmax(lim, 0.0)) * 1.000001 + 1e-6 : *I.reach_p;
    int bxa, bxb, bya, byb;
    bin_range(I.bg, mnx, mxx, mny, mxy, reach, bxa, bxb, bya, byb);
    const int nbx_r = bxb - bxa + 1;
    const int nb_total = nbx_r * (byb - bya + 1);
    unsigned acc = 0;
    for (int k = lane; k < nb_total; k += 64) {
        const int b = (bya + k / nbx_r) * I.bg.nbx + (bxa + k % nbx_r);
        const int c = I.bin_start[b + 1] - I.bin_start[b];
        if (c > 0 && box_gap2(I.bin_box[b], mnx, mxx, mny, mxy) < (uni_lim ? lim : I.bin_thr[b])) acc += (unsigned)c;
    }
    for (int d = 32; d >= 1; d >>= 1) acc += __shfl_xor(acc, d, WAVE);
    if (lane == 0) cost[wv] = acc;
}

template <bool FIXED, bool LOO>
__global__ __launch_bounds__(256) void k_point_density_w(PtsArgs A)
{
    const PointIndex &I = A.idx;
    const int lane = threadIdx.x & 63;
    // the wave index and the work item are wave-uniform: made explicit so the
    // source records below are read with scalar loads
    const int64_t g = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g >= (A.order ? (int64_t)A.n_items : (I.n + 63) / 64)) return;   // whole wave
    const int4 item = uniform_item(A.order ? A.items[A.order[g]] : make_int4((int)g, 0, 1, -1));
    int win0 = 0, win1 = INT_MAX;   // candidate counts are below 2^31
    if (item.z > 1) {
        const int64_t C = A.cost[item.x];
        if (item.y > 0) win0 = (int)(C * item.y / item.z);
        if (item.y < item.z - 1) win1 = (int)(C * (item.y + 1) / item.z);
    }
    const int64_t q0 = (int64_t)item.x * 64;
    const int64_t q = q0 + lane;
    const bool qv = q < I.n;
    float qx = 0.0f, qy = 0.0f;
    if (qv) {
        const PtRec p = I.pts[q];
        qx = p.x;
        qy = p.y;
    }
    const float mnx = uni(wave_min(qv ? qx : FLT_MAX)), mxx = uni(wave_max(qv ? qx : -FLT_MAX));
    const float mny = uni(wave_min(qv ? qy : FLT_MAX)), mxy = uni(wave_max(qv ? qy : -FLT_MAX));
    const double reach = FIXED ? sqrt(fmax(A.thr_fixed, 0.0)) * 1.000001 + 1e-6 : *I.reach_p;
    int bxa, bxb, bya, byb;
    bin_range(I.bg, mnx, mxx, mny, mxy, reach, bxa, bxb, bya, byb);
    bxa = __builtin_amdgcn_readfirstlane(bxa);
    bxb = __builtin_amdgcn_readfirstlane(bxb);
    bya = __builtin_amdgcn_readfirstlane(bya);
    byb = __builtin_amdgcn_readfirstlane(byb);

    const float f_thr_lo = FIXED ? (float)(A.thr_fixed * (1.0 - PD_BAND)) : 0.0f;
    const float f_thr_hi = FIXED ? fmaxf((float)(A.thr_fixed * (1.0 + PD_BAND)), FLT_MIN) : 0.0f;
    const float f_a = FIXED ? (float)A.a_fixed : 0.0f;
    float acc = 0.0f, sum = 0.0f, comp = 0.0f;
    int ord = 0;   // candidate ordinal at the start of the next candidate bin
    for (int by = bya; by <= byb; ++by) {
        for (int bx = bxa; bx <= bxb; ++bx) {
            const int b = by * I.bg.nbx + bx;
            const int bst = I.bin_start[b], ben = I.bin_start[b + 1];
            if (bst == ben) continue;
            const double lim = FIXED ? A.thr_fixed : I.bin_thr[b];
            const BinInfo bb = I.bin_box[b];
            if (!__builtin_amdgcn_readfirstlane((int)(box_gap2(bb, mnx, mxx, mny, mxy) < lim))) continue;
            // FIXED: every pair of the bin and the wave is inside the screen
            // when the largest fl32 offsets give d2f below it (fl32 of a
            // difference, the product and the fma are monotone in |offset|,
            // so no pair's d2f exceeds this one): such a bin's sources need no
            // compare and no select, and add the same terms.  LOO keeps the
            // tests for the bin holding the wave's own queries.
            bool all_in = false;
            if (FIXED) {
                const float gx = fmaxf(bb.x1 - mnx, mxx - bb.x0), gy = fmaxf(bb.y1 - mny, mxy - bb.y0);
                all_in = __builtin_fmaf(gy, gy, gx * gx) < f_thr_lo &&
                         (!LOO || (int64_t)ben <= q0 || (int64_t)bst >= q0 + 64);
                all_in = __builtin_amdgcn_readfirstlane((int)all_in);
            }
            // this part's share of the candidate sequence
            const int o0 = ord, o1 = ord + (ben - bst);
            ord = o1;
            if (o1 <= win0 || o0 >= win1) continue;
            const int st = bst + (win0 > o0 ? win0 - o0 : 0);
            const int en = ben - (o1 > win1 ? o1 - win1 : 0);
            for (int c0 = st; c0 < en; c0 += 32) {
                const int c1 = c0 + 32 < en ? c0 + 32 : en;
                // four records are loaded before their arithmetic, so the
                // scalar loads overlap
                if (FIXED && all_in) {
                    int si = c0;
                    for (; si + 4 <= c1; si += 4) {
                        const PtRec p0 = I.pts[si], p1 = I.pts[si + 1], p2 = I.pts[si + 2], p3 = I.pts[si + 3];
                        for (const PtRec &p : {p0, p1, p2, p3}) {
                            const float dx = p.x - qx, dy = p.y - qy;
                            const float d2f = __builtin_fmaf(dy, dy, dx * dx);
                            acc = __builtin_fmaf(fast_exp2(d2f * f_a), p.coef, acc);
                        }
                    }
                    for (; si < c1; ++si) {
                        const PtRec p = I.pts[si];
                        const float dx = p.x - qx, dy = p.y - qy;
                        const float d2f = __builtin_fmaf(dy, dy, dx * dx);
                        acc = __builtin_fmaf(fast_exp2(d2f * f_a), p.coef, acc);
                    }
                } else if (FIXED) {
                    bool band = false;
                    auto body = [&](const PtRec &p, int si) {
                        const float dx = p.x - qx, dy = p.y - qy;
                        const float d2f = __builtin_fmaf(dy, dy, dx * dx);
                        bool in = d2f < f_thr_lo;
                        band |= (d2f < f_thr_hi) != in;
                        if (LOO && (int64_t)si == q) in = false;
                        const float e = fast_exp2(d2f * f_a);
                        acc = __builtin_fmaf(in ? e : 0.0f, p.coef, acc);
                    };
                    int si = c0;
                    for (; si + 4 <= c1; si += 4) {
                        const PtRec p0 = I.pts[si], p1 = I.pts[si + 1], p2 = I.pts[si + 2], p3 = I.pts[si + 3];
                        body(p0, si);
                        body(p1, si + 1);
                        body(p2, si + 2);
                        body(p3, si + 3);
                    }
                    for (; si < c1; ++si) body(I.pts[si], si);
                    if (__any(band)) {   // rare: exact FP64 test for the chunk
                        acc = 0.0f;
                        for (int sj = c0; sj < c1; ++sj) {
                            const PtRec p = I.pts[sj];
                            const float dx = p.x - qx, dy = p.y - qy;
                            const float d2f = __builtin_fmaf(dy, dy, dx * dx);
                            bool in = sq2(dx, dy) < A.thr_fixed;
                            if (LOO && (int64_t)sj == q) in = false;
                            const float e = fast_exp2(d2f * f_a);
                            acc = __builtin_fmaf(in ? e : 0.0f, p.coef, acc);
                        }
                    }
                } else {
                    auto body = [&](const PtRec &p, int si) {
                        const float dx = p.x - qx, dy = p.y - qy;
                        const float d2f = __builtin_fmaf(dy, dy, dx * dx);
                        bool in = sq2(dx, dy) < p.thr;
                        if (LOO && (int64_t)si == q) in = false;
                        const float e = fast_exp2(d2f * (float)p.a);
                        acc = __builtin_fmaf(in ? e : 0.0f, p.coef, acc);
                    };
                    int si = c0;
                    for (; si + 4 <= c1; si += 4) {
                        const PtRec p0 = I.pts[si], p1 = I.pts[si + 1], p2 = I.pts[si + 2], p3 = I.pts[si + 3];
                        body(p0, si);
                        body(p1, si + 1);
                        body(p2, si + 2);
                        body(p3, si + 3);
                    }
                    for (; si < c1; ++si) body(I.pts[si], si);
                }
                kahan_fold(acc, sum, comp);
                acc = 0.0f;
            }
        }
    }
    if (item.z > 1) A.partials[(int64_t)item.w * 64 + lane] = (double)sum - (double)comp;
    else if (qv) A.den[q] = sum;
}

// Exclusive scan of one int per thread across a 256-thread block.
__device__ __forceinline__ int block_exscan256(int v, int *s_w, int *total)
{
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x >> 6;
    int inc = v;
    for (int d = 1; d < WAVE; d <<= 1) {
        int o = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += o;
    }
    if (lane == WAVE - 1) s_w[wave] = inc;
    __syncthreads();
    int off = 0;
#pragma unroll
    for (int w = 0; w < 4; ++w) off += (w < wave) ? s_w[w] : 0;
    *total = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    __syncthreads();
    return off + inc - v;
}



// ============================================= batched LOO (fixed bandwidths) ==
//
// Up to LM_K candidate bandwidths h_k (ascending) in one pass over point
// pairs.  coef_k of a source depends on k through the candidate's
// edge-correction weight; a pass of K candidates keeps candidate j in slot
// LM_K - K + j of each source's LM_K coefficients (bin order).

#define LM_K MKDE_LM_K

struct LooMultiArgs {
    PointIndex idx;          // positions + bins (thr of the largest candidate)
    const float *coefk;      // [n][MKDE_LM_K] bin order, candidate j of K in slot LM_K - K + j
    const double *a_k;       // [LM_K] by slot (0 for unused slots)
    const double *thr_k;     // [LM_K] by slot, ascending (-1 for unused slots)
    int K;
    float *den1;             // [K][n] bin order
    const int *order;        // work items as for the point kernel
    const int4 *items;
    const unsigned *cost;
    double *partials;        // 64 * MKDE_LM_K per slot
    int n_items;
    const unsigned *n_items_dev;   // non-null: the item count on the device (n_items: grid bound)
};

// Slots K .. K-3 of a source's candidate walk (then the next four while a
// lane is inside): `in` is slot K's lane mask; the block forms the masks of
// slots K-1 .. K-3 and of K-4 (the next block's first) at its start, adds
// the four terms, without selects when every lane is inside slot K-3 (then
// inside all four), else each term selected by its mask.  A slot no lane is
// inside adds exp2(..) * coef selected to +0 (acc + 0 = acc), so finishing a
// block past the last slot with a lane inside changes no sum, and the walk
// ends after the first block whose smallest slot is empty (the cut-off is
// monotone in h).  Two scalar tests per block instead of two per slot.  The
// four exponentials issue before their first use, so no wait state is
// needed after a transcendental; a mask is at least two instructions old
// when v_cndmask reads it.
#define MKDE_LT(I) "v_mul_f32_e32 %[t" #I "], %[d2f], %[a" #I "]\n\t"
#define MKDE_LE(I) "v_exp_f32_e32 %[t" #I "], %[t" #I "]\n\t"
#define MKDE_LS(I) "v_cndmask_b32_e64 %[t" #I "], 0, %[t" #I "], %[m" #I "]\n\t"
#define MKDE_LF(I) "v_fmac_f32_e32 %[c" #I "], %[f" #I "], %[t" #I "]\n\t"
template <int K>
__device__ __forceinline__ void loo_block(uint64_t in, double d2, float d2f, const double (&thr)[LM_K],
                                          const float (&af)[LM_K], const float (&cf)[LM_K], float (&acc)[LM_K])
{
    static_assert(K % 4 == 3, "blocks of four slots");
    if (in == 0) return;
    uint64_t m2, m1, m0, next = 0;
    asm("v_cmp_gt_f64_e64 %0, %3, %6\n\t"
        "v_cmp_gt_f64_e64 %1, %4, %6\n\t"
        "v_cmp_gt_f64_e64 %2, %5, %6"
        : "=&s"(m2), "=&s"(m1), "=&s"(m0)
        : "s"(thr[K - 1]), "s"(thr[K - 2]), "s"(thr[K - 3]), "v"(d2));
    if constexpr (K > 3) asm("v_cmp_gt_f64_e64 %0, %1, %2" : "=s"(next) : "s"(thr[K - 4]), "v"(d2));
    // one piece of assembly for both forms (a scalar branch inside it skips
    // the selects when every lane is inside slot K-3), so the accumulators
    // keep their registers whichever form runs
    float t0, t1, t2, t3;
    asm(MKDE_LT(0) MKDE_LT(1) MKDE_LT(2) MKDE_LT(3) MKDE_LE(0) MKDE_LE(1) MKDE_LE(2) MKDE_LE(3)
        "s_cmp_eq_u64 %[m3], -1\n\t"
        "s_cbranch_scc1 1f\n\t"
        MKDE_LS(0) MKDE_LS(1) MKDE_LS(2) MKDE_LS(3)
        "\n1:\n\t"
        MKDE_LF(0) MKDE_LF(1) MKDE_LF(2) MKDE_LF(3)
        : [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3), [c0] "+v"(acc[K]),
          [c1] "+v"(acc[K - 1]), [c2] "+v"(acc[K - 2]), [c3] "+v"(acc[K - 3])
        : [d2f] "v"(d2f), [a0] "v"(af[K]), [a1] "v"(af[K - 1]), [a2] "v"(af[K - 2]), [a3] "v"(af[K - 3]),
          [f0] "s"(cf[K]), [f1] "s"(cf[K - 1]), [f2] "s"(cf[K - 2]), [f3] "s"(cf[K - 3]), [m0] "s"(in),
          [m1] "s"(m2), [m2] "s"(m1), [m3] "s"(m0)
        : "scc");
    if constexpr (K > 3) loo_block<K - 4>(next, d2, d2f, thr, af, cf, acc);
}

// One source of the batched leave-one-out pass for the wave's 64 queries:
// candidate slot k's term for a lane is exp2(d2f a_k) coef_k if the exact
// cut-off fl64(dx^2 + dy^2) < thr_k holds, else 0, added to acc[k].  Slots are
// visited from the largest bandwidth down and the walk ends at the first one
// no lane is inside (the cut-off is monotone in h; unused slots sit below the
// smallest candidate with thr = -1).  selfm has the bit of the lane whose
// query is this source (or none): that lane gets d2f = +inf, so its terms are
// exp2(-inf) * coef = 0 (leave-one-out).  Each slot's compare and term are
// inline assembly so that its lane mask is consumed right after it is made
// (left to itself the compiler forms all 16 masks up front and runs out of
// scalar registers).
__device__ __forceinline__ void loo_source(const PtRec *pts, const float *coefk, int si, uint64_t selfm, float qx,
                                           float qy, float inf, const double (&thr)[LM_K], const float (&af)[LM_K],
                                           float (&acc)[LM_K])
{
    const PtRec p = pts[si];
    const float4 *cp = (const float4 *)(coefk + (int64_t)si * LM_K);
    const float4 cf0 = cp[0], cf1 = cp[1], cf2 = cp[2], cf3 = cp[3];
    const float cf[LM_K] = {cf0.x, cf0.y, cf0.z, cf0.w, cf1.x, cf1.y, cf1.z, cf1.w,
                            cf2.x, cf2.y, cf2.z, cf2.w, cf3.x, cf3.y, cf3.z, cf3.w};
    const float dx = p.x - qx, dy = p.y - qy;
    float d2f = __builtin_fmaf(dy, dy, dx * dx);
    const double d2 = sq2(dx, dy);
    if (selfm) asm("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(d2f) : "v"(inf), "s"(selfm));
    // slot k's mask is made one step ahead of its use: candidate k-1's compare
    // issues before candidate k's term, so the scalar test of a mask never
    // waits on the compare that made it.  The first compare also takes the
    // source's coefficients, so their scalar loads are issued together and
    // complete before the chain.
    uint64_t in;
    asm("v_cmp_gt_f64_e64 %0, %1, %2"
        : "=s"(in)
        : "s"(thr[LM_K - 1]), "v"(d2), "s"(cf[0]), "s"(cf[1]), "s"(cf[2]), "s"(cf[3]), "s"(cf[4]), "s"(cf[5]),
          "s"(cf[6]), "s"(cf[7]), "s"(cf[8]), "s"(cf[9]), "s"(cf[10]), "s"(cf[11]), "s"(cf[12]), "s"(cf[13]),
          "s"(cf[14]), "s"(cf[15]));
    loo_block<LM_K - 1>(in, d2, d2f, thr, af, cf, acc);
}

// Batched leave-one-out densities of up to 16 fixed bandwidths in one pass
// over point pairs: each wave owns 64 queries in bin order (lane = query) and
// walks the candidate bins of its own bounding box; a source's position and
// coefficients are wave-uniform and come in by scalar loads.  Per (query,
// source) pair the distance is computed twice: d2f = fl32(dx^2 + dy^2) for the
// exponent and d2 = fl64(dx^2 + dy^2), exact, for the cut-off (FP64 compares
// run at the FP32 rate on gfx950, so exactness costs one VALU instruction per
// candidate and no re-check).  The K candidates (ascending) occupy slots
// 16-K .. 15.  Thresholds live in SGPR pairs, the exponent factors in VGPRs
// (wave-uniform values), each source's coefficients in SGPRs.  The sums fold
// into Kahan pairs every 32 sources.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 6))) void k_loo_multi_w(LooMultiArgs A)
{
    const PointIndex &I = A.idx;
    const int lane = threadIdx.x & 63;
    const int off = LM_K - A.K;
    const int64_t n = I.n;
    const int64_t g = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t units = A.n_items_dev ? (int64_t)*A.n_items_dev : A.order ? (int64_t)A.n_items : (n + 63) / 64;
    if (g >= units) return;
    const int4 item = uniform_item(A.order ? A.items[A.order[g]] : make_int4((int)g, 0, 1, -1));
    int win0 = 0, win1 = INT_MAX;   // candidate counts are below 2^31
    if (item.z > 1) {
        const int64_t C = A.cost[item.x];
        if (item.y > 0) win0 = (int)(C * item.y / item.z);
        if (item.y < item.z - 1) win1 = (int)(C * (item.y + 1) / item.z);
    }
    const int64_t q0 = (int64_t)item.x * 64;
    const int64_t q = q0 + lane;
    const bool qv = q < n;
    float qx = 0.0f, qy = 0.0f;
    if (qv) {
        const PtRec p = I.pts[q];
        qx = p.x;
        qy = p.y;
    }
    double thr[LM_K];
    float af[LM_K];
#pragma unroll
    for (int k = 0; k < LM_K; ++k) {
        thr[k] = A.thr_k[k];
        // a wave-uniform value copied into a VGPR: the scalar file holds the
        // thresholds and the sources' coefficients
        const float a = (float)A.a_k[k];
        asm("v_mov_b32 %0, %1" : "=v"(af[k]) : "s"(a));
    }
    const double tmax = A.thr_k[LM_K - 1];
    const float mnx = uni(wave_min(qv ? qx : FLT_MAX)), mxx = uni(wave_max(qv ? qx : -FLT_MAX));
    const float mny = uni(wave_min(qv ? qy : FLT_MAX)), mxy = uni(wave_max(qv ? qy : -FLT_MAX));
    int bxa, bxb, bya, byb;
    bin_range(I.bg, mnx, mxx, mny, mxy, sqrt(tmax) * 1.000001 + 1e-6, bxa, bxb, bya, byb);
    bxa = __builtin_amdgcn_readfirstlane(bxa);
    bxb = __builtin_amdgcn_readfirstlane(bxb);
    bya = __builtin_amdgcn_readfirstlane(bya);
    byb = __builtin_amdgcn_readfirstlane(byb);
    const int q0i = (int)q0;   // bin-order indexes are below 2^30
    const float inf = INFINITY;

    float acc[LM_K], sum[LM_K], comp[LM_K];
#pragma unroll
    for (int k = 0; k < LM_K; ++k) { acc[k] = 0.0f; sum[k] = 0.0f; comp[k] = 0.0f; }
    int ord = 0;
    for (int by = bya; by <= byb; ++by) {
        for (int bx = bxa; bx <= bxb; ++bx) {
            const int b = by * I.bg.nbx + bx;
            const int bst = I.bin_start[b], ben = I.bin_start[b + 1];
            if (bst == ben) continue;
            if (!__builtin_amdgcn_readfirstlane((int)(box_gap2(I.bin_box[b], mnx, mxx, mny, mxy) < tmax))) continue;
            const int o0 = ord, o1 = ord + (ben - bst);
            ord = o1;
            if (o1 <= win0 || o0 >= win1) continue;
            const int st = bst + (win0 > o0 ? win0 - o0 : 0);
            const int en = ben - (o1 > win1 ? o1 - win1 : 0);
            for (int c0 = st; c0 < en; c0 += 32) {
                const int c1 = c0 + 32 < en ? c0 + 32 : en;
                for (int si = c0; si < c1; ++si) {
                    // the lane whose query is this source, if any (scalar)
                    const unsigned rel = (unsigned)(si - q0i);
                    const uint64_t selfm = rel < 64u ? (uint64_t)1 << rel : (uint64_t)0;
                    loo_source(I.pts, A.coefk, si, selfm, qx, qy, inf, thr, af, acc);
                }
#pragma unroll
                for (int k = 0; k < LM_K; ++k) {
                    kahan_fold(acc[k], sum[k], comp[k]);
                    acc[k] = 0.0f;
                }
            }
        }
    }
    // outputs by candidate index k - off (rows of den1, rows of a part's partials)
    if (item.z > 1) {
        double *pp = A.partials + (int64_t)item.w * 64 * LM_K + lane;
#pragma unroll
        for (int k = 0; k < LM_K; ++k)
            if (k >= off) pp[(k - off) * 64] = (double)sum[k] - (double)comp[k];
    } else {
#pragma unroll
        for (int k = 0; k < LM_K; ++k)
            if (qv && k >= off) A.den1[(int64_t)(k - off) * n + q] = sum[k];
    }
}

// Parts of split waves: FP64 partials added in part order, rounded once.
// splits[k] = {wave, slot0, parts}; m = values per query (1, or MKDE_LM_K).
__global__ void k_point_merge(int64_t n, int m, int K, const unsigned *n_split_dev, const int4 *splits,
                              const double *partials, float *den)
{
    if (n_split_dev && blockIdx.x >= *n_split_dev) return;   // grid sized for the most splits
    const int4 sp = splits[blockIdx.x];
    const int lane = threadIdx.x & 63, kk = threadIdx.x >> 6;   // 64 x 4 threads
    const int64_t q = (int64_t)sp.x * 64 + lane;
    if (q >= n) return;
    for (int k = kk; k < K; k += 4) {
        double t = 0.0;
        for (int j = 0; j < sp.z; ++j) t += partials[((int64_t)(sp.y + j) * m + k) * 64 + lane];
        den[(int64_t)k * n + q] = (float)t;
    }
}


// ==================================================== range candidates ====
// Candidate sets of N1 with the QUERY's own threshold (tests only).  Pass 1
// counts (out_idx == nullptr), pass 2 fills at offsets (bin-order indexes).

struct RangeArgs {
    PointIndex idx;       // the query threshold is the point's own pts[q].thr
    int64_t *counts;      // bin order
    const int64_t *offsets;
    int32_t *out_idx;
};

__global__ __launch_bounds__(256) void k_range(RangeArgs A)
{
    __shared__ int s_bins[256];
    __shared__ int s_wcnt[4];
    __shared__ float s_sx[256];
    __shared__ float s_sy[256];
    __shared__ float s_red[4][4];

    const PointIndex &I = A.idx;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t q = (int64_t)blockIdx.x * 256 + tid;
    const bool qv = q < I.n;
    float qx = 0.0f, qy = 0.0f;
    double qt = -1.0;
    if (qv) {
        PtRec p = I.pts[q];
        qx = p.x;
        qy = p.y;
        qt = p.thr;
    }
    float mnx = qv ? qx : FLT_MAX, mxx = qv ? qx : -FLT_MAX, mny = qv ? qy : FLT_MAX, mxy = qv ? qy : -FLT_MAX;
    for (int d = 32; d >= 1; d >>= 1) {
        mnx = fminf(mnx, __shfl_xor(mnx, d, WAVE));
        mxx = fmaxf(mxx, __shfl_xor(mxx, d, WAVE));
        mny = fminf(mny, __shfl_xor(mny, d, WAVE));
        mxy = fmaxf(mxy, __shfl_xor(mxy, d, WAVE));
    }
    if (lane == 0) {
        s_red[wave][0] = mnx;
        s_red[wave][1] = mxx;
        s_red[wave][2] = mny;
        s_red[wave][3] = mxy;
    }
    __syncthreads();
    mnx = fminf(fminf(s_red[0][0], s_red[1][0]), fminf(s_red[2][0], s_red[3][0]));
    mxx = fmaxf(fmaxf(s_red[0][1], s_red[1][1]), fmaxf(s_red[2][1], s_red[3][1]));
    mny = fminf(fminf(s_red[0][2], s_red[1][2]), fminf(s_red[2][2], s_red[3][2]));
    mxy = fmaxf(fmaxf(s_red[0][3], s_red[1][3]), fmaxf(s_red[2][3], s_red[3][3]));

    int bxa, bxb, bya, byb;
    const double reach = *I.reach_p;
    const double qthr_max = reach * reach;
    bin_range(I.bg, mnx, mxx, mny, mxy, reach, bxa, bxb, bya, byb);
    const int nbx_r = bxb - bxa + 1;
    const int nb_total = nbx_r * (byb - bya + 1);
    int64_t cnt_q = 0;
    int64_t wpos = (A.out_idx && qv) ? A.offsets[q] : 0;
    for (int base = 0; base < nb_total; base += 256) {
        const int kk = base + tid;
        bool cand = false;
        int b = 0;
        if (kk < nb_total) {
            b = (bya + kk / nbx_r) * I.bg.nbx + (bxa + kk % nbx_r);
            if (I.bin_start[b + 1] > I.bin_start[b])
                cand = box_gap2(I.bin_box[b], mnx, mxx, mny, mxy) < qthr_max;
        }
        int nb;
        int slot = block_compact256(cand, s_wcnt, &nb);
        if (cand) s_bins[slot] = b;
        __syncthreads();
        for (int kb = 0; kb < nb; ++kb) {
            const int bb = s_bins[kb];
            const int st = I.bin_start[bb], en = I.bin_start[bb + 1];
            for (int off = st; off < en; off += 256) {
                const int cnt = (en - off) < 256 ? (en - off) : 256;
                if (tid < cnt) {
                    PtRec p = I.pts[off + tid];
                    s_sx[tid] = p.x;
                    s_sy[tid] = p.y;
                }
                __syncthreads();
                if (qv) {
                    for (int s = 0; s < cnt; ++s) {
                        const double d2 = sq2(qx - s_sx[s], qy - s_sy[s]);
                        if (d2 < qt) {
                            if (A.out_idx) A.out_idx[wpos++] = off + s;
                            cnt_q++;
                        }
                    }
                }
                __syncthreads();
            }
        }
        __syncthreads();
    }
    if (qv && !A.out_idx) A.counts[q] = cnt_q;
}

// ============================================================ boundary ====

__global__ void k_valid(int64_t ncells, const float *mask, float nodata, uint8_t *valid)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncells) valid[i] = mask[i] != nodata ? 1 : 0;
}

// Valid-cell runs of the mask rows, one wave per row, columns 64 at a time:
// a run starts at a valid cell whose left neighbour is not valid (or at
// column 0) and ends at a valid cell whose right neighbour is not valid (or
// at the last column).  k_run_count counts each row's runs; k_run_fill
// writes them, in column order, from the row's first index (the exclusive
// scan of the counts): a start is run number (starts before it in the row),
// an end belongs to the last start at or before it.
__device__ __forceinline__ void run_flags(const uint8_t *row, int64_t ncols, int64_t c, bool &start, bool &end)
{
    const bool v = c < ncols && row[c];
    start = v && (c == 0 || !row[c - 1]);
    end = v && (c == ncols - 1 || !row[c + 1]);
}

__global__ __launch_bounds__(256) void k_run_count(int64_t ncols, int64_t nrows, const uint8_t *valid, unsigned *cnt)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nrows) return;   // whole wave
    const uint8_t *row = valid + r * ncols;
    unsigned n = 0;
    for (int64_t c0 = 0; c0 < ncols; c0 += WAVE) {
        bool s, e;
        run_flags(row, ncols, c0 + lane, s, e);
        n += (unsigned)__popcll(__ballot(s));
    }
    if (lane == 0) cnt[r] = n;
}

__global__ __launch_bounds__(256) void k_run_fill(int64_t ncols, int64_t nrows, const uint8_t *valid,
                                                  const unsigned *first, int2 *se)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= nrows) return;   // whole wave
    const uint8_t *row = valid + r * ncols;
    const uint64_t le = ~0ull >> (63 - lane);   // lanes 0 .. lane
    int64_t k = first[r];
    for (int64_t c0 = 0; c0 < ncols; c0 += WAVE) {
        const int64_t c = c0 + lane;
        bool s, e;
        run_flags(row, ncols, c, s, e);
        const uint64_t ms = __ballot(s);
        const int64_t upto = k + __popcll(ms & le);   // starts at columns <= c
        if (s) se[upto - 1].x = (int)c;
        if (e) se[upto - 1].y = (int)c;
        k += __popcll(ms);
    }
}

// run_ptr[r] = first index of row r's runs, run_ptr[nrows] = the total.
__global__ void k_run_ptr(int64_t nrows, const unsigned *first, const unsigned *total, int64_t *ptr)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrows) ptr[r] = first[r];
    else if (r == nrows) ptr[r] = *total;
}

// MarkBoundary: a valid cell is a boundary cell iff it is on the outer rows /
// columns or one of its 8 neighbours is nodata (tiled:3993-4059).  Boundary
// cell centres are appended to blist (order irrelevant: only minima are taken).
// Row r = blockIdx.y (+ gridDim.y strides), column from blockIdx.x: no
// 64-bit division per cell.  With blist the boundary centres are appended
// (bcount: one counter); without, only counted, in counter slots.
__global__ __launch_bounds__(256) void k_boundary(GridDesc g, const uint8_t *valid, uint8_t *flags, float2 *blist,
                                                  unsigned long long *bcount)
{
    const int64_t nc = g.ncols, nr = g.nrows;
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    for (int64_t r = blockIdx.y; r < nr; r += gridDim.y) {
    const int64_t i = r * nc + c;
    uint8_t f = 0;
    if (valid[i]) {
        if (r == 0 || r == nr - 1 || c == 0 || c == nc - 1) {
            f = 1;
        } else {
            const uint8_t *up = valid + (r - 1) * nc + c, *mid = valid + r * nc + c, *dn = valid + (r + 1) * nc + c;
            f = !(up[-1] & up[0] & up[1] & mid[-1] & mid[1] & dn[-1] & dn[0] & dn[1]);
        }
    }
    flags[i] = f;
    if (f) {
        if (blist) {
            unsigned long long k = atomicAdd(bcount, 1ull);
            const double nrm05 = (double)nr - 0.5;
            blist[k] = make_float2(cell_x(g.xll, g.cs, c), cell_y(g.yll, g.cs, nrm05, r));
        } else {
            atomicAdd(&bcount[((blockIdx.y * gridDim.x + blockIdx.x) % MKDE_CNT_SLOTS) * MKDE_CNT_STRIDE], 1ull);
        }
    }
    }
}

// CalcDist2Boundary (CPU semantics, tiled:4091-4131): min over boundary cells
// of fl64(dx^2 + dy^2), DBL_MAX if none.  Boundary centres streamed via LDS.
// A float screen skips a cell when fl32(dx*dx + dy*dy) exceeds the running
// minimum by more than a relative 2^-20 (its error is below 2^-22 relative),
// so only cells that can lower the minimum take the exact FP64 path.
__global__ __launch_bounds__(256) void k_dist_boundary(int64_t n, const float *x, const float *y, const float2 *blist,
                                                       int64_t nb, double *dist)
{
    __shared__ float2 s_b[1024];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float px = i < n ? x[i] : 0.0f, py = i < n ? y[i] : 0.0f;
    double m = DBL_MAX;
    float mf = FLT_MAX;   // screen: cells with d2f > mf cannot lower m
    for (int64_t base = 0; base < nb; base += 1024) {
        const int cnt = (nb - base) < 1024 ? (int)(nb - base) : 1024;
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) s_b[t] = blist[base + t];
        __syncthreads();
        for (int t = 0; t < cnt; ++t) {
            const float2 b = s_b[t];
            const float dx = b.x - px, dy = b.y - py;
            const float d2f = __builtin_fmaf(dy, dy, dx * dx);
            if (d2f <= mf) {
                const double d2 = sq2(dx, dy);
                if (d2 < m) {
                    m = d2;
                    mf = (float)(m * (1.0 + 1.0 / 1048576.0));
                    if (!(mf < FLT_MAX)) mf = FLT_MAX;
                }
            }
        }
        __syncthreads();
    }
    if (i < n) dist[i] = m;
}

// ---- boundary hierarchy and nearest-boundary search
//
// k_bt_leaf: one thread per 16x16 leaf: bit mask of its boundary cells and
// the bounding box of their centres (cell_x is non-decreasing in the column,
// cell_y non-increasing in the row, so the extremes are at the extreme
// indices).  k_bt_up: parents' boxes as unions of up to 4 children.
__global__ void k_bt_leaf(GridDesc g, const uint8_t *flags, int nx, int ny, uint4 *bits, float4 *box)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int bx = (int)(t % nx), by = (int)(t / nx);
    const int64_t c0 = (int64_t)bx * MKDE_BT_S, r0 = (int64_t)by * MKDE_BT_S;
    unsigned m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int cmin = INT_MAX, cmax = -1, rmin = INT_MAX, rmax = -1;
    for (int r = 0; r < MKDE_BT_S; ++r) {
        if (r0 + r >= g.nrows) break;
        const uint8_t *row = flags + (r0 + r) * g.ncols + c0;
        for (int c = 0; c < MKDE_BT_S; ++c) {
            if (c0 + c >= g.ncols) break;
            if (row[c]) {
                const int k = r * MKDE_BT_S + c;
                m[k >> 5] |= 1u << (k & 31);
                cmin = min(cmin, c);
                cmax = max(cmax, c);
                rmin = min(rmin, r);
                rmax = max(rmax, r);
            }
        }
    }
    bits[2 * t] = make_uint4(m[0], m[1], m[2], m[3]);
    bits[2 * t + 1] = make_uint4(m[4], m[5], m[6], m[7]);
    float4 b = make_float4(FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX);
    if (cmax >= 0) {
        const double nrm05 = (double)g.nrows - 0.5;
        b.x = cell_x(g.xll, g.cs, c0 + cmin);
        b.z = cell_x(g.xll, g.cs, c0 + cmax);
        b.y = cell_y(g.yll, g.cs, nrm05, r0 + rmax);
        b.w = cell_y(g.yll, g.cs, nrm05, r0 + rmin);
    }
    box[t] = b;
}

__global__ void k_bt_up(int cnx, int cny, const float4 *child, int nx, int ny, float4 *parent)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nx * ny) return;
    const int px = (int)(t % nx), py = (int)(t / nx);
    float4 b = make_float4(FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const int cx = 2 * px + dx, cy = 2 * py + dy;
            if (cx >= cnx || cy >= cny) continue;
            const float4 c = child[(int64_t)cy * cnx + cx];
            if (c.x > c.z) continue;
            b.x = fminf(b.x, c.x);
            b.y = fminf(b.y, c.y);
            b.z = fmaxf(b.z, c.z);
            b.w = fmaxf(b.w, c.w);
        }
    parent[t] = b;
}

// Lower bound of fl64(dx^2 + dy^2) over the cells of a node (gap1 is an exact
// lower bound of |fl32(u - v)| over an interval; both squares are exact).
__device__ __forceinline__ double node_lb(float4 b, float px, float py)
{
    return b.x > b.z ? DBL_MAX : sq2(gap1(px, b.x, b.z), gap1(py, b.y, b.w));
}

#define BT_STACK 48
#define BT_WG 128

// CalcDist2Boundary by depth-first search, nearest child first, pruning
// nodes whose lower bound is not below the best squared distance so far:
// the result is the brute-force minimum over all boundary cells, exactly.
// Points are taken in bin order (neighbouring lanes search alike) and the
// result is scattered to the caller's order.
__global__ __launch_bounds__(BT_WG) void k_dist_tree(int64_t n, const float *xs, const float *ys, const int *perm,
                                                     GridDesc g, BndTree T, double *dist)
{
    __shared__ unsigned s_stack[BT_STACK * BT_WG];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * BT_WG + tid;
    if (i >= n) return;
    const float px = xs[i], py = ys[i];
    const double nrm05 = (double)g.nrows - 0.5;
    double best = DBL_MAX;
    int sp = 0;
    const int top = T.nlev - 1;
    s_stack[(sp++) * BT_WG + tid] = (unsigned)top << 28;
    while (sp > 0) {
        const unsigned e = s_stack[(--sp) * BT_WG + tid];
        const int lev = (int)(e >> 28), iy = (int)((e >> 14) & 0x3fff), ix = (int)(e & 0x3fff);
        const int64_t node = T.off[lev] + (int64_t)iy * T.nx[lev] + ix;
        if (node_lb(T.box[node], px, py) >= best) continue;
        if (lev == 0) {
            const uint4 m0 = T.bits[2 * node], m1 = T.bits[2 * node + 1];
            const unsigned m[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            const int64_t c0 = (int64_t)ix * MKDE_BT_S, r0 = (int64_t)iy * MKDE_BT_S;
            for (int w = 0; w < 8; ++w) {
                unsigned v = m[w];
                while (v) {
                    const int k = w * 32 + __ffs(v) - 1;
                    v &= v - 1;
                    const float cx = cell_x(g.xll, g.cs, c0 + (k & (MKDE_BT_S - 1)));
                    const float cy = cell_y(g.yll, g.cs, nrm05, r0 + (k >> 4));
                    const double d2 = sq2(cx - px, cy - py);
                    best = d2 < best ? d2 : best;
                }
            }
            continue;
        }
        // children sorted by lower bound, pushed farthest first
        const int cl = lev - 1;
        unsigned ce[4];
        double cb[4];
        int nc = 0;
        for (int dy = 0; dy < 2; ++dy)
            for (int dx = 0; dx < 2; ++dx) {
                const int cx = 2 * ix + dx, cy = 2 * iy + dy;
                if (cx >= T.nx[cl] || cy >= T.ny[cl]) continue;
                const double lb = node_lb(T.box[T.off[cl] + (int64_t)cy * T.nx[cl] + cx], px, py);
                if (lb >= best) continue;
                int k = nc++;
                while (k > 0 && cb[k - 1] < lb) {   // descending lower bounds
                    cb[k] = cb[k - 1];
                    ce[k] = ce[k - 1];
                    --k;
                }
                cb[k] = lb;
                ce[k] = ((unsigned)cl << 28) | ((unsigned)cy << 14) | (unsigned)cx;
            }
        for (int k = 0; k < nc; ++k) s_stack[(sp++) * BT_WG + tid] = ce[k];
    }
    dist[perm[i]] = best;
}

// ===================================================== edge correction ====
//
// EdgeCorrectionWeightsExact with O(R) work per point instead of O(R^2).
// The integrand is separable on the mask grid:
//   K(c, r) = exp2(dx2_c a) * exp2(dy2_r a) / (2 pi h2f)
// and for a fixed row the cells inside the cut-off form one interval of
// columns (the reference's test is monotone in |dx|).  One wave handles one
// point: columns of the NARROW box (tiled:2716-2727) are processed in chunks
// of EC_CW (1024), with an FP64 prefix sum of exp2(dx2 a) in LDS; lanes then
// take rows, find each row's exact column interval, intersect it with the
// row's valid-cell runs and add exp2(dy2 a) * (prefix differences).
// Output ew = 1/(cellArea * sum / (2 pi h2f)); points whose boundary distance
// is >= 9 h2f were given weight 1 by k_ec_active.


struct EcArgs {
    GridDesc g;
    MaskRuns runs;
    const float *colx, *rowy;   // cell-centre tables of the mask grid
    const float *x, *y;     // original order
    const float *h;         // original order (nullable when h_fixed > 0 or h_dev)
    float h_fixed;
    const float *h_dev;     // non-null: every point's bandwidth is *h_dev (a device value)
    const int *count;       // non-null: {narrow, wide} list lengths on the device
    const int *list;        // active points
    float *ew;              // original order
    unsigned long long *n_eval;
    int cx32;               // cell centres of the mask's columns are exact in FP32 (host-checked)
};

__device__ __forceinline__ float ec_bandwidth(const EcArgs &A, int p)
{
    return A.h_dev ? *A.h_dev : (A.h_fixed > 0.0f ? A.h_fixed : A.h[p]);
}

// NARROW box with the reference's macro semantics (Utilities.h:16-17,30-31;
// tiled:2716-2727): size_t arithmetic for row_upper / col_upper.
__device__ __forceinline__ void narrow_box(const GridDesc &g, float px, float py, float h, int &rl, int &ru, int &cl,
                                           int &cu)
{
    const uint64_t nR = (uint64_t)g.nrows, nC = (uint64_t)g.ncols;
    const double yll = g.yll, xll = g.xll, cs = g.cs, hh = h;
    const int k_top = (int)(((double)py + 3.0 * hh - yll) / cs);
    const int r = (int)(int32_t)(uint32_t)(nR - 1 - (uint64_t)(int64_t)k_top);
    rl = 0 < r ? r : 0;
    const int k_bot = (int)(((double)py - 3.0 * hh - yll) / cs);
    const uint64_t yu = nR - 1 - (uint64_t)(int64_t)k_bot;
    ru = (int)(int32_t)(uint32_t)((nR - 1 < yu) ? nR - 1 : yu);
    const int k_left = (int)(((double)px - 3.0 * hh - xll) / cs);
    cl = 0 < k_left ? k_left : 0;
    const int k_right = (int)(((double)px + 3.0 * hh - xll) / cs);
    const uint64_t kr = (uint64_t)(int64_t)k_right;
    cu = (int)(int32_t)(uint32_t)((nC - 1 < kr) ? nC - 1 : kr);
}

// Column centres of the mask grid: computed (cell_x) for ColCompute, and
// dx = fl32(cell_x(c) - px) of a point: computed (DxCompute) or from the
// point's table in LDS (DxLds, written while its prefix sums are formed);
// both give the same dx.
struct ColCompute {
    double xll, cs;
    // columns of one raster row fit in 32 bits (one-instruction conversion)
    __device__ __forceinline__ float operator()(int64_t c) const { return cell_x(xll, cs, (int64_t)(int)c); }
};
// CX32: the grid's centres fl32(xll + (c + 0.5) cs) (FP64 arithmetic) equal
// fl32(fma(c + 0.5, cs, xll)) in FP32 for every column (checked on the host
// column by column: then the FP32 form, three instructions instead of five,
// gives the same centres).
template <bool CX32>
struct DxCompute {
    double xll, cs;
    float px;
    __device__ __forceinline__ float operator()(int64_t c) const
    {
        if constexpr (CX32) return __builtin_fmaf((float)(int)c + 0.5f, (float)cs, (float)xll) - px;
        else return ColCompute{xll, cs}(c) - px;
    }
};
struct DxLds {
    const float *dx;   // dx of columns c0, c0 + 1, ...
    int c0;
    __device__ __forceinline__ float operator()(int64_t c) const { return dx[(int)c - c0]; }
};

template <class DX>
__device__ __forceinline__ bool pred_col(const DX &dxf, int64_t c, float px, double dy2, double thr)
{
    const float dx = dxf(c);
    return __dadd_rn(sqd(dx), dy2) < thr;
}

// Exact column interval [lo, hi] within [cl, cu] of the cells of one row that
// are inside the cut-off; lo > hi when empty.  cn is the column of [cl, cu]
// nearest to px (|dx| is smallest there, and the predicate is monotone in the
// distance from cn on either side).  The first guess (FP32 square root,
// multiplication by 1/cs) only seeds the search: the four predicates that
// confirm it (both ends, one step beyond each end) are evaluated together, so
// table loads of their centres are in flight at once; when the seed is off,
// the row is empty unless cn is inside, and the ends are moved by the exact
// predicate until they are the last columns inside.
template <class DX>
__device__ __forceinline__ void row_interval(const DX &colx, double xll, double inv_cs, int cl, int cu, int cn,
                                             float px, double dy2, double thr, int &lo, int &hi)
{
    const double rem = thr - dy2;
    const double w = rem > 0.0 ? (double)sqrtf((float)rem) : 0.0;
    const double fh = floor(((double)px + w - xll) * inv_cs - 0.5);
    const double fl = ceil(((double)px - w - xll) * inv_cs - 0.5);
    int h_e = (int)fmin(fmax(fh, (double)cn), (double)cu);
    int l_e = (int)fmin(fmax(fl, (double)cl), (double)cn);
    const int h1 = h_e < cu ? h_e + 1 : h_e, l1 = l_e > cl ? l_e - 1 : l_e;
    const bool p_h = pred_col(colx, h_e, px, dy2, thr);
    const bool p_h1 = pred_col(colx, h1, px, dy2, thr) & (h_e < cu);
    const bool p_l = pred_col(colx, l_e, px, dy2, thr);
    const bool p_l1 = pred_col(colx, l1, px, dy2, thr) & (l_e > cl);
    if (!(p_h & !p_h1 & p_l & !p_l1)) {
        // cn is inside whenever any column is (p_h implies it)
        if (!pred_col(colx, cn, px, dy2, thr)) {
            lo = 1;
            hi = 0;
            return;
        }
        while (h_e < cu && pred_col(colx, h_e + 1, px, dy2, thr)) ++h_e;
        while (h_e > cn && !pred_col(colx, h_e, px, dy2, thr)) --h_e;
        while (l_e > cl && pred_col(colx, l_e - 1, px, dy2, thr)) --l_e;
        while (l_e < cn && !pred_col(colx, l_e, px, dy2, thr)) ++l_e;
    }
    lo = l_e;
    hi = h_e;
}

// Points whose boundary distance is below the cut-off get a work-list entry:
// count[0] narrow points from the front of list (box at most cw_narrow
// columns wide: k_ec, one wave per point, column chunks of cw_narrow),
// count[1] wide points from the back (wider boxes up to cw_wide columns:
// k_ec_wide, the whole box in LDS), count[2] the widest wide box.
__global__ __launch_bounds__(256) void k_ec_active(int64_t n, GridDesc g, const float *x, const float *y,
                                                   const double *dist, const float *h, float h_fixed,
                                                   const float *h_dev, float *ew, int *list, int *count,
                                                   int cw_narrow, int cw_wide)
{
    // list slots are reserved per workgroup (LDS counters, then one global
    // atomic per list): per-wave atomics on one address serialise
    __shared__ int s_n[2], s_base[2], s_wmax;
    const int tid = threadIdx.x;
    if (tid == 0) {
        s_n[0] = s_n[1] = 0;
        s_wmax = 0;
    }
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + tid;
    int kind = -1, slot = 0;   // 0 narrow, 1 wide
    if (i < n) {
        const float hh = h_dev ? *h_dev : (h_fixed > 0.0f ? h_fixed : h[i]);
        const double h2 = (double)(float)(hh * hh);
        if (dist[i] >= 9.0 * h2) {
            ew[i] = 1.0f;
        } else {
            int rl, ru, cl, cu;
            narrow_box(g, x[i], y[i], hh, rl, ru, cl, cu);
            const int w = cu - cl + 1;
            kind = (rl <= ru && w > cw_narrow && w <= cw_wide) ? 1 : 0;
            slot = atomicAdd(&s_n[kind], 1);
            if (kind == 1) atomicMax(&s_wmax, w);
        }
    }
    __syncthreads();
    if (tid < 2 && s_n[tid]) s_base[tid] = atomicAdd(&count[tid], s_n[tid]);
    if (tid == 2 && s_wmax) atomicMax(&count[2], s_wmax);
    __syncthreads();
    if (kind == 0) list[s_base[0] + slot] = (int)i;
    else if (kind == 1) list[n - 1 - (s_base[1] + slot)] = (int)i;
}

// Cell-centre tables of a grid: colx[c] = cell_x(c), rowy[r] = cell_y(r).
__global__ void k_centres(GridDesc g, float *colx, float *rowy)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < g.ncols) colx[i] = cell_x(g.xll, g.cs, i);
    if (i < g.nrows) rowy[i] = cell_y(g.yll, g.cs, (double)g.nrows - 0.5, i);
}

// First run index in [rs, re) whose end >= c (binary search).
__device__ __forceinline__ int64_t first_run_ending_at_or_after(const int2 *runs, int64_t rs, int64_t re, int c)
{
    int64_t lo = rs, hi = re;
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (runs[mid].y < c) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// One row r of the edge-correction integral of a point: the exact column
// interval of the row (row_interval) clipped to the columns [ca, cb] whose
// prefix sums of Ex are in s_pd, intersected with the row's valid runs;
// acc += exp2(dy2 a) * (sum of Ex over those cells), cnt += their number.
// Row and column centres are computed (cell_y / cell_x, the values of the
// mask's centre tables) rather than loaded, and the row's run range [rs, re)
// and, for a single-run row, its run r0 come loaded ahead by ec_rows, so the
// row's arithmetic waits on no memory.
template <class DX>
__device__ __forceinline__ void ec_row(const EcArgs &A, const DX &dxf, const double *s_pd, int ca, int cb, int cl,
                                       int cu, int cn,
                                       int r, int64_t rs, int64_t re, int2 r0, float px, float py, double thr,
                                       double a, double xll, double inv_cs, double yll, double cs, double nrm05,
                                       double &acc, unsigned long long &cnt)
{
    const float cy = cell_y(yll, cs, nrm05, r);
    const double dy2 = sqd(cy - py);
    if (!(dy2 < thr)) return;
    int lo, hi;
    row_interval(dxf, xll, inv_cs, cl, cu, cn, px, dy2, thr, lo, hi);
    if (lo < ca) lo = ca;
    if (hi > cb) hi = cb;
    if (lo > hi) return;
    double seg = 0.0;
    if (re - rs == 1) {
        const int s = r0.x > lo ? r0.x : lo;
        const int e = r0.y < hi ? r0.y : hi;
        if (s <= e) {
            seg = s_pd[e - ca] - (s > ca ? s_pd[s - 1 - ca] : 0.0);
            cnt += (unsigned long long)(e - s + 1);
        }
    } else {
        for (int64_t k = first_run_ending_at_or_after(A.runs.run_se, rs, re, lo); k < re; ++k) {
            const int2 rr = A.runs.run_se[k];
            if (rr.x > hi) break;
            const int s = rr.x > lo ? rr.x : lo;
            const int e = rr.y < hi ? rr.y : hi;
            seg += s_pd[e - ca] - (s > ca ? s_pd[s - 1 - ca] : 0.0);
            cnt += (unsigned long long)(e - s + 1);
        }
    }
    const float ey = fast_exp2((float)(dy2 * a));
    acc = __builtin_fma(seg, (double)ey, acc);
}

// Rows r, r + step, ... <= ru of a point's integral (ec_row each).  The run
// range of the next row and the single run of the current one are loaded at
// the top of an iteration, so their latency hides behind the row's interval
// search.
template <class DX>
__device__ __forceinline__ void ec_rows(const EcArgs &A, const DX &dxf, const double *s_pd, int ca, int cb, int cl,
                                        int cu, int cn, int r, int ru, int step, float px, float py, double thr,
                                        double a, double &acc, unsigned long long &cnt)
{
    const GridDesc &g = A.g;
    const double xll = g.xll, yll = g.yll, cs = g.cs, inv_cs = 1.0 / cs, nrm05 = (double)g.nrows - 0.5;
    int64_t rs = 0, re = 0;
    if (r <= ru) {
        rs = A.runs.run_ptr[r];
        re = A.runs.run_ptr[r + 1];
    }
    for (; r <= ru; r += step) {
        int2 r0 = make_int2(1, 0);
        if (re - rs == 1) r0 = A.runs.run_se[rs];
        const int rn = r + step;
        int64_t rs_n = 0, re_n = 0;
        if (rn <= ru) {
            rs_n = A.runs.run_ptr[rn];
            re_n = A.runs.run_ptr[rn + 1];
        }
        ec_row(A, dxf, s_pd, ca, cb, cl, cu, cn, r, rs, re, r0, px, py, thr, a, xll, inv_cs, yll, cs, nrm05, acc,
               cnt);
        rs = rs_n;
        re = re_n;
    }
}

// EC weight of point p by one wave: column chunks of EC_CW with the FP64
// prefix sums of Ex in s_pd, rows on lanes.
template <int EC_CW, bool CX32>
__device__ __forceinline__ void ec_point_t(const EcArgs &A, double *s_pd, int p)
{
    const int lane = threadIdx.x;
    const GridDesc g = A.g;
    const double xll = g.xll, cs = g.cs;
    const float px = A.x[p], py = A.y[p];
    const float hh = ec_bandwidth(A, p);
    const double h2 = (double)(float)(hh * hh);
    const double thr = 9.0 * h2;
    const double a = -MKDE_LOG2E / (2.0 * h2);

    int rl, ru, cl, cu;
    narrow_box(g, px, py, hh, rl, ru, cl, cu);
    double acc = 0.0;
    unsigned long long cnt = 0;
    if (rl <= ru && cl <= cu) {
        const int cn = clampi((int64_t)floor(((double)px - xll) / cs), cl, cu);
        for (int ca = cl; ca <= cu; ca += EC_CW) {
            const int cb = (ca + EC_CW - 1 < cu) ? ca + EC_CW - 1 : cu;
            const int w = cb - ca + 1;
            // lane owns columns [ca + lane*per, ...); an odd per puts the
            // lanes' FP64 prefix entries in distinct LDS banks
            const int per = ((w + WAVE - 1) / WAVE) | 1;
            const int my0 = lane * per;
            // FP64 prefix (53-bit significand, more than the 48 of a
            // float-float pair; FP64 adds cost one VALU instruction each)
            double run = 0.0;
            for (int t = 0; t < per; ++t) {
                const int k = my0 + t;
                if (k < w) {
                    const float dx = DxCompute<CX32>{xll, cs, px}(ca + k);
                    const float ex = fast_exp2((float)(sqd(dx) * a));
                    run += (double)ex;
                    s_pd[k] = run;
                }
            }
            // exclusive wave scan of the lane totals
            double inc = run;
            for (int d = 1; d < WAVE; d <<= 1) {
                const double o = __shfl_up(inc, d, WAVE);
                if (lane >= d) inc += o;
            }
            const double offv = inc - run;
            for (int t = 0; t < per; ++t) {
                const int k = my0 + t;
                if (k < w) s_pd[k] += offv;
            }
            __syncthreads();
            // rows
            ec_rows(A, DxCompute<CX32>{xll, cs, px}, s_pd, ca, cb, cl, cu, cn, rl + lane, ru, WAVE, px, py, thr, a,
                    acc, cnt);
            __syncthreads();
        }
    }
    // wave reduction
    for (int d = 32; d >= 1; d >>= 1) {
        acc += __shfl_xor(acc, d, WAVE);
        cnt += __shfl_xor(cnt, d, WAVE);
    }
    if (lane == 0) {
        const double ec = acc * (double)(float)(g.cs * g.cs) /
                          (h2 * (double)MKDE_TWO_PI_F);
        A.ew[p] = (float)(1.0 / ec);
        if (A.n_eval) slot_add(A.n_eval, cnt);
    }
}

// One wave per point of the narrow part of the list: a workgroup takes list
// items blockIdx.x, blockIdx.x + gridDim.x, ... below the narrow count (the
// device count when given, else one item per workgroup).
template <int EC_CW>
__device__ __forceinline__ void ec_point(const EcArgs &A, double *s_pd, int p)
{
    if (A.cx32) ec_point_t<EC_CW, true>(A, s_pd, p);
    else ec_point_t<EC_CW, false>(A, s_pd, p);
}

template <int EC_CW>
__global__ __launch_bounds__(64) void k_ec(EcArgs A)
{
    __shared__ double s_pd[EC_CW];   // prefix sums of Ex over the chunk's columns
    const int lim = A.count ? A.count[0] : (int)gridDim.x;
    for (int it = blockIdx.x; it < lim; it += gridDim.x) ec_point<EC_CW>(A, s_pd, A.list[it]);
}

// Wide points (NARROW box wider than the column chunk of k_ec): one
// 256-thread workgroup per point and the whole box's prefix sums in dynamic
// LDS, so every row is visited once instead of once per chunk.  Same row
// arithmetic as k_ec; with DXT the rows' predicates read each column's
// dx = fl32(cell_x - px) from LDS (written with the prefix, 4 B per column
// next to the 8 B sum) instead of recomputing the cell centre in FP64 (the
// same value).  The prefix is scanned across the workgroup (four waves).
template <bool DXT, bool CX32>
__device__ __forceinline__ void ec_point_wide_t(const EcArgs &A, double *s_pd, float *s_dx, double *s_wt,
                                                double *s_acc, unsigned long long *s_cnt, int p)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const GridDesc g = A.g;
    const double xll = g.xll, cs = g.cs;
    const float px = A.x[p], py = A.y[p];
    const float hh = ec_bandwidth(A, p);
    const double h2 = (double)(float)(hh * hh);
    const double thr = 9.0 * h2;
    const double a = -MKDE_LOG2E / (2.0 * h2);

    int rl, ru, cl, cu;
    narrow_box(g, px, py, hh, rl, ru, cl, cu);   // non-empty: k_ec_active checked
    const int cn = clampi((int64_t)floor(((double)px - xll) / cs), cl, cu);
    const int w = cu - cl + 1;
    // thread owns columns [cl + tid*per, ...); an odd per puts the threads'
    // FP64 prefix entries in distinct LDS banks
    const int per = ((w + 255) / 256) | 1;
    const int my0 = tid * per;
    double run = 0.0;
    for (int t = 0; t < per; ++t) {
        const int k = my0 + t;
        if (k < w) {
            const float dx = DxCompute<CX32>{xll, cs, px}(cl + k);
            const float ex = fast_exp2((float)(sqd(dx) * a));
            run += (double)ex;
            s_pd[k] = run;
            if (DXT) s_dx[k] = dx;
        }
    }
    // exclusive scan of the thread totals: within each wave, then the waves
    double inc = run;
    for (int d = 1; d < WAVE; d <<= 1) {
        const double o = __shfl_up(inc, d, WAVE);
        if (lane >= d) inc += o;
    }
    if (lane == WAVE - 1) s_wt[wave] = inc;
    __syncthreads();
    double woff = 0.0;
    for (int v = 0; v < wave; ++v) woff += s_wt[v];
    const double offv = woff + (inc - run);
    for (int t = 0; t < per; ++t) {
        const int k = my0 + t;
        if (k < w) s_pd[k] += offv;
    }
    __syncthreads();
    double acc = 0.0;
    unsigned long long cnt = 0;
    if (DXT) ec_rows(A, DxLds{s_dx, cl}, s_pd, cl, cu, cl, cu, cn, rl + tid, ru, 256, px, py, thr, a, acc, cnt);
    else ec_rows(A, DxCompute<CX32>{xll, cs, px}, s_pd, cl, cu, cl, cu, cn, rl + tid, ru, 256, px, py, thr, a, acc,
                 cnt);
    for (int d = 32; d >= 1; d >>= 1) {
        acc += __shfl_xor(acc, d, WAVE);
        cnt += __shfl_xor(cnt, d, WAVE);
    }
    if (lane == 0) {
        s_acc[wave] = acc;
        s_cnt[wave] = cnt;
    }
    __syncthreads();
    if (tid == 0) {
        const double tot = ((s_acc[0] + s_acc[1]) + s_acc[2]) + s_acc[3];
        const double ec = tot * (double)(float)(g.cs * g.cs) / (h2 * (double)MKDE_TWO_PI_F);
        A.ew[p] = (float)(1.0 / ec);
        if (A.n_eval) slot_add(A.n_eval, s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3]);
    }
}

// One workgroup per point of the wide part of the list (stored from the back
// of the list), items blockIdx.x, blockIdx.x + gridDim.x, ... below the wide
// count (the device count when given, else one item per workgroup).
template <bool DXT>
__device__ __forceinline__ void ec_point_wide(const EcArgs &A, double *s_pd, float *s_dx, double *s_wt,
                                              double *s_acc, unsigned long long *s_cnt, int p)
{
    if (A.cx32) ec_point_wide_t<DXT, true>(A, s_pd, s_dx, s_wt, s_acc, s_cnt, p);
    else ec_point_wide_t<DXT, false>(A, s_pd, s_dx, s_wt, s_acc, s_cnt, p);
}

template <bool DXT>
__global__ __launch_bounds__(256) void k_ec_wide(EcArgs A, int64_t n, int wmax)
{
    extern __shared__ double s_pd[];   // wmax prefix sums, then (DXT) wmax dx values
    float *s_dx = (float *)(s_pd + wmax);
    __shared__ double s_wt[4];
    __shared__ double s_acc[4];
    __shared__ unsigned long long s_cnt[4];
    const int lim = A.count ? A.count[1] : (int)gridDim.x;
    for (int it = blockIdx.x; it < lim; it += gridDim.x) {
        ec_point_wide<DXT>(A, s_pd, s_dx, s_wt, s_acc, s_cnt, A.list[n - 1 - it]);
        __syncthreads();   // the next point rewrites the LDS this one read
    }
}

// Exact count of (point, valid cell) pairs inside the cut-off on a tile (the
// density kernel's algorithmic work), one wave per point, rows on lanes.
struct CountArgs {
    GridDesc g;
    MaskRuns runs;      // valid runs of the tile (all-valid tiles have one run per row)
    PointIndex idx;
    unsigned long long *total;
};

__global__ __launch_bounds__(64) void k_count_pairs(CountArgs A)
{
    const int lane = threadIdx.x;
    const int64_t i = blockIdx.x;
    const PtRec P = A.idx.pts[i];
    const GridDesc g = A.g;
    unsigned long long cnt = 0;
    if (P.thr > 0.0) {
        const double xll = g.xll, yll = g.yll, cs = g.cs, nrm05 = (double)g.nrows - 0.5;
        const double h = sqrt(P.thr / 9.0);
        int rl, ru, cl, cu;
        narrow_box(g, P.x, P.y, (float)(h * 1.000001) + 1e-6f, rl, ru, cl, cu);
        if (rl <= ru && cl <= cu) {
            const int cn = clampi((int64_t)floor(((double)P.x - xll) / cs), cl, cu);
            for (int r = rl + lane; r <= ru; r += WAVE) {
                const float cy = cell_y(yll, cs, nrm05, r);
                const double dy2 = sqd(cy - P.y);
                if (!(dy2 < P.thr)) continue;
                int lo, hi;
                row_interval(DxCompute<false>{xll, cs, P.x}, xll, 1.0 / cs, cl, cu, cn, P.x, dy2, P.thr, lo, hi);
                if (lo > hi) continue;
                const int64_t rs = A.runs.run_ptr[r], re = A.runs.run_ptr[r + 1];
                for (int64_t k = first_run_ending_at_or_after(A.runs.run_se, rs, re, lo); k < re; ++k) {
                    const int2 rr = A.runs.run_se[k];
                    if (rr.x > hi) break;
                    const int s = rr.x > lo ? rr.x : lo;
                    const int e = rr.y < hi ? rr.y : hi;
                    cnt += (unsigned long long)(e - s + 1);
                }
            }
        }
    }
    for (int d = 32; d >= 1; d >>= 1) cnt += __shfl_xor(cnt, d, WAVE);
    if (lane == 0) slot_add(A.total, cnt);
}

// reach = sqrt(max bin threshold) (world units), slightly enlarged so that the
// bin range computed from it is a superset.
__global__ __launch_bounds__(256) void k_reach(int nbins, const double *bin_thr, double *reach)
{
    __shared__ double s[4];
    double m = 0.0;
    for (int b = threadIdx.x; b < nbins; b += 256) m = fmax(m, bin_thr[b]);
    for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, WAVE));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmax(fmax(s[0], s[1]), fmax(s[2], s[3]));
        reach[0] = m > 0.0 ? sqrt(m) * 1.000001 + 1e-6 : 0.0;
    }
}

// g = exp(S / n): the geometric mean of the pilot densities (compGML, tiled:3807-3814).
__global__ void k_gml(const double *S, int64_t n, double *g)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) g[0] = exp(S[0] / (double)n);
}

// ============================================================ reductions ====

// partial[k * nblk + b] = sum over the block's points of log(max(den, 1e-14f))
// (or of log(den) when clamp == 0), in FP64.  Sequential second pass.
__global__ __launch_bounds__(256) void k_sumlog_partial(int64_t n, const float *den, int clamp, double *partial)
{
    __shared__ double s[4];
    const int k = blockIdx.y;
    const float *d = den + (int64_t)k * n;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        float v = d[i];
        if (clamp) v = v < MKDE_EPS_DEN ? MKDE_EPS_DEN : v;
        acc += log((double)v);
    }
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, WAVE);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) partial[(int64_t)k * gridDim.x + blockIdx.x] = ((s[0] + s[1]) + s[2]) + s[3];
}

// One workgroup per candidate: thread t adds partials t, t + 256, ... in
// order, then a fixed-shape tree over the 256 thread sums (deterministic).
__global__ __launch_bounds__(256) void k_sum_final(int nblk, int K, const double *partial, double *out)
{
    __shared__ double s_w[256];
    const int k = blockIdx.x;
    const double *pk = partial + (int64_t)k * nblk;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += 256) s += pk[b];
    s_w[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if ((int)threadIdx.x < w) s_w[threadIdx.x] += s_w[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = s_w[0];
}

// Abramson update h_i = h0 * powf(den0_i / g, alpha) (tiled:571, kde_kernel_kdtr.cu:1007-1037),
// den0 in bin order, h written in original order.
__global__ void k_abramson(int64_t n, const int *perm, const float *den0_sorted, const double *g, float h0,
                           float alpha, float *h_orig)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ratio = (float)((double)den0_sorted[i] / g[0]);
    h_orig[perm[i]] = h0 * powf(ratio, alpha);
}

// scatter a bin-order array back to original order
__global__ void k_unsort_f32(int64_t n, const int *perm, const float *src, float *dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[perm[i]] = src[i];
}

// gather an original-order array into bin order
__global__ void k_sort_f32(int64_t n, const int *perm, const float *src, float *dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// coefk[k*n + i] = w * ew_k / (2 pi h2_k), bin order
// coef_out[i * stride]: one candidate's coefficient of every point (bin order)
__global__ void k_coef_k(int64_t n, const int *perm, const float *w_orig, const float *ew_k_orig, double h2k,
                         float *coef_out, int stride)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    coef_out[i * stride] = (float)((double)w_orig[o] * (double)ew_k_orig[o] / (h2k * (double)MKDE_TWO_PI_F));
}

// k_coef_k with the bandwidth read from the device: h2k = fl32(h*h).
__global__ void k_coef_k_dev(int64_t n, const int *perm, const float *w_orig, const float *ew_k_orig, const float *h_dev,
                             float *coef_out, int stride)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float h = *h_dev;
    const double h2k = (double)(float)(h * h);
    const int o = perm[i];
    coef_out[i * stride] = (float)((double)w_orig[o] * (double)ew_k_orig[o] / (h2k * (double)MKDE_TWO_PI_F));
}

__global__ void k_fill_f32(int64_t n, float v, float *dst)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = v;
}

// ====================================================== host launchers ====

extern "C++" {
namespace mkde_dev {

// ------------------------------------------------------------- scratch --
// Device memory of the allocation-free entry point
// (mkde_density_at_points_noalloc): the point kernels' work list, their FP64
// partial sums and the bin-order densities are carved from this file-scope
// array instead of buffers the library allocates.  Every device the library
// runs on holds one copy of it until the process ends; all calls on a device
// share it.  Host code never passes the array's name to a kernel (on the host
// it is a stand-in, not a device address): scratch_address() resolves the
// device address once per device, with that device current, and callers use
// that pointer.
static __device__ __align__(256) unsigned char scratch[MKDE_SCRATCH_BYTES];

hipError_t scratch_address(int device, void **out)
{
    static std::mutex mu;
    static void *addr[MKDE_MAX_DEVICES] = {};
    if (!out || device < 0 || device >= MKDE_MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lk(mu);
    if (!addr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device) return hipErrorInvalidDevice;   // the caller makes the device current
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        addr[device] = p;
    }
    *out = addr[device];
    return hipSuccess;
}

static inline unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
static inline unsigned nblk_(int64_t n) { return nblk(n, 256); }

hipError_t prep_points(hipStream_t s, int64_t n, const int *perm, const float *xs, const float *ys,
                       const float *w, const float *h, const float *ew, float h_fixed, PtRec *out)
{
    if (n > 0) hipLaunchKernelGGL(k_prep_points, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, xs, ys, w, h, ew,
                                  h_fixed, out);
    return hipGetLastError();
}

hipError_t bin_stats(hipStream_t s, int nbins, const int *bin_start, const PtRec *pts, BinInfo *box, double *thr)
{
    hipLaunchKernelGGL(k_bin_stats, dim3(nblk(nbins, 4)), dim3(256), 0, s, nbins, bin_start, pts, box, thr);
    return hipGetLastError();
}

// Density blocks of a single tile or of a tiled raster (tile-major numbering)
int64_t raster_blocks(const GridDesc &g, const Tiling &tl)
{
    if (tl.tw > 0) {
        const int64_t nty = (g.nrows + tl.th - 1) / tl.th;
        return (int64_t)tl.ntx * nty * tl.bpt;
    }
    return ((g.ncols + MKDE_DT_BW - 1) / MKDE_DT_BW) * ((g.nrows + 63) / 64);
}

static void scan_u32(hipStream_t s, int n, const unsigned *in, unsigned *out, unsigned *chunks, unsigned *total)
{
    const int nchunks = (n + 1023) / 1024;
    hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nchunks), dim3(256), 0, s, n, in, out, chunks);
    hipLaunchKernelGGL(k_scan_chunks, dim3(1), dim3(256), 0, s, nchunks, chunks, total);
    hipLaunchKernelGGL(k_scan_add, dim3(nblk_(n)), dim3(256), 0, s, n, out, chunks);
}

// Costs, split parts and the three scans; W.counts receives {items, slots, splits}.
hipError_t density_plan(hipStream_t s, const GridDesc &g, const float *mask, const PointIndex &idx, const Tiling &tl,
                        const DensityWork &W, const Tuning &T)
{
    const int nbx = (int)((g.ncols + MKDE_DT_BW - 1) / MKDE_DT_BW);
    const int nblk = (int)raster_blocks(g, tl);
    hipLaunchKernelGGL(k_block_cost, dim3((unsigned)nblk), dim3(256), 0, s, g, mask, idx, nbx, tl, W.cost);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, (int64_t)2, 0.0f, (float *)W.total);
    const unsigned nred = (unsigned)std::min<int64_t>(1024, ((int64_t)nblk + 255) / 256);
    hipLaunchKernelGGL(k_cost_total, dim3(nred), dim3(256), 0, s, nblk, W.cost, W.total);
    hipLaunchKernelGGL(k_parts, dim3(nblk_(nblk)), dim3(256), 0, s, nblk, W.cost, W.total, DT_TMIN, T.dt_target, W.parts,
                       W.item0, W.sidx);
    // in place is not allowed: parts -> item0 scan reads parts; sparts / sflag
    // were written to item0 / sidx and are scanned into slot0 / sidx2 below
    scan_u32(s, nblk, W.item0, W.slot0, W.chunks, W.counts + 1);
    scan_u32(s, nblk, W.sidx, W.sidx2, W.chunks, W.counts + 2);
    scan_u32(s, nblk, W.parts, W.item0, W.chunks, W.counts + 0);
    return hipGetLastError();
}

// n_dev non-null: the item count is read on the device (W.counts[0]) and
// n_items is only the grid bound (the most items the work list can have).
hipError_t density_items(hipStream_t s, int nblk, const DensityWork &W, int n_items, int4 *items, int4 *splits,
                         int *order, const unsigned *n_dev)
{
    hipLaunchKernelGGL(k_items_fill, dim3(nblk_(nblk)), dim3(256), 0, s, nblk, W.parts, W.item0, W.slot0, W.sidx2,
                       items, splits);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, (int64_t)32, 0.0f, (float *)W.hist);
    hipLaunchKernelGGL(k_item_hist, dim3(nblk_(n_items)), dim3(256), 0, s, n_items, n_dev, items, W.cost, W.hist);
    hipLaunchKernelGGL(k_bucket_scan, dim3(1), dim3(64), 0, s, W.hist);
    hipLaunchKernelGGL(k_item_order, dim3(nblk_(n_items)), dim3(256), 0, s, n_items, n_dev, items, W.cost, W.hist,
                       order);
    return hipGetLastError();
}

// order == nullptr: one item per block in raster order; else n_items items.
hipError_t density_tile(hipStream_t s, const GridDesc &g, const float *mask, float *out, const PointIndex &idx,
                        unsigned long long *stats, const int *order, const int4 *items, const unsigned *cost,
                        double *partials, int n_items, const Tiling &tl)
{
    const int nbx = (int)((g.ncols + MKDE_DT_BW - 1) / MKDE_DT_BW);
    TileArgs A{g, mask, out, idx, stats, order, items, cost, partials, nbx, tl};
    dim3 grid((unsigned)(order ? n_items : raster_blocks(g, tl)));
    if (stats) hipLaunchKernelGGL(k_density_tile<true>, grid, dim3(DT_NT), 0, s, A);
    else hipLaunchKernelGGL(k_density_tile<false>, grid, dim3(DT_NT), 0, s, A);
    return hipGetLastError();
}

hipError_t density_merge(hipStream_t s, const GridDesc &g, const Tiling &tl, const float *mask, int n_split,
                         const int4 *splits, const double *partials, float *out)
{
    if (n_split <= 0) return hipSuccess;
    const int nbx = (int)((g.ncols + MKDE_DT_BW - 1) / MKDE_DT_BW);
    hipLaunchKernelGGL(k_density_merge, dim3((unsigned)n_split), dim3(256), 0, s, g, tl, nbx, mask, splits, partials,
                       out);
    return hipGetLastError();
}

// Work list of the point kernels (64-query waves): costs, split parts (T =
// max(tmin, total / target)) and scans; W.counts = {items, slots, splits}.
// kind 0: density at points (h_fixed > 0: one bandwidth), 1: batched
// leave-one-out with thresholds thr_k[0..K-1] ascending.
hipError_t point_plan(hipStream_t s, const PointIndex &idx, int kind, float h_fixed, const double *thr_k, int K,
                      const DensityWork &W, const Tuning &T)
{
    const int nw = (int)((idx.n + 63) / 64);
    double lim_fixed = 0.0;
    const double *lim_dev = nullptr;
    if (kind == 1) {
        lim_dev = thr_k + (K - 1);
    } else if (h_fixed > 0.0f) {
        lim_fixed = 9.0 * (double)(float)(h_fixed * h_fixed);
    }
    hipLaunchKernelGGL(k_wave_cost, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, s, idx, lim_fixed, lim_dev, W.cost);
    hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, s, (int64_t)2, 0.0f, (float *)W.total);
    const unsigned nred = (unsigned)std::min<int64_t>(1024, ((int64_t)nw + 255) / 256);
    hipLaunchKernelGGL(k_cost_total, dim3(nred), dim3(256), 0, s, nw, W.cost, W.total);
    const unsigned tmin = (unsigned)MKDE_PTS_TMIN;
    const unsigned long long target = kind == 1 ? T.loo_target : T.pts_target;
    hipLaunchKernelGGL(k_parts, dim3(nblk_(nw)), dim3(256), 0, s, nw, W.cost, W.total, tmin, target, W.parts,
                       W.item0, W.sidx);
    scan_u32(s, nw, W.item0, W.slot0, W.chunks, W.counts + 1);
    scan_u32(s, nw, W.sidx, W.sidx2, W.chunks, W.counts + 2);
    scan_u32(s, nw, W.parts, W.item0, W.chunks, W.counts + 0);
    return hipGetLastError();
}

hipError_t point_density(hipStream_t s, const PointIndex &idx, int loo, float *den_sorted, float h_fixed,
                         const int *order, const int4 *items, const unsigned *cost, double *partials, int n_items)
{
    PtsArgs A{idx, den_sorted, loo, 0.0, 0.0, order, items, cost, partials, n_items};
    if (idx.n <= 0) return hipSuccess;
    const int64_t units = order ? n_items : (idx.n + 63) / 64;
    const unsigned grid = (unsigned)((units + 3) / 4);
    if (h_fixed > 0.0f) {
        const double h2 = (double)(float)(h_fixed * h_fixed);
        A.thr_fixed = 9.0 * h2;
        A.a_fixed = -MKDE_LOG2E / (2.0 * h2);
        if (loo) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_point_density_w<true, true>), dim3(grid), dim3(256), 0, s, A);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_point_density_w<true, false>), dim3(grid), dim3(256), 0, s, A);
    } else {
        if (loo) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_point_density_w<false, true>), dim3(grid), dim3(256), 0, s, A);
        else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_point_density_w<false, false>), dim3(grid), dim3(256), 0, s, A);
    }
    return hipGetLastError();
}

// n_items_dev non-null: the item count is read on the device and n_items is
// only the grid bound.
hipError_t loo_multi(hipStream_t s, const PointIndex &idx, const float *coefk, const double *a_k,
                     const double *thr_k, int K, float *den1, const int *order, const int4 *items,
                     const unsigned *cost, double *partials, int n_items,
                     const unsigned *n_items_dev)
{
    LooMultiArgs A{idx, coefk, a_k, thr_k, K, den1, order, items, cost, partials, n_items, n_items_dev};
    if (idx.n <= 0) return hipSuccess;
    const int64_t units = order ? n_items : (idx.n + 63) / 64;
    hipLaunchKernelGGL(k_loo_multi_w, dim3((unsigned)((units + 3) / 4)), dim3(256), 0, s, A);
    return hipGetLastError();
}

// n_split_dev non-null: the split count is read on the device and n_split is
// only the grid bound.
hipError_t point_merge(hipStream_t s, int64_t n, int m, int K, int n_split, const int4 *splits, const double *partials,
                       float *den, const unsigned *n_split_dev)
{
    if (n_split > 0)
        hipLaunchKernelGGL(k_point_merge, dim3((unsigned)n_split), dim3(256), 0, s, n, m, K, n_split_dev, splits,
                           partials, den);
    return hipGetLastError();
}

hipError_t range_count(hipStream_t s, const PointIndex &idx, int64_t *counts)
{
    RangeArgs A{idx, counts, nullptr, nullptr};
    if (idx.n > 0) hipLaunchKernelGGL(k_range, dim3(nblk(idx.n, 256)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t range_fill(hipStream_t s, const PointIndex &idx, const int64_t *offsets, int32_t *out_idx)
{
    RangeArgs A{idx, nullptr, offsets, out_idx};
    if (idx.n > 0) hipLaunchKernelGGL(k_range, dim3(nblk(idx.n, 256)), dim3(256), 0, s, A);
    return hipGetLastError();
}

hipError_t valid_cells(hipStream_t s, int64_t ncells, const float *mask, float nodata, uint8_t *valid)
{
    hipLaunchKernelGGL(k_valid, dim3(nblk(ncells, 256)), dim3(256), 0, s, ncells, mask, nodata, valid);
    return hipGetLastError();
}

// Valid-cell runs of every row of a mask: per-row counts, their exclusive scan
// (first[r]) and run_ptr; the total is left in *total for the caller to size
// the run list before runs_fill.  first, cnt: nrows each; chunks:
// (nrows + 1023) / 1024.
hipError_t runs_count(hipStream_t s, const GridDesc &g, const uint8_t *valid, unsigned *cnt, unsigned *first,
                      unsigned *chunks, unsigned *total, int64_t *ptr)
{
    hipLaunchKernelGGL(k_run_count, dim3(nblk(g.nrows, 4)), dim3(256), 0, s, g.ncols, g.nrows, valid, cnt);
    scan_u32(s, (int)g.nrows, cnt, first, chunks, total);
    hipLaunchKernelGGL(k_run_ptr, dim3(nblk_(g.nrows + 1)), dim3(256), 0, s, g.nrows, first, total, ptr);
    return hipGetLastError();
}

hipError_t runs_fill(hipStream_t s, const GridDesc &g, const uint8_t *valid, const unsigned *first, int2 *se)
{
    hipLaunchKernelGGL(k_run_fill, dim3(nblk(g.nrows, 4)), dim3(256), 0, s, g.ncols, g.nrows, valid, first, se);
    return hipGetLastError();
}

hipError_t boundary(hipStream_t s, const GridDesc &g, const uint8_t *valid, uint8_t *flags, float2 *blist,
                    unsigned long long *bcount)
{
    if (g.ncols <= 0 || g.nrows <= 0) return hipSuccess;
    const unsigned gy = (unsigned)std::min<int64_t>(g.nrows, 65535);
    hipLaunchKernelGGL(k_boundary, dim3(nblk(g.ncols, 256), gy), dim3(256), 0, s, g, valid, flags, blist, bcount);
    return hipGetLastError();
}

hipError_t bt_build(hipStream_t s, const GridDesc &g, const uint8_t *flags, uint4 *bits, float4 *box, const BndTree &T)
{
    const int64_t n0 = (int64_t)T.nx[0] * T.ny[0];
    hipLaunchKernelGGL(k_bt_leaf, dim3(nblk(n0, 256)), dim3(256), 0, s, g, flags, T.nx[0], T.ny[0], bits, box);
    for (int l = 1; l < T.nlev; ++l) {
        const int64_t nl = (int64_t)T.nx[l] * T.ny[l];
        hipLaunchKernelGGL(k_bt_up, dim3(nblk(nl, 256)), dim3(256), 0, s, T.nx[l - 1], T.ny[l - 1], box + T.off[l - 1],
                           T.nx[l], T.ny[l], box + T.off[l]);
    }
    return hipGetLastError();
}

hipError_t dist_tree(hipStream_t s, int64_t n, const float *xs, const float *ys, const int *perm, const GridDesc &g,
                     const BndTree &T, double *dist)
{
    if (n > 0) hipLaunchKernelGGL(k_dist_tree, dim3(nblk(n, BT_WG)), dim3(BT_WG), 0, s, n, xs, ys, perm, g, T, dist);
    return hipGetLastError();
}

hipError_t dist_boundary(hipStream_t s, int64_t n, const float *x, const float *y, const float2 *blist, int64_t nb,
                         double *dist)
{
    if (n > 0) hipLaunchKernelGGL(k_dist_boundary, dim3(nblk(n, 256)), dim3(256), 0, s, n, x, y, blist, nb, dist);
    return hipGetLastError();
}

// h_dev non-null: every point's bandwidth is the device value *h_dev.
hipError_t ec_active(hipStream_t s, int64_t n, const GridDesc &g, const float *x, const float *y, const double *dist,
                     const float *h, float h_fixed, const float *h_dev, float *ew, int *list, int *count,
                     int cw_narrow, int cw_wide)
{
    if (n > 0)
        hipLaunchKernelGGL(k_ec_active, dim3(nblk(n, 256)), dim3(256), 0, s, n, g, x, y, dist, h, h_fixed, h_dev, ew,
                           list, count, cw_narrow, cw_wide);
    return hipGetLastError();
}

hipError_t centres(hipStream_t s, const GridDesc &g, float *colx, float *rowy)
{
    const int64_t m = g.ncols > g.nrows ? g.ncols : g.nrows;
    hipLaunchKernelGGL(k_centres, dim3(nblk(m, 256)), dim3(256), 0, s, g, colx, rowy);
    return hipGetLastError();
}

// Widest box k_ec_wide takes, and its LDS: per column the FP64 prefix sum
// (8 B), at most 64,000 B, and up to EC_DXT_MAX columns also the point's dx
// (4 B more per column).  The dx table saves the rows' FP64 cell centres
// (~15% of a launch up to ~2,900 columns); wider boxes keep five workgroups
// per CU without it (measured slower with it from ~3,200 columns).
#define EC_DXT_MAX 2900
int ec_wide_max() { return 8000; }
static size_t ec_wide_lds(int w)
{
    return (size_t)w * (w <= EC_DXT_MAX ? sizeof(double) + sizeof(float) : sizeof(double));
}
static void launch_ec_wide(hipStream_t s, unsigned grid, int w, const EcArgs &A, int64_t n)
{
    if (w <= EC_DXT_MAX) hipLaunchKernelGGL(k_ec_wide<true>, dim3(grid), dim3(256), ec_wide_lds(w), s, A, n, w);
    else hipLaunchKernelGGL(k_ec_wide<false>, dim3(grid), dim3(256), ec_wide_lds(w), s, A, n, w);
}

// Most columns of the NARROW box of a point at the fixed bandwidth h on grid
// g, for every point within 3h of a valid cell (narrow_box: the box spans
// 6h / cs columns between two truncations, so at most floor(6h / cs) + 2;
// one more for the rounding of the double expressions).  A point whose box is
// wider (only one lying more than 3h left of the raster, whose box the size_t
// arithmetic stretches to the whole row) goes to the chunked k_ec.
int ec_box_cols(const GridDesc &g, float h)
{
    const double w = std::floor(6.0 * (double)h / g.cs + 1e-6) + 3.0;
    return (int)std::min<double>((double)g.ncols, std::max(w, 1.0));
}

static long long env_ll(const char *name, long long dflt)
{
    const char *e = getenv(name);
    const long long v = e ? atoll(e) : 0;
    return v > 0 ? v : dflt;
}

// Default launch shapes (measured on one MI355X), each overridable from the
// environment:
//   ec_chunk 1024 (MKDE_EC_CHUNK; 512 / 1024 / 2048: the adaptive job's EC
//     within 1%, the 64-candidate search 314 / 296 / 307 ms);
//   ec_wide_min = ec_chunk (MKDE_EC_WIDE_MIN): every box k_ec would cut into
//     two or more chunks runs in k_ec_wide;
//   loo_target 32768 (MKDE_LOO_TARGET; 2048 / 8192 / 32768 / 65536 / 131072:
//     search 237 / 200 / 187 / 191 / 194 ms);
//   pts_target 16384 (MKDE_PTS_TARGET; 4096 / 16384 / 65536: bandwidth phase
//     of the adaptive 1M job 8.2 / 7.5 / 8.0 ms);
//   dt_target 4096 (MKDE_DENSITY_TARGET; adaptive 1M job / 10M-point tiles,
//     density ms: 1024: 30.9 / 32.7, 2048: 30.9 / 27.5, 4096: 31.0 / 27.8,
//     16384: 31.7 / 27.9, 65536: 33.3 / 29.2).
Tuning default_tuning()
{
    Tuning t;
    const long long cw = env_ll("MKDE_EC_CHUNK", 1024);
    t.ec_chunk = (cw == 512 || cw == 2048) ? (int)cw : 1024;
    t.ec_wide_min = (int)env_ll("MKDE_EC_WIDE_MIN", t.ec_chunk);
    t.loo_target = (unsigned long long)env_ll("MKDE_LOO_TARGET", 32768);
    t.pts_target = (unsigned long long)env_ll("MKDE_PTS_TARGET", 16384);
    t.dt_target = (unsigned long long)env_ll("MKDE_DENSITY_TARGET", (long long)DT_TARGET);
    return t;
}

// The n_narrow points at the front of list run in k_ec (one wave each, column
// chunks), the n_wide points at the back in k_ec_wide (one workgroup each,
// max_w columns of LDS).
hipError_t ec(hipStream_t s, int64_t n, int n_narrow, int n_wide, int max_w, const GridDesc &g, const MaskRuns &runs,
              const float *colx, const float *rowy, const float *x, const float *y, const float *h, float h_fixed,
              const int *list, float *ew, unsigned long long *n_eval, int cw, int cx32)
{
    EcArgs A{g, runs, colx, rowy, x, y, h, h_fixed, nullptr, nullptr, list, ew, n_eval, cx32};
    if (n_narrow > 0) {
        if (cw == 512) hipLaunchKernelGGL(k_ec<512>, dim3((unsigned)n_narrow), dim3(64), 0, s, A);
        else if (cw == 1024) hipLaunchKernelGGL(k_ec<1024>, dim3((unsigned)n_narrow), dim3(64), 0, s, A);
        else hipLaunchKernelGGL(k_ec<2048>, dim3((unsigned)n_narrow), dim3(64), 0, s, A);
    }
    if (n_wide > 0) {
        if (max_w < 1 || max_w > ec_wide_max()) return hipErrorInvalidValue;
        launch_ec_wide(s, (unsigned)n_wide, max_w, A, n);
    }
    return hipGetLastError();
}

// The same EC pass with every size on the device (no host read-back, so the
// launches can be captured in a graph): the bandwidth is *h_dev, the list
// lengths count[0] (narrow) and count[1] (wide) as k_ec_active left them.
// Fixed grids loop over the list.  w_wide is the cw_wide k_ec_active was given
// (the widest wide box): the wide kernel takes that much LDS, as many
// workgroups as fit on 256 CUs, and is left out when no box can be wide
// (w_wide <= cw_narrow).
hipError_t ec_dev(hipStream_t s, int64_t n, const GridDesc &g, const MaskRuns &runs, const float *colx,
                  const float *rowy, const float *x, const float *y, const float *h_dev, const int *count,
                  const int *list, float *ew, int cw, int cw_narrow, int w_wide, int cx32)
{
    if (n <= 0) return hipSuccess;
    if (w_wide > ec_wide_max()) return hipErrorInvalidValue;
    EcArgs A{g, runs, colx, rowy, x, y, nullptr, 0.0f, h_dev, count, list, ew, nullptr, cx32};
    const unsigned gn = (unsigned)std::min<int64_t>(n, 256 * 32);   // 8 waves per SIMD
    if (cw == 512) hipLaunchKernelGGL(k_ec<512>, dim3(gn), dim3(64), 0, s, A);
    else if (cw == 1024) hipLaunchKernelGGL(k_ec<1024>, dim3(gn), dim3(64), 0, s, A);
    else hipLaunchKernelGGL(k_ec<2048>, dim3(gn), dim3(64), 0, s, A);
    if (w_wide > cw_narrow) {
        const size_t lds = ec_wide_lds(w_wide);
        const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, (160u << 10) / (lds + 256)));
        const unsigned gw = (unsigned)std::min<int64_t>(n, 256 * per_cu);
        launch_ec_wide(s, gw, w_wide, A, n);
    }
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fold_slots(unsigned long long *slots)
{
    __shared__ unsigned long long s_w[4];
    unsigned long long v = 0;
    for (int k = threadIdx.x; k < MKDE_CNT_SLOTS; k += 256) v += slots[k * MKDE_CNT_STRIDE];
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) slots[MKDE_CNT_SLOTS * MKDE_CNT_STRIDE] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Words of a slot counter (slots, then the folded total).
int64_t slot_words() { return (int64_t)MKDE_CNT_SLOTS * MKDE_CNT_STRIDE + 1; }

// slots[slot_words() - 1] = sum of the slots.
hipError_t fold_slots(hipStream_t s, unsigned long long *slots)
{
    hipLaunchKernelGGL(k_fold_slots, dim3(1), dim3(256), 0, s, slots);
    return hipGetLastError();
}

hipError_t count_pairs(hipStream_t s, const GridDesc &g, const MaskRuns &runs, const PointIndex &idx,
                       unsigned long long *total)
{
    CountArgs A{g, runs, idx, total};
    if (idx.n > 0) hipLaunchKernelGGL(k_count_pairs, dim3((unsigned)idx.n), dim3(64), 0, s, A);
    return hipGetLastError();
}

hipError_t reach(hipStream_t s, int nbins, const double *bin_thr, double *reach_out)
{
    hipLaunchKernelGGL(k_reach, dim3(1), dim3(256), 0, s, nbins, bin_thr, reach_out);
    return hipGetLastError();
}

hipError_t gml(hipStream_t s, const double *S, int64_t n, double *g)
{
    hipLaunchKernelGGL(k_gml, dim3(1), dim3(64), 0, s, S, n, g);
    return hipGetLastError();
}

hipError_t sumlog(hipStream_t s, int64_t n, int K, const float *den, int clamp, double *partial, int nb, double *out)
{
    hipLaunchKernelGGL(k_sumlog_partial, dim3((unsigned)nb, (unsigned)K), dim3(256), 0, s, n, den, clamp, partial);
    hipLaunchKernelGGL(k_sum_final, dim3((unsigned)K), dim3(256), 0, s, nb, K, partial, out);
    return hipGetLastError();
}

hipError_t abramson(hipStream_t s, int64_t n, const int *perm, const float *den0_sorted, const double *g, float h0,
                    float alpha, float *h_orig)
{
    if (n > 0) hipLaunchKernelGGL(k_abramson, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, den0_sorted, g, h0,
                                  alpha, h_orig);
    return hipGetLastError();
}

hipError_t unsort_f32(hipStream_t s, int64_t n, const int *perm, const float *src, float *dst)
{
    if (n > 0) hipLaunchKernelGGL(k_unsort_f32, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, src, dst);
    return hipGetLastError();
}

hipError_t sort_f32(hipStream_t s, int64_t n, const int *perm, const float *src, float *dst)
{
    if (n > 0) hipLaunchKernelGGL(k_sort_f32, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, src, dst);
    return hipGetLastError();
}

hipError_t coef_k(hipStream_t s, int64_t n, const int *perm, const float *w, const float *ew_k, double h2k,
                  float *coef_out, int stride)
{
    if (n > 0)
        hipLaunchKernelGGL(k_coef_k, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, w, ew_k, h2k, coef_out, stride);
    return hipGetLastError();
}

hipError_t coef_k_dev(hipStream_t s, int64_t n, const int *perm, const float *w, const float *ew_k, const float *h_dev,
                      float *coef_out, int stride)
{
    if (n > 0)
        hipLaunchKernelGGL(k_coef_k_dev, dim3(nblk(n, 256)), dim3(256), 0, s, n, perm, w, ew_k, h_dev, coef_out, stride);
    return hipGetLastError();
}

hipError_t fill_f32(hipStream_t s, int64_t n, float v, float *dst)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_f32, dim3(nblk(n, 256)), dim3(256), 0, s, n, v, dst);
    return hipGetLastError();
}

}  // namespace mkde_dev
}
