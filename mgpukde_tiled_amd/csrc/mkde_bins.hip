// mkde_bins.hip -- the point index's bin order built on the device.
//
// The reference orders points by a host k-d tree (KDtree::Create / Split,
// KDtree.cu:53-154, called from BuildCPUKDtree tiled:4241-4258, flattened by
// BuildGPUKDtree tiled:4260-4275); this library bins them on a
// uniform grid (BinGrid, sized on the host from the points' bounding box) and
// needs, per upload: each point's bin, the bin order (bin-major, the caller's
// order within a bin), the first position of every bin, and the coordinates
// in bin order.  A stable radix sort of the bin numbers (rocPRIM) gives the
// same permutation as a stable counting sort: ties keep the caller's order,
// so every later summation order is fixed by the input alone.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <rocprim/device/device_radix_sort.hpp>

#include "mkde_device.h"

// Bin of point i: the grid cell of its coordinates, with the arithmetic of the
// query side (bin_range): floor of the FP64 offset over the bin size, clamped.
__global__ __launch_bounds__(256) void k_bin_key(int64_t n, const float *x, const float *y, BinGrid bg,
                                                 unsigned *key, int *idx)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int bx = (int)floor(((double)x[i] - bg.x0) / bg.size);
    int by = (int)floor(((double)y[i] - bg.y0) / bg.size);
    bx = bx < 0 ? 0 : (bx > bg.nbx - 1 ? bg.nbx - 1 : bx);
    by = by < 0 ? 0 : (by > bg.nby - 1 ? bg.nby - 1 : by);
    key[i] = (unsigned)(by * bg.nbx + bx);
    idx[i] = (int)i;
}

// First position of every bin from the sorted bin numbers: position p (0..n)
// opens the bins after its predecessor's, up to and including its own, so
// each of the nbins + 1 entries is written once and start[nbins] = n.
__global__ __launch_bounds__(256) void k_bin_start(int64_t n, int nbins, const unsigned *skey, int *start)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p > n) return;
    const int lo = p == 0 ? 0 : (int)skey[p - 1] + 1;
    const int hi = p == n ? nbins : (int)skey[p];
    for (int b = lo; b <= hi; ++b) start[b] = (int)p;
}

__global__ __launch_bounds__(256) void k_gather_xy(int64_t n, const int *perm, const float *x, const float *y,
                                                   float *xs, float *ys)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int i = perm[p];
    xs[p] = x[i];
    ys[p] = y[i];
}

namespace mkde_dev {

static int key_bits(int nbins)
{
    int b = 1;
    while (b < 31 && (1 << b) < nbins) ++b;
    return b;
}

// Temporary storage of the sort for n points in nbins bins.
hipError_t bin_sort_bytes(int64_t n, int nbins, size_t *bytes)
{
    *bytes = 0;
    return rocprim::radix_sort_pairs(nullptr, *bytes, (const unsigned *)nullptr, (unsigned *)nullptr,
                                     (const int *)nullptr, (int *)nullptr, (size_t)n, 0, key_bits(nbins));
}

// Bin order of n points: key / skey / idx are n-element scratch, perm gets the
// caller's index of each bin-order position, start the nbins + 1 bin starts,
// xs / ys the coordinates in bin order.  tmp holds tmp_bytes from
// bin_sort_bytes.
hipError_t bin_points(hipStream_t s, int64_t n, const float *x, const float *y, const BinGrid &bg, int nbins,
                      unsigned *key, unsigned *skey, int *idx, int *perm, void *tmp, size_t tmp_bytes, float *xs,
                      float *ys, int *start)
{
    const unsigned nb = (unsigned)((n + 255) / 256), nb1 = (unsigned)((n + 1 + 255) / 256);
    if (n > 0) {
        hipLaunchKernelGGL(k_bin_key, dim3(nb), dim3(256), 0, s, n, x, y, bg, key, idx);
        size_t bytes = tmp_bytes;
        hipError_t e = rocprim::radix_sort_pairs(tmp, bytes, (const unsigned *)key, skey, (const int *)idx, perm,
                                                 (size_t)n, 0, key_bits(nbins), s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_gather_xy, dim3(nb), dim3(256), 0, s, n, (const int *)perm, x, y, xs, ys);
    }
    hipLaunchKernelGGL(k_bin_start, dim3(nb1), dim3(256), 0, s, n, nbins, (const unsigned *)skey, start);
    return hipGetLastError();
}

}  // namespace mkde_dev
