// ref_kdtree_harness.cpp -- a thin C-ABI harness around the reference's own
// k-d tree (KDtree.h / KDtree.cu, compiled from where they lie in the reference
// tree by oracle/Makefile into oracle/_ref/).  TEST INFRASTRUCTURE ONLY: the
// tests use it to check that the oracle's brute-force candidate predicate
// (orc_range_candidates) yields exactly the sets the reference's range search
// returns.  Nothing in the product links it.
//
// Mirrors BuildCPUKDtree (tiled:4241-4258: max_level = (int)(log2(N)/2) + 1)
// and the CPU branch of LogLikelihood (tiled:3309-3334: range = 9 * h2 with
// h2 = h*h of a float h; SearchRange per query).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "KDtree.h"

extern "C" {

// For every query i (a data point, as the reference queries its own points)
// run SearchRange with range 9*fl32(h_i*h_i).  Writes per-query counts into
// counts[n]; when idx != nullptr also writes the candidate indexes, sorted
// ascending per query, at offsets[i] (offsets has n+1 entries, prefilled by a
// counting call).  Returns the total number of candidates.
int64_t ref_kdtree_range(int n, const float *x, const float *y, const float *h, int64_t *counts,
                         const int64_t *offsets, int32_t *idx)
{
    std::vector<Point> pts(n);
    for (int i = 0; i < n; i++) {
        pts[i].coords[0] = x[i];
        pts[i].coords[1] = y[i];
    }
    KDtree tree;
    int max_level = (int)(std::log((double)pts.size()) / std::log(2.0) / 2) + 1;
    tree.Create(pts, max_level);
    int64_t total = 0;
    std::vector<int> ri;
    std::vector<double> rd;
    for (int i = 0; i < n; i++) {
        double h2 = h[i] * h[i];
        double range = 9.0 * h2;
        Point q;
        q.coords[0] = x[i];
        q.coords[1] = y[i];
        ri.clear();
        rd.clear();
        tree.SearchRange(q, range, ri, rd);
        counts[i] = (int64_t)ri.size();
        total += (int64_t)ri.size();
        if (idx) {
            std::sort(ri.begin(), ri.end());
            for (size_t k = 0; k < ri.size(); k++) idx[offsets[i] + (int64_t)k] = ri[k];
        }
    }
    return total;
}

int ref_kdtree_num_nodes(int n, const float *x, const float *y)
{
    std::vector<Point> pts(n);
    for (int i = 0; i < n; i++) {
        pts[i].coords[0] = x[i];
        pts[i].coords[1] = y[i];
    }
    KDtree tree;
    int max_level = (int)(std::log((double)pts.size()) / std::log(2.0) / 2) + 1;
    tree.Create(pts, max_level);
    return tree.GetNumNodes();
}
}
