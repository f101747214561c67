// Test-only driver around the REFERENCE's own KD-tree traversal.
//
// The reference headers gpu_cpu_shared/{kd_tree.cuh, kd_tree_common.cuh,
// misc_common.cuh, small_vector.cuh, defines_common.cuh, gpu_grid.cuh} are
// header-only C++ that also compiles for the host (DEVICE expands to nothing
// without __CUDACC__, gpu_cpu_shared/defines_common.cuh:4-8).  oracle/Makefile
// compiles this file against those headers where they lie under
// /root/reference and writes the result to oracle/_ref/ only; no reference
// source is copied into this repository.  The library exposes
// KDTree<GPUFemGrid2D>::findElement (gpu_cpu_shared/kd_tree.cuh:40-141) and
// isPointInTriagle (gpu_cpu_shared/kd_tree_common.cuh:18-61) so tests can
// pin the oracle's restatement bit-for-bit.
//
// The standard headers below are the ones the reference normally receives
// from its precompiled header (CMakeLists.txt:162-197).
#include <cassert>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <utility>

#include "gpu_grid.cuh"
#include "kd_tree.cuh"
#include "kd_tree_common.cuh"

extern "C" {

int ref_point_in_triangle(float px, float py, float ax, float ay, float bx,
                          float by, float cx, float cy, float *xi,
                          float *eta) {
    return NSFem::isPointInTriagle(NSFem::Point2D(px, py), NSFem::Point2D(ax, ay),
                                   NSFem::Point2D(bx, by), NSFem::Point2D(cx, cy),
                                   *xi, *eta)
               ? 1
               : 0;
}

// One findElement call per query point.  nodes: 2 words per KDNode (the
// struct is {unsigned flags; union{float; unsigned}}, 8 bytes).
void ref_kd_find_many(const uint32_t *nodes, const int32_t *leaf,
                      const float *xy, const int32_t *elems, int node_count,
                      const float *bbox, int n, const float *pts,
                      int32_t *element, float *xi, float *eta,
                      int32_t *nearest) {
    static_assert(sizeof(NSFem::KDNode) == 8, "KDNode layout");
    GPUSimulation::GPUFemGrid2D grid(xy, elems, node_count);
    NSFem::BBox2D box(NSFem::Point2D(bbox[0], bbox[1]),
                      NSFem::Point2D(bbox[2], bbox[3]));
    NSFem::KDTree<GPUSimulation::GPUFemGrid2D> tree(
        box, reinterpret_cast<const NSFem::KDNode *>(nodes), leaf, &grid);
    for (int i = 0; i < n; ++i) {
        float a = 0.0f, b = 0.0f;
        int nn = -1;
        const int e =
            tree.findElement(NSFem::Point2D(pts[2 * i], pts[2 * i + 1]), a, b, nn);
        element[i] = e;
        xi[i] = e >= 0 ? a : 0.0f;
        eta[i] = e >= 0 ? b : 0.0f;
        nearest[i] = e >= 0 ? -1 : nn;
    }
}
}
