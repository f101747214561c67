// Host KD-tree build in the reference's layout (KDTreeBuilder::build,
// cpp/kd_tree_builder.cpp:81-148; KDNode, gpu_cpu_shared/kd_tree_common.cuh:
// 63-154).  The tree must be the reference's tree node for node, since the
// element a point is located in depends on the leaf order:
//   * the root box is FemGrid2D's bbox, whose max is the LAST node's
//     coordinates (BBox2D::expand, gpu_cpu_shared/misc_common.cuh:59-65);
//   * split = (min[axis] + max[axis]) * 0.5f in float, axes x, y, x, ...;
//   * a triangle goes left if any vertex is <= split and right if any vertex
//     is >= split (both when it touches the split);
//   * leaf when depth >= maxDepth or <= minLeafSize triangles;
//   * nodes are numbered depth-first, left subtree first, so the right child
//     index is the next free index after the left subtree.
// "Any vertex <= split" is evaluated as min(vertex coords) <= split on
// precomputed per-triangle extents, which is the same predicate.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

#include "nsfem_internal.hpp"

namespace nsfem {
namespace {

struct Builder {
    const std::vector<float> *tmin[2];
    const std::vector<float> *tmax[2];
    int max_depth, min_leaf;
    std::vector<uint32_t> *nodes;
    std::vector<int32_t> *leaf;
};

inline uint32_t f2u(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// Returns the index after the subtree (the next free node index).
int64_t build_rec(const Builder &B, std::vector<int32_t> &idx, float mn0, float mn1, float mx0,
                  float mx1, int axis, int level) {
    std::vector<uint32_t> &nodes = *B.nodes;
    if (level >= B.max_depth || idx.size() <= (size_t)B.min_leaf) {
        const size_t off = B.leaf->size();
        B.leaf->insert(B.leaf->end(), idx.begin(), idx.end());
        nodes.push_back(2u | ((uint32_t)idx.size() << 2));
        nodes.push_back((uint32_t)off);
        return (int64_t)nodes.size() / 2;
    }
    const float mn[2] = {mn0, mn1}, mx[2] = {mx0, mx1};
    const float split = (mn[axis] + mx[axis]) * 0.5f;
    const std::vector<float> &lo = *B.tmin[axis];
    const std::vector<float> &hi = *B.tmax[axis];
    std::vector<int32_t> li, ri;
    li.reserve(idx.size() / 2 + 16);
    ri.reserve(idx.size() / 2 + 16);
    for (const int32_t e : idx) {
        if (lo[e] <= split) li.push_back(e);
        if (hi[e] >= split) ri.push_back(e);
    }
    std::vector<int32_t>().swap(idx);  // release the parent's list early
    const int64_t cur = (int64_t)nodes.size() / 2;
    nodes.push_back(0u);
    nodes.push_back(0u);
    const int na = (axis + 1) % 2;
    float lmx[2] = {mx[0], mx[1]};
    lmx[axis] = split;
    float rmn[2] = {mn[0], mn[1]};
    rmn[axis] = split;
    const int64_t right = build_rec(B, li, mn[0], mn[1], lmx[0], lmx[1], na, level + 1);
    nodes[2 * cur] = (uint32_t)axis | ((uint32_t)right << 2);
    nodes[2 * cur + 1] = f2u(split);
    return build_rec(B, ri, rmn[0], rmn[1], mx[0], mx[1], na, level + 1);
}

}  // namespace

void kd_build_host(int node_count, const float *xy, int elem_count, const int32_t *elems,
                   int max_depth, int min_leaf, HostKdTree &out) {
    NSFEM_REQUIRE(node_count > 0 && elem_count > 0, NSFEM_INVALID_ARGUMENT,
                  "KD tree needs nodes and elements");
    for (int64_t k = 0; k < 6ll * elem_count; ++k)
        NSFEM_REQUIRE(elems[k] >= 0 && elems[k] < node_count, NSFEM_INVALID_ARGUMENT,
                      "element node index out of range");
    if (max_depth > 29) max_depth = 29;
    if (max_depth < 0) {
        const double dd = std::round(8.0 + (double)1.3f * std::log((double)elem_count));
        max_depth = std::min(29, (int)dd);
    }
    if (min_leaf < 0) min_leaf = 16;
    // FemGrid2D bbox with the expand() defect (cpp/grid.cpp:30)
    float mnx = FLT_MAX, mny = FLT_MAX;
    for (int i = 0; i < node_count; ++i) {
        mnx = xy[2 * i] < mnx ? xy[2 * i] : mnx;
        mny = xy[2 * i + 1] < mny ? xy[2 * i + 1] : mny;
    }
    const float mxx = xy[2 * (node_count - 1)], mxy = xy[2 * (node_count - 1) + 1];
    std::vector<float> tmin[2], tmax[2];
    for (int a = 0; a < 2; ++a) {
        tmin[a].resize(elem_count);
        tmax[a].resize(elem_count);
    }
    for (int e = 0; e < elem_count; ++e) {
        for (int a = 0; a < 2; ++a) {
            const float p = xy[2 * elems[6 * e] + a], q = xy[2 * elems[6 * e + 1] + a],
                        r = xy[2 * elems[6 * e + 2] + a];
            tmin[a][e] = std::min(p, std::min(q, r));
            tmax[a][e] = std::max(p, std::max(q, r));
        }
    }
    out.nodes.clear();
    out.leaf.clear();
    out.nodes.reserve((size_t)elem_count / 4);
    out.leaf.reserve((size_t)elem_count * 3);
    Builder B{{&tmin[0], &tmin[1]}, {&tmax[0], &tmax[1]}, max_depth, min_leaf, &out.nodes,
              &out.leaf};
    std::vector<int32_t> idx(elem_count);
    for (int i = 0; i < elem_count; ++i) idx[i] = i;
    build_rec(B, idx, mnx, mny, mxx, mxy, 0, 0);
    out.bbox[0] = mnx;
    out.bbox[1] = mny;
    out.bbox[2] = mxx;
    out.bbox[3] = mxy;
    out.max_depth = max_depth;
    NSFEM_REQUIRE(out.nodes.size() / 2 < (1u << 30), NSFEM_INVALID_ARGUMENT,
                  "KD tree too large for 30-bit child indices");
}

}  // namespace nsfem
