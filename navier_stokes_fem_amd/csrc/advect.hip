// Semi-Lagrangian advection with KD-tree point location for gfx950.
// Replaces gpu/advection.cu (advect :14-57, P2ShapeEval :5-12) and the
// traversal of gpu_cpu_shared/kd_tree.cuh (findElement :40-141,
// nearestNeghbourProcessLeaf :149-169, isPointInTriagle
// kd_tree_common.cuh:18-61).
//
// Bit-exactness: every float operation is written in the reference's
// order and this file is compiled without FMA contraction (the pragma
// below, and -ffp-contract=off on the command line), so the start point,
// the barycentric test, xi/eta and the P2 interpolation round exactly as
// the reference's IEEE float code does.  Leaf offsets are read as exact
// integers (the reference converts them through float,
// kd_tree_common.cuh:148-150, which only differs for offsets >= 2^24).
//
// Layout in HBM: tree nodes as 8-byte pairs (flags, split|offset) in the
// reference's KDNode layout, leaf triangle indices, node coordinates as
// float2, elements as 6 ints, plus tri_xy: the 3 vertex coordinates of each
// element (24 bytes) so that a leaf test reads one contiguous record
// instead of three indices and three dependent coordinate gathers.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cfloat>

#include "nsfem_internal.hpp"

namespace nsfem {
namespace {

struct TreeView {
    const uint2 *nodes;
    const int32_t *leaf;
    const float2 *xy;
    const int32_t *elem;
    const float2 *tri;  // 3 float2 per element
};

__device__ __forceinline__ bool point_in_triangle(float px, float py, float2 A, float2 B,
                                                  float2 Cc, float &xi, float &eta) {
    const float abx = B.x - A.x, aby = B.y - A.y;
    const float acx = Cc.x - A.x, acy = Cc.y - A.y;
    const float apx = px - A.x, apy = py - A.y;
    const float D = (abx * acy) - (aby * acx);
    const float dsign = D > 0 ? 1.0f : -1.0f;
    const float b1 = (apx * acy - apy * acx) * dsign;
    if (b1 < 0 || b1 > D) return false;
    const float b2 = (apy * abx - apx * aby) * dsign;
    if (b2 < 0 || b1 > D - b2) return false;
    const float dinv = 1.0f / fabsf(D);
    xi = b1 * dinv;
    eta = b2 * dinv;
    return true;
}

// Nearest-node search by the reference's stack unwinding
// (kd_tree.cuh:95-139).  The stack state after the initial descent and leaf
// scan is rebuilt by descending again; entries pack node | (count << 30).
__device__ __noinline__ int nn_search(const TreeView T, float px, float py) {
    unsigned stack[32];
    int sp = 0;
    int cur = 0;
    uint2 nd = T.nodes[0];
    while ((nd.x & 3u) != 2u) {
        const float pa = (nd.x & 3u) ? py : px;
        const float s = __uint_as_float(nd.y);
        stack[sp++] = (unsigned)cur | (1u << 30);
        cur = pa <= s ? cur + 1 : (int)(nd.x >> 2);
        nd = T.nodes[cur];
    }
    stack[sp++] = (unsigned)cur;
    float min_d = FLT_MAX;
    int closest = -1;
    while (sp > 0) {
        const int top = (int)(stack[sp - 1] & 0x3fffffffu);
        const uint2 tn = T.nodes[top];
        if ((tn.x & 3u) == 2u) {
            const int off = (int)tn.y, cnt = (int)(tn.x >> 2);
            for (int i = 0; i < cnt; ++i) {
                const int e = T.leaf[off + i];
                for (int j = 0; j < 6; ++j) {
                    const int ni = T.elem[6 * e + j];
                    const float2 o = T.xy[ni];
                    const float d = (px - o.x) * (px - o.x) + (py - o.y) * (py - o.y);
                    if (d < min_d) {
                        min_d = d;
                        closest = ni;
                    }
                }
            }
            --sp;
        } else {
            while (sp > 0 && (stack[sp - 1] >> 30) == 2u) --sp;
            if (sp == 0) break;
            const int act = (int)(stack[sp - 1] & 0x3fffffffu);
            const unsigned cnt = stack[sp - 1] >> 30;
            const uint2 an = T.nodes[act];
            const float pa = (an.x & 3u) ? py : px;
            const float s = __uint_as_float(an.y);
            if (cnt == 0u) {
                const int next = pa <= s ? act + 1 : (int)(an.x >> 2);
                stack[sp - 1] = (unsigned)act | (1u << 30);
                stack[sp++] = (unsigned)next;
            } else {
                const float dd = pa - s;
                const float d2 = dd * dd;
                if (d2 < min_d) {
                    const int far = pa <= s ? (int)(an.x >> 2) : act + 1;
                    stack[sp - 1] = (unsigned)act | (2u << 30);
                    stack[sp++] = (unsigned)far;
                } else {
                    --sp;
                }
            }
        }
    }
    return closest;
}

// findElement: descent (left on <=), first containing triangle in leaf
// order, else nearest node.
__device__ __forceinline__ int find_element(const TreeView &T, float px, float py, float &xi,
                                            float &eta, int &nearest) {
    nearest = -1;
    if (isnan(px) || isnan(py)) return -1;
    int cur = 0;
    uint2 nd = T.nodes[0];
    while ((nd.x & 3u) != 2u) {
        const float pa = (nd.x & 3u) ? py : px;
        const float s = __uint_as_float(nd.y);
        cur = pa <= s ? cur + 1 : (int)(nd.x >> 2);
        nd = T.nodes[cur];
    }
    const int off = (int)nd.y, cnt = (int)(nd.x >> 2);
    for (int i = 0; i < cnt; ++i) {
        const int e = T.leaf[off + i];
        const float2 a = T.tri[3 * e], b = T.tri[3 * e + 1], c = T.tri[3 * e + 2];
        if (point_in_triangle(px, py, a, b, c, xi, eta)) return e;
    }
    nearest = nn_search(T, px, py);
    return -1;
}

__device__ __forceinline__ void p2_shape(float xi, float eta, float (&o)[6]) {
    o[0] = 1 - 3 * xi - 3 * eta + 2 * xi * xi + 4 * xi * eta + 2 * eta * eta;
    o[1] = 2 * xi * xi - xi;
    o[2] = 2 * eta * eta - eta;
    o[3] = 4 * xi * eta;
    o[4] = 4 * eta - 4 * xi * eta - 4 * eta * eta;
    o[5] = 4 * xi - 4 * xi * xi - 4 * xi * eta;
}

__device__ __forceinline__ void p2_shape(double xi, double eta, double (&o)[6]) {
    o[0] = 1 - 3 * xi - 3 * eta + 2 * xi * xi + 4 * xi * eta + 2 * eta * eta;
    o[1] = 2 * xi * xi - xi;
    o[2] = 2 * eta * eta - eta;
    o[3] = 4 * xi * eta;
    o[4] = 4 * eta - 4 * xi * eta - 4 * eta * eta;
    o[5] = 4 * xi - 4 * xi * xi - 4 * xi * eta;
}

__global__ __launch_bounds__(256) void k_advect_f32(TreeView T, int n, const float *__restrict__ u,
                                                    const float *__restrict__ v, float dt,
                                                    float *__restrict__ uo,
                                                    float *__restrict__ vo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 pos = T.xy[i];
    const float ui = u[i], vi = v[i];
    const float sx = pos.x - ui * dt;
    const float sy = pos.y - vi * dt;
    float xi = 0.f, eta = 0.f;
    int nn = -1;
    const int e = find_element(T, sx, sy, xi, eta, nn);
    if (e > -1) {
        float c[6];
        p2_shape(xi, eta, c);
        float ur = 0, vr = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int idx = T.elem[6 * e + k];
            ur += c[k] * u[idx];
            vr += c[k] * v[idx];
        }
        uo[i] = ur;
        vo[i] = vr;
    } else if (nn >= 0) {
        uo[i] = u[nn];
        vo[i] = v[nn];
    } else {
        uo[i] = __builtin_nanf("");
        vo[i] = __builtin_nanf("");
    }
}

__global__ __launch_bounds__(256) void k_advect_f64(TreeView T, int n,
                                                    const double *__restrict__ u,
                                                    const double *__restrict__ v, float dt,
                                                    double *__restrict__ uo,
                                                    double *__restrict__ vo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 pos = T.xy[i];
    const float sx = pos.x - (float)u[i] * dt;
    const float sy = pos.y - (float)v[i] * dt;
    float xi = 0.f, eta = 0.f;
    int nn = -1;
    const int e = find_element(T, sx, sy, xi, eta, nn);
    if (e > -1) {
        double c[6];
        p2_shape((double)xi, (double)eta, c);
        double ur = 0, vr = 0;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int idx = T.elem[6 * e + k];
            ur += c[k] * u[idx];
            vr += c[k] * v[idx];
        }
        uo[i] = ur;
        vo[i] = vr;
    } else if (nn >= 0) {
        uo[i] = u[nn];
        vo[i] = v[nn];
    } else {
        uo[i] = __builtin_nan("");
        vo[i] = __builtin_nan("");
    }
}

__global__ __launch_bounds__(256) void k_locate(TreeView T, int n, const float2 *__restrict__ pts,
                                                int32_t *__restrict__ elem, float *__restrict__ xi,
                                                float *__restrict__ eta,
                                                int32_t *__restrict__ nearest) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 p = pts[i];
    float a = 0.f, b = 0.f;
    int nn = -1;
    const int e = find_element(T, p.x, p.y, a, b, nn);
    elem[i] = e;
    xi[i] = e >= 0 ? a : 0.f;
    eta[i] = e >= 0 ? b : 0.f;
    nearest[i] = e >= 0 ? -1 : nn;
}

__global__ void k_tri_xy(int ne, const int32_t *__restrict__ elem, const float2 *__restrict__ xy,
                         float2 *__restrict__ tri) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    tri[3 * e] = xy[elem[6 * e]];
    tri[3 * e + 1] = xy[elem[6 * e + 1]];
    tri[3 * e + 2] = xy[elem[6 * e + 2]];
}

TreeView view(const KdTreeGpu &t) {
    NSFEM_REQUIRE(t.ready, NSFEM_NOT_READY, "KD tree not uploaded");
    return TreeView{reinterpret_cast<const uint2 *>(t.nodes.ptr), t.leaf.ptr,
                    reinterpret_cast<const float2 *>(t.xy.ptr), t.elements.ptr,
                    reinterpret_cast<const float2 *>(t.tri_xy.ptr)};
}

}  // namespace

void build_tri_xy(Device &d, KdTreeGpu &t) {
    t.tri_xy.resize((size_t)t.elements_count * 6);
    if (t.elements_count == 0) return;
    k_tri_xy<<<(t.elements_count + 255) / 256, 256, 0, d.stream>>>(
        t.elements_count, t.elements.ptr, reinterpret_cast<const float2 *>(t.xy.ptr),
        reinterpret_cast<float2 *>(t.tri_xy.ptr));
    NSFEM_HIP(hipGetLastError());
}

void launch_advect_f32(Device &d, const KdTreeGpu &t, int n, const float *u, const float *v,
                       float dt, float *uo, float *vo) {
    if (n <= 0) return;
    NSFEM_REQUIRE(n <= t.mesh_nodes, NSFEM_INVALID_ARGUMENT,
                  "advect: more velocity nodes than mesh nodes");
    k_advect_f32<<<(n + 255) / 256, 256, 0, d.stream>>>(view(t), n, u, v, dt, uo, vo);
    NSFEM_HIP(hipGetLastError());
}

void launch_advect_f64(Device &d, const KdTreeGpu &t, int n, const double *u, const double *v,
                       double dt, double *uo, double *vo) {
    if (n <= 0) return;
    NSFEM_REQUIRE(n <= t.mesh_nodes, NSFEM_INVALID_ARGUMENT,
                  "advect: more velocity nodes than mesh nodes");
    k_advect_f64<<<(n + 255) / 256, 256, 0, d.stream>>>(view(t), n, u, v, (float)dt, uo, vo);
    NSFEM_HIP(hipGetLastError());
}

void launch_locate(Device &d, const KdTreeGpu &t, int n, const float *pts, int32_t *elem,
                   float *xi, float *eta, int32_t *nearest) {
    if (n <= 0) return;
    k_locate<<<(n + 255) / 256, 256, 0, d.stream>>>(
        view(t), n, reinterpret_cast<const float2 *>(pts), elem, xi, eta, nearest);
    NSFEM_HIP(hipGetLastError());
}

}  // namespace nsfem
