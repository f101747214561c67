/*
 * CPU oracle for the CG + semi-Lagrangian advection path.
 * TEST INFRASTRUCTURE ONLY -- see nsfem_oracle.h for the rules and the list
 * of checks that pin each function.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "nsfem_oracle.h"

#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* threads                                                             */
/* ------------------------------------------------------------------ */
static int g_threads = 0;

void or_set_threads(int n) {
    g_threads = n > 0 ? n : 0;
    omp_set_num_threads(n > 0 ? n : omp_get_num_procs());
}

int or_get_threads(void) {
    return g_threads > 0 ? g_threads : omp_get_num_procs();
}

/* ------------------------------------------------------------------ */
/* CSR linear algebra                                                  */
/* ------------------------------------------------------------------ */

/* spRMult: res[row] = sum_i values[i]*mult[col[i]] in row order
 * (gpu/matrix_math.cu:19-40). */
void or_spmv(int rows, const int32_t *rs, const int32_t *col,
             const double *val, const double *x, double *y) {
#pragma omp parallel for schedule(static, 4096)
    for (int r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int32_t i = rs[r]; i < rs[r + 1]; ++i) s += val[i] * x[col[i]];
        y[r] = s;
    }
}

/* spRMultSub: res = lhs - A*mult; res may alias lhs
 * (gpu/matrix_math.cu:66-87). */
void or_spmv_sub(int rows, const int32_t *rs, const int32_t *col,
                 const double *val, const double *lhs, const double *x,
                 double *y) {
#pragma omp parallel for schedule(static, 4096)
    for (int r = 0; r < rows; ++r) {
        double s = 0.0;
        for (int32_t i = rs[r]; i < rs[r + 1]; ++i) s += val[i] * x[col[i]];
        y[r] = lhs[r] - s;
    }
}

/* Deterministic blocked dot product: partial sums over blocks of 8192,
 * added in block order (thesis 04_implementation_details.tex:288; the GPU
 * kernel it replaces is gpu/matrix_math.cu:162-185). */
#define OR_DOT_BLOCK 8192
double or_dot(int64_t n, const double *a, const double *b) {
    const int64_t nb = (n + OR_DOT_BLOCK - 1) / OR_DOT_BLOCK;
    if (nb == 0) return 0.0;
    double stack_part[64];
    double *part = nb <= 64 ? stack_part : (double *)malloc(sizeof(double) * nb);
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < nb; ++k) {
        const int64_t lo = k * OR_DOT_BLOCK;
        const int64_t hi = lo + OR_DOT_BLOCK < n ? lo + OR_DOT_BLOCK : n;
        double s = 0.0;
        for (int64_t i = lo; i < hi; ++i) s += a[i] * b[i];
        part[k] = s;
    }
    double s = 0.0;
    for (int64_t k = 0; k < nb; ++k) s += part[k];
    if (part != stack_part) free(part);
    return s;
}

/* saxpy: y += a*x (gpu/matrix_math.cu:106-113). */
void or_axpy(int64_t n, double a, const double *x, double *y) {
#pragma omp parallel for schedule(static, 8192)
    for (int64_t i = 0; i < n; ++i) y[i] += a * x[i];
}

/* saxpby: out = a*x + b*y (gpu/matrix_math.cu:131-145). */
void or_axpby(int64_t n, double a, double b, const double *x, const double *y,
              double *out) {
#pragma omp parallel for schedule(static, 8192)
    for (int64_t i = 0; i < n; ++i) out[i] = a * x[i] + b * y[i];
}

/* Triplet -> CSR (cpp/gpu_sparse_matrix.cpp:7-49).  The reference receives
 * an SMM::TripletMatrix whose iteration yields each (row, col) once, with
 * duplicates already summed, in ascending column order per row (asserted at
 * :29).  Here: stable counting sort by row, stable insertion sort by column
 * inside a row, then duplicates summed in input order. */
int or_triplet_to_csr(int rows, int64_t count, const int32_t *ti,
                      const int32_t *tj, const double *tv, int32_t *row_start,
                      int32_t *col, double *val, int64_t *nnz_out) {
    int64_t *cnt = (int64_t *)calloc((size_t)rows + 1, sizeof(int64_t));
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (count > 0 ? count : 1));
    if (!cnt || !perm) { free(cnt); free(perm); return -1; }
    for (int64_t k = 0; k < count; ++k) {
        if (ti[k] < 0 || ti[k] >= rows) { free(cnt); free(perm); return -3; }
        cnt[ti[k] + 1]++;
    }
    for (int r = 0; r < rows; ++r) cnt[r + 1] += cnt[r];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * ((size_t)rows + 1));
    memcpy(pos, cnt, sizeof(int64_t) * ((size_t)rows + 1));
    for (int64_t k = 0; k < count; ++k) perm[pos[ti[k]]++] = k;
    free(pos);
    /* stable insertion sort by column inside each row */
    for (int r = 0; r < rows; ++r) {
        for (int64_t a = cnt[r] + 1; a < cnt[r + 1]; ++a) {
            const int64_t key = perm[a];
            int64_t b = a - 1;
            while (b >= cnt[r] && tj[perm[b]] > tj[key]) {
                perm[b + 1] = perm[b];
                --b;
            }
            perm[b + 1] = key;
        }
    }
    int64_t nnz = 0;
    for (int r = 0; r < rows; ++r) {
        if (row_start) row_start[r] = (int32_t)nnz;
        int64_t a = cnt[r];
        while (a < cnt[r + 1]) {
            const int32_t c = tj[perm[a]];
            double s = 0.0;
            int first = 1;
            while (a < cnt[r + 1] && tj[perm[a]] == c) {
                s = first ? tv[perm[a]] : s + tv[perm[a]];
                first = 0;
                ++a;
            }
            if (col) {
                col[nnz] = c;
                val[nnz] = s;
            }
            ++nnz;
        }
    }
    if (row_start) row_start[rows] = (int32_t)nnz;
    *nnz_out = nnz;
    free(cnt);
    free(perm);
    return 0;
}

/* CG.  Set-up as cpp/gpu_simulation_device.cpp:425-451 (r0 = b - A x0 by
 * spRMultSub, p = r, rr = (r, r), early exit, maxIterations -1 -> rows),
 * loop body in the order of gpu/matrix_math.cu:256-286:
 *   ap = A p; pAp = (ap, p); alpha = rr/pAp; x += alpha p; r -= alpha ap;
 *   rr' = (r, r); converged? ; beta = rr'/rr; p = 1*r + beta*p.
 * Differences from the reference, by design: FP64, relative+absolute
 * tolerance, x always written. */
int or_cg(int rows, const int32_t *rs, const int32_t *col, const double *val,
          const double *b, const double *x0, double *x, int max_iter,
          double rtol, double atol, or_cg_result *res) {
    double *r = (double *)malloc(sizeof(double) * (size_t)rows);
    double *p = (double *)malloc(sizeof(double) * (size_t)rows);
    double *ap = (double *)malloc(sizeof(double) * (size_t)rows);
    if (x != x0) memcpy(x, x0, sizeof(double) * (size_t)rows);
    or_spmv_sub(rows, rs, col, val, b, x, r);
    memcpy(p, r, sizeof(double) * (size_t)rows);
    double rr = or_dot(rows, r, r);
    const double bnorm = sqrt(or_dot(rows, b, b));
    const double r0 = sqrt(rr);
    const double ref = bnorm > 0.0 ? bnorm : r0;
    double tol = rtol * ref;
    if (atol > tol) tol = atol;
    const double tol2 = tol * tol;
    if (max_iter < 0) max_iter = rows;
    int status = -2, it = 0;
    if (rr <= tol2) {
        status = 0;
    } else {
        for (it = 0; it < max_iter;) {
            or_spmv(rows, rs, col, val, p, ap);
            const double pap = or_dot(rows, ap, p);
            if (!(pap > 0.0)) { status = -7; break; }
            const double alpha = rr / pap;
            or_axpy(rows, alpha, p, x);
            or_axpy(rows, -alpha, ap, r);
            const double rr_new = or_dot(rows, r, r);
            ++it;
            if (rr_new <= tol2) { rr = rr_new; status = 0; break; }
            const double beta = rr_new / rr;
            or_axpby(rows, 1.0, beta, r, p, p);
            rr = rr_new;
        }
    }
    if (res) {
        res->iterations = it;
        res->converged = status == 0;
        res->rhs_norm = bnorm;
        res->initial_residual = r0;
        res->recursive_residual = sqrt(rr);
        or_spmv_sub(rows, rs, col, val, b, x, ap);
        res->final_residual = sqrt(or_dot(rows, ap, ap));
    }
    free(r);
    free(p);
    free(ap);
    return status;
}

/* ------------------------------------------------------------------ */
/* KD tree                                                             */
/* ------------------------------------------------------------------ */
static inline uint32_t f2u(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}
static inline float u2f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}

/* BBox2D() starts at (FLT_MAX, FLT_MAX)/(-FLT_MAX, -FLT_MAX); expand()
 * takes min(point, min) for the minimum but max(point.x, point.x) for the
 * maximum (gpu_cpu_shared/misc_common.cuh:40-65), applied to every node in
 * file order by FemGrid2D::loadJSON (cpp/grid.cpp:27-33). */
void or_grid_bbox(int node_count, const float *xy, float bbox[4]) {
    float mnx = FLT_MAX, mny = FLT_MAX, mxx = -FLT_MAX, mxy = -FLT_MAX;
    for (int i = 0; i < node_count; ++i) {
        const float px = xy[2 * i], py = xy[2 * i + 1];
        mnx = px < mnx ? px : mnx; /* NSFemGPU::min(a, b) = a < b ? a : b */
        mny = py < mny ? py : mny;
        mxx = px;                  /* max(point.x, point.x) */
        mxy = py;
    }
    bbox[0] = mnx; bbox[1] = mny; bbox[2] = mxx; bbox[3] = mxy;
}

typedef struct {
    uint32_t *nodes;
    int64_t count, cap;
    int32_t *leaf;
    int64_t lcount, lcap;
    int failed;
} kd_builder;

static void kd_push_node(kd_builder *b, uint32_t w0, uint32_t w1) {
    if (b->count == b->cap) {
        b->cap = b->cap ? 2 * b->cap : 1024;
        b->nodes = (uint32_t *)realloc(b->nodes, sizeof(uint32_t) * 2 * (size_t)b->cap);
    }
    b->nodes[2 * b->count] = w0;
    b->nodes[2 * b->count + 1] = w1;
    b->count++;
}

static void kd_push_leaf(kd_builder *b, const int32_t *idx, int64_t n) {
    if (b->lcount + n > b->lcap) {
        while (b->lcount + n > b->lcap) b->lcap = b->lcap ? 2 * b->lcap : 4096;
        b->leaf = (int32_t *)realloc(b->leaf, sizeof(int32_t) * (size_t)b->lcap);
    }
    memcpy(b->leaf + b->lcount, idx, sizeof(int32_t) * (size_t)n);
    b->lcount += n;
}

/* KDTreeBuilder::build, recursive part (cpp/kd_tree_builder.cpp:94-148). */
static int64_t kd_build_rec(kd_builder *b, const float *xy, const int32_t *el,
                            const int32_t *idx, int64_t n, const float mn[2],
                            const float mx[2], int axis, int level,
                            int max_depth, int min_leaf) {
    if (level >= max_depth || n <= (int64_t)min_leaf) {
        const int64_t off = b->lcount;
        kd_push_leaf(b, idx, n);
        /* KDNode::makeLeaf: flags = 2 | (count << 2); offset as uint */
        kd_push_node(b, 2u | ((uint32_t)n << 2), (uint32_t)off);
        return b->count;
    }
    const float split = (mn[axis] + mx[axis]) * 0.5f;
    float lmax[2] = {mx[0], mx[1]};
    lmax[axis] = split;
    float rmin[2] = {mn[0], mn[1]};
    rmin[axis] = split;
    const int new_axis = (axis + 1) % 2;
    int32_t *li = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t *ri = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t nl = 0, nr = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int32_t e = idx[k];
        const float a = xy[2 * el[6 * e + 0] + axis];
        const float bb = xy[2 * el[6 * e + 1] + axis];
        const float c = xy[2 * el[6 * e + 2] + axis];
        if (a <= split || bb <= split || c <= split) li[nl++] = e;
        if (a >= split || bb >= split || c >= split) ri[nr++] = e;
    }
    const int64_t cur = b->count;
    kd_push_node(b, 0u, 0u);
    const int64_t right = kd_build_rec(b, xy, el, li, nl, mn, lmax, new_axis,
                                       level + 1, max_depth, min_leaf);
    free(li);
    /* KDNode::makeInternal: flags = axis | (right << 2); split bits */
    b->nodes[2 * cur] = (uint32_t)axis | ((uint32_t)right << 2);
    b->nodes[2 * cur + 1] = f2u(split);
    const int64_t ret = kd_build_rec(b, xy, el, ri, nr, rmin, mx, new_axis,
                                     level + 1, max_depth, min_leaf);
    free(ri);
    return ret;
}

int or_kd_build(int node_count, const float *xy, int elem_count,
                const int32_t *elems, int max_depth, int min_leaf_size,
                or_kdtree *out) {
    memset(out, 0, sizeof(*out));
    /* KDTreeBuilder(int maxDepth, int minLeafSize) clamps maxDepth to 29;
     * the default builder uses maxDepth -1 and minLeafSize 16
     * (cpp/kd_tree_builder.cpp:7-15); build() resolves -1 with the pbrt
     * formula (:87). */
    if (max_depth > 29) max_depth = 29;
    if (max_depth < 0) {
        const double d = round(8.0 + (double)1.3f * log((double)elem_count));
        max_depth = (int)d < 29 ? (int)d : 29;
    }
    if (min_leaf_size < 0) min_leaf_size = 16;
    float bbox[4];
    or_grid_bbox(node_count, xy, bbox);
    int32_t *idx = (int32_t *)malloc(sizeof(int32_t) * (size_t)(elem_count > 0 ? elem_count : 1));
    for (int i = 0; i < elem_count; ++i) idx[i] = i;
    kd_builder b;
    memset(&b, 0, sizeof(b));
    const float mn[2] = {bbox[0], bbox[1]};
    const float mx[2] = {bbox[2], bbox[3]};
    kd_build_rec(&b, xy, elems, idx, elem_count, mn, mx, 0, 0, max_depth,
                 min_leaf_size);
    free(idx);
    out->node_count = (int32_t)b.count;
    out->nodes = b.nodes;
    out->leaf_index_count = b.lcount;
    out->leaf_indices = b.leaf;
    memcpy(out->bbox, bbox, sizeof(bbox));
    out->max_depth = max_depth;
    return 0;
}

void or_kd_free(or_kdtree *t) {
    free(t->nodes);
    free(t->leaf_indices);
    memset(t, 0, sizeof(*t));
}

/* isPointInTriagle (gpu_cpu_shared/kd_tree_common.cuh:18-61). */
int or_point_in_triangle(float px, float py, float ax, float ay, float bx,
                         float by, float cx, float cy, float *xi, float *eta) {
    const float abx = bx - ax, aby = by - ay;
    const float acx = cx - ax, acy = cy - ay;
    const float apx = px - ax, apy = py - ay;
    const float D = (abx * acy) - (aby * acx);
    const int dsign = D > 0 ? 1 : -1;
    const float b1 = (apx * acy - apy * acx) * (float)dsign;
    if (b1 < 0 || b1 > D) return 0;
    const float b2 = (apy * abx - apx * aby) * (float)dsign;
    if (b2 < 0 || b1 > D - b2) return 0;
    const float dinv = 1 / fabsf(D);
    *xi = b1 * dinv;
    *eta = b2 * dinv;
    return 1;
}

static inline int kd_leaf_offset(uint32_t w1, int float_offset) {
    /* getTriangleOffset() returns float (kd_tree_common.cuh:148-150); the
     * int conversion at kd_tree.cuh:58,157 then reads the rounded value. */
    return float_offset ? (int)(float)w1 : (int)w1;
}

/* nearestNeghbourProcessLeaf (gpu_cpu_shared/kd_tree.cuh:149-169). */
static void kd_nn_leaf(uint32_t w0, uint32_t w1, const int32_t *leaf,
                       const float *xy, const int32_t *el, float px, float py,
                       float *min_d, int *closest, int float_offset) {
    const int off = kd_leaf_offset(w1, float_offset);
    const int cnt = (int)(w0 >> 2);
    for (int i = 0; i < cnt; ++i) {
        const int e = leaf[off + i];
        for (int j = 0; j < 6; ++j) {
            const int nidx = el[6 * e + j];
            const float ox = xy[2 * nidx], oy = xy[2 * nidx + 1];
            const float d = (px - ox) * (px - ox) + (py - oy) * (py - oy);
            if (d < *min_d) {
                *min_d = d;
                *closest = nidx;
            }
        }
    }
}

/* KDTree::findElement (gpu_cpu_shared/kd_tree.cuh:40-141). */
int or_kd_find_element(const uint32_t *nodes, const int32_t *leaf,
                       const float *xy, const int32_t *el, float px, float py,
                       float *xi, float *eta, int *nearest, int float_offset) {
    *nearest = -1;
    /* A NaN coordinate satisfies neither `<=` nor `>` and the reference's
     * descent would cycle back to the root forever; report "not found". */
    if (isnan(px) || isnan(py)) return -1;
    int stack_node[32], stack_cnt[32];
    int sp = 0;
    const float pt[2] = {px, py};
    int cur = 0;
    uint32_t w0 = nodes[0], w1 = nodes[1];
    while ((w0 & 3u) != 2u) {
        const int axis = (int)(w0 & 3u);
        const float s = u2f(w1);
        const int child = pt[axis] <= s ? cur + 1 : (int)(w0 >> 2);
        stack_node[sp] = cur; stack_cnt[sp] = 1; ++sp;
        cur = child;
        w0 = nodes[2 * cur];
        w1 = nodes[2 * cur + 1];
    }
    stack_node[sp] = cur; stack_cnt[sp] = 0; ++sp;
    {
        const int off = kd_leaf_offset(w1, float_offset);
        const int cnt = (int)(w0 >> 2);
        for (int i = 0; i < cnt; ++i) {
            const int e = leaf[off + i];
            const int a = el[6 * e], b = el[6 * e + 1], c = el[6 * e + 2];
            if (or_point_in_triangle(px, py, xy[2 * a], xy[2 * a + 1], xy[2 * b],
                                     xy[2 * b + 1], xy[2 * c], xy[2 * c + 1], xi,
                                     eta))
                return e;
        }
    }
    float min_d = FLT_MAX;
    while (sp > 0) {
        const int top = stack_node[sp - 1];
        const uint32_t t0 = nodes[2 * top], t1 = nodes[2 * top + 1];
        if ((t0 & 3u) == 2u) {
            kd_nn_leaf(t0, t1, leaf, xy, el, px, py, &min_d, nearest, float_offset);
            --sp;
        } else {
            while (sp > 0 && stack_cnt[sp - 1] == 2) --sp;
            if (sp == 0) break;
            const int act = stack_node[sp - 1];
            const uint32_t a0 = nodes[2 * act], a1 = nodes[2 * act + 1];
            const int axis = (int)(a0 & 3u);
            const float s = u2f(a1);
            if (stack_cnt[sp - 1] == 0) {
                const int next = pt[axis] <= s ? act + 1 : (int)(a0 >> 2);
                stack_cnt[sp - 1]++;
                stack_node[sp] = next; stack_cnt[sp] = 0; ++sp;
            } else {
                const float dd = pt[axis] - s;
                const float d2 = dd * dd;
                if (d2 < min_d) {
                    const int far = pt[axis] <= s ? (int)(a0 >> 2) : act + 1;
                    stack_cnt[sp - 1]++;
                    stack_node[sp] = far; stack_cnt[sp] = 0; ++sp;
                } else {
                    --sp;
                }
            }
        }
    }
    return -1;
}

void or_kd_find_many(const uint32_t *nodes, const int32_t *leaf,
                     const float *xy, const int32_t *el, int n,
                     const float *pts, int32_t *element, float *xi,
                     float *eta, int32_t *nearest, int float_offset) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        float a = 0.0f, b = 0.0f;
        int nn = -1;
        const int e = or_kd_find_element(nodes, leaf, xy, el, pts[2 * i],
                                         pts[2 * i + 1], &a, &b, &nn, float_offset);
        element[i] = e;
        xi[i] = e >= 0 ? a : 0.0f;
        eta[i] = e >= 0 ? b : 0.0f;
        nearest[i] = e >= 0 ? -1 : nn;
    }
}

/* P2ShapeEval (gpu/advection.cu:5-12), in float. */
static inline void p2_shape_f(float xi, float eta, float out[6]) {
    out[0] = 1 - 3 * xi - 3 * eta + 2 * xi * xi + 4 * xi * eta + 2 * eta * eta;
    out[1] = 2 * xi * xi - xi;
    out[2] = 2 * eta * eta - eta;
    out[3] = 4 * xi * eta;
    out[4] = 4 * eta - 4 * xi * eta - 4 * eta * eta;
    out[5] = 4 * xi - 4 * xi * xi - 4 * xi * eta;
}

static inline void p2_shape_d(double xi, double eta, double out[6]) {
    out[0] = 1 - 3 * xi - 3 * eta + 2 * xi * xi + 4 * xi * eta + 2 * eta * eta;
    out[1] = 2 * xi * xi - xi;
    out[2] = 2 * eta * eta - eta;
    out[3] = 4 * xi * eta;
    out[4] = 4 * eta - 4 * xi * eta - 4 * eta * eta;
    out[5] = 4 * xi - 4 * xi * xi - 4 * xi * eta;
}

/* advect (gpu/advection.cu:14-57). */
void or_advect_f32(const uint32_t *nodes, const int32_t *leaf,
                   const float *xy, const int32_t *el, int n, const float *u,
                   const float *v, float dt, float *u_out, float *v_out,
                   int float_offset) {
#pragma omp parallel for schedule(dynamic, 512)
    for (int i = 0; i < n; ++i) {
        const float px = xy[2 * i], py = xy[2 * i + 1];
        const float sx = px - u[i] * dt;
        const float sy = py - v[i] * dt;
        float xi = 0.0f, eta = 0.0f;
        int nn = -1;
        const int e = or_kd_find_element(nodes, leaf, xy, el, sx, sy, &xi, &eta,
                                         &nn, float_offset);
        if (e > -1) {
            float c[6];
            p2_shape_f(xi, eta, c);
            float ur = 0, vr = 0;
            for (int k = 0; k < 6; ++k) {
                const int idx = el[6 * e + k];
                ur += c[k] * u[idx];
                vr += c[k] * v[idx];
            }
            u_out[i] = ur;
            v_out[i] = vr;
        } else if (nn >= 0) {
            u_out[i] = u[nn];
            v_out[i] = v[nn];
        } else {
            u_out[i] = NAN;
            v_out[i] = NAN;
        }
    }
}

void or_advect_f64(const uint32_t *nodes, const int32_t *leaf,
                   const float *xy, const int32_t *el, int n,
                   const double *u, const double *v, double dt, double *u_out,
                   double *v_out, int float_offset) {
    const float dtf = (float)dt;
#pragma omp parallel for schedule(dynamic, 512)
    for (int i = 0; i < n; ++i) {
        const float px = xy[2 * i], py = xy[2 * i + 1];
        const float sx = px - (float)u[i] * dtf;
        const float sy = py - (float)v[i] * dtf;
        float xi = 0.0f, eta = 0.0f;
        int nn = -1;
        const int e = or_kd_find_element(nodes, leaf, xy, el, sx, sy, &xi, &eta,
                                         &nn, float_offset);
        if (e > -1) {
            double c[6];
            p2_shape_d((double)xi, (double)eta, c);
            double ur = 0, vr = 0;
            for (int k = 0; k < 6; ++k) {
                const int idx = el[6 * e + k];
                ur += c[k] * u[idx];
                vr += c[k] * v[idx];
            }
            u_out[i] = ur;
            v_out[i] = vr;
        } else if (nn >= 0) {
            u_out[i] = u[nn];
            v_out[i] = v[nn];
        } else {
            u_out[i] = NAN;
            v_out[i] = NAN;
        }
    }
}

/* ------------------------------------------------------------------ */
/* FEM assembly                                                        */
/* ------------------------------------------------------------------ */

/* TriangleIntegrator (include/assembly.h:259-304): 8 points, the 8th is
 * the value-initialised (0, 0) with weight 0. */
static const double QX[8] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0, 1.0 / 3.0, 0.0};
static const double QY[8] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5, 1.0 / 3.0, 0.0};
static const double QW[8] = {3.0 / 120.0, 3.0 / 120.0, 3.0 / 120.0,
                             8.0 / 120.0, 8.0 / 120.0, 8.0 / 120.0,
                             27.0 / 120.0, 0.0};

/* P1::eval / P1::del (include/assembly.h:154-198). */
static void p1_eval(double xi, double eta, double o[3]) {
    o[0] = 1 - xi - eta;
    o[1] = xi;
    o[2] = eta;
}
static void p1_del(double xi, double eta, double o[2][3]) {
    (void)xi; (void)eta;
    o[0][0] = -1.0; o[0][1] = 1.0; o[0][2] = 0.0;
    o[1][0] = -1.0; o[1][1] = 0.0; o[1][2] = 1.0;
}
/* P2::del (include/assembly.h:237-256). */
static void p2_del(double xi, double eta, double o[2][6]) {
    o[0][0] = -3 + 4 * eta + 4 * xi;
    o[0][1] = -1 + 4 * xi;
    o[0][2] = 0.0;
    o[0][3] = 4 * eta;
    o[0][4] = -4 * eta;
    o[0][5] = 4 - 4 * eta - 8 * xi;
    o[1][0] = -3 + 4 * eta + 4 * xi;
    o[1][1] = 0;
    o[1][2] = -1 + 4 * eta;
    o[1][3] = 4 * xi;
    o[1][4] = 4 - 8 * eta - 4 * xi;
    o[1][5] = -4 * xi;
}

typedef struct {
    double mass2[6][6];      /* LocalMassFunctor<P2>::shapeSquared */
    double delsq2[12][12];   /* LocalStiffnessFunctor<P2>::delPSq */
    double delsq1[6][6];     /* LocalStiffnessFunctor<P1>::delPSq */
    double comb_vd[3][12];   /* LocalDivergenceFunctor<P1, P2>::combined */
    double comb_pd[6][6];    /* LocalDivergenceFunctor<P2, P1>::combined */
} fem_tables;

static void fem_tables_init(fem_tables *t) {
    memset(t, 0, sizeof(*t));
    for (int q = 0; q < 8; ++q) {
        const double x = QX[q], y = QY[q], w = QW[q];
        double f2[6], d2[2][6], f1[3], d1[2][3];
        p2_shape_d(x, y, f2);
        p2_del(x, y, d2);
        p1_eval(x, y, f1);
        p1_del(x, y, d1);
        const double *d2f = &d2[0][0];
        const double *d1f = &d1[0][0];
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) t->mass2[i][j] += (f2[i] * f2[j]) * w;
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) t->delsq2[i][j] += (d2f[i] * d2f[j]) * w;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) t->delsq1[i][j] += (d1f[i] * d1f[j]) * w;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 6; ++j)
                for (int k = 0; k < 2; ++k)
                    t->comb_vd[i][j + k * 6] += (f1[i] * d2[k][j]) * w;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 2; ++k)
                    t->comb_pd[i][j + k * 3] += (f2[i] * d1[k][j]) * w;
    }
}

/* differentialOperator (include/assembly.h:131-150). */
static void diff_op(const double *nd, double *det, double B[2][2]) {
    const double xk = nd[0], yk = nd[1], xl = nd[2], yl = nd[3], xm = nd[4], ym = nd[5];
    B[0][0] = ym - yk;
    B[0][1] = yk - yl;
    B[1][0] = xk - xm;
    B[1][1] = xl - xk;
    *det = B[0][0] * B[1][1] - B[0][1] * B[1][0];
}

/* linTriangleTmJacobian (include/assembly.h:98-115). */
static double lin_jac(const double *nd) {
    const double a = nd[2] - nd[0];
    const double b = nd[3] - nd[1];
    const double c = nd[4] - nd[0];
    const double d = nd[5] - nd[1];
    return a * d - b * c;
}

/* LocalStiffnessFunctor::operator() (include/assembly.h:749-787). */
static void local_stiffness(const double *delsq, int S, const double *nd,
                            double *out /* S*S */) {
    double B[2][2], J;
    diff_op(nd, &J, B);
    J = 1.0 / fabs(J);
    const int W = 2 * S;
    for (int i = 0; i < S; ++i) {
        for (int j = 0; j < S; ++j) {
            double acc = 0.0;
            const double sq0 = delsq[i * W + j];
            const double sq1 = delsq[i * W + S + j];
            const double sq2 = delsq[(S + i) * W + j];
            const double sq3 = delsq[(S + i) * W + S + j];
            for (int k = 0; k < 2; ++k) {
                acc += sq0 * B[k][0] * B[k][0] + sq1 * B[k][0] * B[k][1] +
                       sq2 * B[k][0] * B[k][1] + sq3 * B[k][1] * B[k][1];
            }
            out[i * S + j] = acc * J;
        }
    }
}

typedef struct {
    int32_t *i, *j;
    double *v;
    int64_t n, cap;
} triplets;

static void tr_add(triplets *t, int32_t i, int32_t j, double v) {
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1 << 16;
        t->i = (int32_t *)realloc(t->i, sizeof(int32_t) * (size_t)t->cap);
        t->j = (int32_t *)realloc(t->j, sizeof(int32_t) * (size_t)t->cap);
        t->v = (double *)realloc(t->v, sizeof(double) * (size_t)t->cap);
    }
    t->i[t->n] = i;
    t->j[t->n] = j;
    t->v[t->n] = v;
    t->n++;
}

static void tr_free(triplets *t) {
    free(t->i); free(t->j); free(t->v);
    memset(t, 0, sizeof(*t));
}

static int tr_to_csr(const triplets *t, int rows, int cols, or_csr *out) {
    int64_t nnz = 0;
    out->rows = rows;
    out->cols = cols;
    out->row_start = (int32_t *)malloc(sizeof(int32_t) * ((size_t)rows + 1));
    int rc = or_triplet_to_csr(rows, t->n, t->i, t->j, t->v, out->row_start,
                               NULL, NULL, &nnz);
    if (rc) return rc;
    out->col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nnz > 0 ? nnz : 1));
    out->val = (double *)malloc(sizeof(double) * (size_t)(nnz > 0 ? nnz : 1));
    rc = or_triplet_to_csr(rows, t->n, t->i, t->j, t->v, out->row_start,
                           out->col, out->val, &nnz);
    out->nnz = nnz;
    return rc;
}

static void get_element(const float *xy, const int32_t *el, int e, int idx[6],
                        double nd[12]) {
    for (int k = 0; k < 6; ++k) {
        idx[k] = el[6 * e + k];
        nd[2 * k] = (double)xy[2 * idx[k]];
        nd[2 * k + 1] = (double)xy[2 * idx[k] + 1];
    }
}

void or_csr_free(or_csr *m) {
    free(m->row_start); free(m->col); free(m->val);
    memset(m, 0, sizeof(*m));
}

void or_matrices_free(or_matrices *m) {
    or_csr_free(&m->velocity_mass);
    or_csr_free(&m->diffusion);
    or_csr_free(&m->velocity_weights);
    or_csr_free(&m->pressure_stiffness);
    or_csr_free(&m->pressure_weights);
    or_csr_free(&m->velocity_divergence);
    or_csr_free(&m->pressure_divergence);
}

/* assembleAllMatrices (include/assembly.h:1302-1475) with assembleMatrix
 * (:1034-1054), assembleBCMatrix (:994-1032), assembleDivergenceMatrix
 * (:1627-1662) and the local functors (:725-882). */
int or_assemble(int V, int P, const float *xy, int E, const int32_t *el,
                int nvb, const int32_t *vb, int npb, const int32_t *pb,
                double nu, double dt, or_matrices *out) {
    memset(out, 0, sizeof(*out));
    fem_tables T;
    fem_tables_init(&T);
    unsigned char *vmask = (unsigned char *)calloc((size_t)V, 1);
    unsigned char *pmask = (unsigned char *)calloc((size_t)P, 1);
    for (int k = 0; k < nvb; ++k) vmask[vb[k]] = 1;
    for (int k = 0; k < npb; ++k) pmask[pb[k]] = 1;

    triplets tm = {0}, td = {0}, tw = {0}, tp = {0}, tpw = {0}, tvd = {0}, tpd = {0};
    int idx[6];
    double nd[12];
    double mloc[36], kloc[36], ploc[9];
    for (int e = 0; e < E; ++e) {
        get_element(xy, el, e, idx, nd);
        const double jabs = fabs(lin_jac(nd));
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) mloc[i * 6 + j] = T.mass2[i][j] * jabs;
        local_stiffness(&T.delsq2[0][0], 6, nd, kloc);
        local_stiffness(&T.delsq1[0][0], 3, nd, ploc);
        /* velocity mass: assembleMatrix */
        for (int r = 0; r < 6; ++r)
            for (int c = 0; c < 6; ++c) tr_add(&tm, idx[r], idx[c], mloc[r * 6 + c]);
        /* diffusion: mass + stiffness*dt*viscosity, assembleBCMatrix */
        for (int r = 0; r < 6; ++r) {
            const int gr = idx[r];
            for (int c = 0; c < 6; ++c) {
                const int gc = idx[c];
                const double val = mloc[r * 6 + c] + kloc[r * 6 + c] * dt * nu;
                if (!vmask[gr] && !vmask[gc]) tr_add(&td, gr, gc, val);
                else if (!vmask[gr] && vmask[gc]) tr_add(&tw, gc, gr, val);
            }
        }
        /* pressure stiffness (P1), assembleBCMatrix */
        for (int r = 0; r < 3; ++r) {
            const int gr = idx[r];
            for (int c = 0; c < 3; ++c) {
                const int gc = idx[c];
                const double val = ploc[r * 3 + c];
                if (!pmask[gr] && !pmask[gc]) tr_add(&tp, gr, gc, val);
                else if (!pmask[gr] && pmask[gc]) tr_add(&tpw, gc, gr, val);
            }
        }
        /* divergence matrices (LocalDivergenceFunctor::operator(), :812-826) */
        double B[2][2], J;
        diff_op(nd, &J, B);
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 6; ++c) {
                const double dx = B[0][0] * T.comb_vd[r][c] + B[0][1] * T.comb_vd[r][c + 6];
                const double dy = B[1][0] * T.comb_vd[r][c] + B[1][1] * T.comb_vd[r][c + 6];
                tr_add(&tvd, idx[r], idx[c], dx);
                tr_add(&tvd, idx[r], idx[c] + V, dy);
            }
        }
        for (int r = 0; r < 6; ++r) {
            for (int c = 0; c < 3; ++c) {
                const double dx = B[0][0] * T.comb_pd[r][c] + B[0][1] * T.comb_pd[r][c + 3];
                const double dy = B[1][0] * T.comb_pd[r][c] + B[1][1] * T.comb_pd[r][c + 3];
                tr_add(&tpd, idx[r], idx[c], dx);
                tr_add(&tpd, idx[r] + V, idx[c], dy);
            }
        }
    }
    /* out.addEntry(boundaryNode, boundaryNode, 1) for each boundary node
     * (:1028-1031); every node once. */
    for (int i = 0; i < V; ++i) if (vmask[i]) tr_add(&td, i, i, 1.0);
    for (int i = 0; i < P; ++i) if (pmask[i]) tr_add(&tp, i, i, 1.0);

    int rc = 0;
    rc |= tr_to_csr(&tm, V, V, &out->velocity_mass);
    rc |= tr_to_csr(&td, V, V, &out->diffusion);
    rc |= tr_to_csr(&tw, V, V, &out->velocity_weights);
    rc |= tr_to_csr(&tp, P, P, &out->pressure_stiffness);
    rc |= tr_to_csr(&tpw, P, P, &out->pressure_weights);
    rc |= tr_to_csr(&tvd, P, 2 * V, &out->velocity_divergence);
    rc |= tr_to_csr(&tpd, 2 * V, P, &out->pressure_divergence);
    /* velocityDivergenceMatrix *= -dtInv; pressureDivergenceMatrix *= -dt
     * (:1449-1450, :1461-1462) */
    const double dt_inv = 1.0 / dt;
    for (int64_t k = 0; k < out->velocity_divergence.nnz; ++k)
        out->velocity_divergence.val[k] *= -dt_inv;
    for (int64_t k = 0; k < out->pressure_divergence.nnz; ++k)
        out->pressure_divergence.val[k] *= -dt;
    tr_free(&tm); tr_free(&td); tr_free(&tw); tr_free(&tp); tr_free(&tpw);
    tr_free(&tvd); tr_free(&tpd);
    free(vmask);
    free(pmask);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Time step                                                           */
/* ------------------------------------------------------------------ */
void or_impose_velocity_bc(int nvb, const int32_t *vb, const double *gu,
                           const double *gv, double *u, double *v) {
    for (int k = 0; k < nvb; ++k) {
        u[vb[k]] = gu[k];
        v[vb[k]] = gv[k];
    }
}

/* Dirichlet right-hand-side patch: rhs[b] = g; rhs[c] -= W[b][c]*g for the
 * weights row of b (include/assembly.h:503-524, :583-603). */
static void bc_patch(const or_csr *W, int nb, const int32_t *bn,
                     const double *g, double *rhs, double sign) {
    for (int k = 0; k < nb; ++k) {
        const int b = bn[k];
        rhs[b] = g[k];
        for (int32_t i = W->row_start[b]; i < W->row_start[b + 1]; ++i)
            rhs[W->col[i]] -= sign * (W->val[i] * g[k]);
    }
}

/* planted mistake of the composition tests (OR_FAULT_*), 0 = none */
static int g_step_fault = OR_FAULT_NONE;

void or_set_step_fault(int fault) {
    g_step_fault = (fault > OR_FAULT_NONE && fault < OR_FAULT_COUNT) ? fault : OR_FAULT_NONE;
}

int or_get_step_fault(void) { return g_step_fault; }

static void scale_vec(int64_t n, double a, double *x) {
    for (int64_t i = 0; i < n; ++i) x[i] *= a;
}

static double rel_res(const or_cg_result *r) {
    return r->rhs_norm > 0 ? r->final_residual / r->rhs_norm : r->final_residual;
}

/* semiLagrangianSolve loop body (include/assembly.h:1516-1554):
 * advect, diffusionSolve u and v (:489-558), applyPressure (:560-712). */
int or_step(const or_matrices *m, const uint32_t *kd_nodes,
            const int32_t *leaf, const float *xy, const int32_t *el, int V,
            int P, int nvb, const int32_t *vb, const double *gu,
            const double *gv, int npb, const int32_t *pb, const double *gp,
            double dt, double rtol, int max_iter, double *u, double *v,
            double *p, or_step_stats *st) {
    const int fault = g_step_fault;
    double *ut = (double *)malloc(sizeof(double) * 2 * (size_t)V);
    double *vt = ut + V;
    double *rhs = (double *)malloc(sizeof(double) * 2 * (size_t)V);
    double *prhs = (double *)malloc(sizeof(double) * (size_t)P);
    double *du = (double *)malloc(sizeof(double) * 2 * (size_t)V);
    double *uv = (double *)malloc(sizeof(double) * 2 * (size_t)V);
    or_cg_result res;
    int status = 0, rc;
    const double vel_bc_sign = fault == OR_FAULT_VEL_BC_SIGN ? -1.0 : 1.0;
    const double p_bc_sign = fault == OR_FAULT_P_BC_SIGN ? -1.0 : 1.0;
    memset(st, 0, sizeof(*st));

    if (fault == OR_FAULT_SKIP_ADVECTION) {
        memcpy(ut, u, sizeof(double) * (size_t)V);
        memcpy(vt, v, sizeof(double) * (size_t)V);
    } else {
        or_advect_f64(kd_nodes, leaf, xy, el, V, u, v, dt, ut, vt, 0);
    }

    /* diffusion, u then v; x0 = advected velocity (:530-537) */
    or_spmv(V, m->velocity_mass.row_start, m->velocity_mass.col,
            m->velocity_mass.val, ut, rhs);
    bc_patch(&m->velocity_weights, nvb, vb, gu, rhs, vel_bc_sign);
    rc = or_cg(V, m->diffusion.row_start, m->diffusion.col, m->diffusion.val,
               rhs, ut, u, max_iter, rtol, 0.0, &res);
    st->cg_iterations[0] = res.iterations; st->cg_residual[0] = rel_res(&res);
    if (rc && !status) status = rc;
    or_spmv(V, m->velocity_mass.row_start, m->velocity_mass.col,
            m->velocity_mass.val, vt, rhs + V);
    bc_patch(&m->velocity_weights, nvb, vb, gv, rhs + V, vel_bc_sign);
    rc = or_cg(V, m->diffusion.row_start, m->diffusion.col, m->diffusion.val,
               rhs + V, vt, v, max_iter, rtol, 0.0, &res);
    st->cg_iterations[1] = res.iterations; st->cg_residual[1] = rel_res(&res);
    if (rc && !status) status = rc;

    /* pressure: rhs = velocityDivergence * [u; v] + BC; x0 = previous p */
    const int swap = fault == OR_FAULT_SWAP_COMPONENTS;
    memcpy(uv, swap ? v : u, sizeof(double) * (size_t)V);
    memcpy(uv + V, swap ? u : v, sizeof(double) * (size_t)V);
    or_spmv(P, m->velocity_divergence.row_start, m->velocity_divergence.col,
            m->velocity_divergence.val, uv, prhs);
    if (fault == OR_FAULT_DIV_SIGN) scale_vec(P, -1.0, prhs);
    if (fault == OR_FAULT_DIV_SCALE) scale_vec(P, dt * dt, prhs);
    bc_patch(&m->pressure_weights, npb, pb, gp, prhs, p_bc_sign);
    rc = or_cg(P, m->pressure_stiffness.row_start, m->pressure_stiffness.col,
               m->pressure_stiffness.val, prhs, p, p, max_iter, rtol, 0.0, &res);
    st->cg_iterations[2] = res.iterations; st->cg_residual[2] = rel_res(&res);
    if (rc && !status) status = rc;

    /* velocity correction: rhs = pressureDivergence * p; mass solves with
     * x0 = velocity input; then velocityOut += velocityIn; BC (:631-708) */
    or_spmv(2 * V, m->pressure_divergence.row_start, m->pressure_divergence.col,
            m->pressure_divergence.val, p, rhs);
    if (fault == OR_FAULT_GRAD_SIGN) scale_vec(2 * (int64_t)V, -1.0, rhs);
    if (fault == OR_FAULT_GRAD_SCALE) scale_vec(2 * (int64_t)V, 1.0 / (dt * dt), rhs);
    rc = or_cg(V, m->velocity_mass.row_start, m->velocity_mass.col,
               m->velocity_mass.val, rhs, u, du, max_iter, rtol, 0.0, &res);
    st->cg_iterations[3] = res.iterations; st->cg_residual[3] = rel_res(&res);
    if (rc && !status) status = rc;
    rc = or_cg(V, m->velocity_mass.row_start, m->velocity_mass.col,
               m->velocity_mass.val, rhs + V, v, du + V, max_iter, rtol, 0.0, &res);
    st->cg_iterations[4] = res.iterations; st->cg_residual[4] = rel_res(&res);
    if (rc && !status) status = rc;
    const double keep = fault == OR_FAULT_NO_UPDATE_ADD ? 0.0 : 1.0;
    for (int i = 0; i < V; ++i) {
        u[i] = du[i] + keep * u[i];
        v[i] = du[V + i] + keep * v[i];
    }
    or_impose_velocity_bc(nvb, vb, gu, gv, u, v);
    st->status = status;
    free(ut); free(rhs); free(prhs); free(du); free(uv);
    return status;
}
