/*
 * CPU oracle for the CG + semi-Lagrangian advection path of
 * vasil-pashov/Navier_Stokes_FEM.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product (navier_stokes_fem_amd) never links, loads or
 * calls it.
 *
 * Every function restates one piece of the reference algorithm in plain C and
 * cites the reference file:line it follows.  Floating-point conventions:
 *   - FP32 parts (KD-tree build and point location, FP32 advection) use IEEE
 *     single precision with no FMA contraction (built with -ffp-contract=off),
 *     which is what the reference's host twin computes (x86-64 SSE, no FMA).
 *   - The CG, the CSR products and the FEM assembly are re-typed to double:
 *     the reference is FP32 everywhere (gpu_cpu_shared/misc_common.cuh:8), and
 *     the plain method of this project is FP64.
 *   - Dot products use deterministic blocked sums (block 8192, partials added
 *     in block order) as the reference's CPU CG does
 *     (tex/Chapters/04_implementation_details.tex:288).
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   or_spmv / or_spmv_sub / or_dot / or_axpy / or_axpby : dense FP64 products
 *       on random small CSR matrices (closed form).
 *   or_triplet_to_csr : scipy.sparse coo->csr on random triplets.
 *   or_cg : dense Cholesky solve (numpy) on assembled cavity matrices;
 *       true residual <= rtol*||b||; A-orthogonality of directions.  Parity
 *       with the reference's CPU solver itself (SMM::ConjugateGradient of
 *       github.com/vasil-pashov/sparse_matrix_math@master, not vendored) is
 *       unpinned: its stopping rule and FP32 iterate sequence are unknown.
 *   or_kd_build : structural invariants (leaf coverage of every triangle,
 *       split = box midpoint, right-child = next free index, leaf cut-off),
 *       hand-derived trees on tiny meshes, bbox max = last node.
 *   or_kd_find_element : bit-exact against the reference's own
 *       gpu_cpu_shared/kd_tree.cuh compiled from /root/reference into
 *       oracle/_ref (golden vectors in tests/golden/), and brute force over all
 *       triangles (containment / nearest node).
 *   or_advect_* : zero velocity is the identity; P2 polynomials reproduced;
 *       partition of unity of the P2 shape functions.
 *   or_assemble : mass matrix row sums = element areas (closed form),
 *       stiffness annihilates constants and reproduces the discrete Laplacian
 *       of linear functions, divergence of a linear field, symmetry.
 *   or_step : the composition, by flows whose one-step outcome the
 *       mathematics fixes (tests/step_cases.py): a uniform flow is a fixed
 *       point; a potential flow is projected away with its pressure p dt =
 *       phi - dt|u|^2/2, errors falling under refinement; a solid-body
 *       rotation stays put with the centripetal pressure; first order in dt
 *       (Runge estimate); mirror symmetry; exact Dirichlet values.  Every
 *       planted mistake of or_set_step_fault fails at least one of them.
 */
#ifndef NSFEM_ORACLE_H
#define NSFEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Threading of the baseline.  0 = all hardware threads.              */
/* ------------------------------------------------------------------ */
void or_set_threads(int n);
int or_get_threads(void);

/* ------------------------------------------------------------------ */
/* CSR linear algebra (gpu/matrix_math.cu)                             */
/* ------------------------------------------------------------------ */
void or_spmv(int rows, const int32_t *row_start, const int32_t *col,
             const double *val, const double *x, double *y);
void or_spmv_sub(int rows, const int32_t *row_start, const int32_t *col,
                 const double *val, const double *lhs, const double *x,
                 double *y);
double or_dot(int64_t n, const double *a, const double *b);
void or_axpy(int64_t n, double a, const double *x, double *y);
void or_axpby(int64_t n, double a, double b, const double *x, const double *y,
              double *out);

/* Triplet -> CSR by counting sort, duplicates summed in input order
 * (cpp/gpu_sparse_matrix.cpp:7-49).  Output arrays are caller-allocated:
 * row_start[rows+1]; col/val sized by *nnz_out which is first queried with
 * col==NULL.  Returns 0 on success. */
int or_triplet_to_csr(int rows, int64_t count, const int32_t *ti,
                      const int32_t *tj, const double *tv, int32_t *row_start,
                      int32_t *col, double *val, int64_t *nnz_out);

typedef struct {
    int iterations;
    int converged;          /* 1 if ||r|| <= tol was reached */
    double rhs_norm;        /* ||b|| */
    double initial_residual;/* ||b - A x0|| */
    double final_residual;  /* true ||b - A x|| recomputed at exit */
    double recursive_residual; /* ||r_k|| of the CG recurrence at exit */
} or_cg_result;

/* Textbook CG (thesis Alg. CG, tex/Chapters/01_chapter_FEM.tex:644-661), in
 * the operation order of gpu/matrix_math.cu:251-287 and the set-up of
 * cpp/gpu_simulation_device.cpp:385-448.  Stopping rule:
 * ||r|| <= max(rtol*ref, atol) with ref = ||b|| (or ||r0|| if b == 0).
 * max_iter < 0 means rows (cpp/gpu_simulation_device.cpp:449-451).
 * x is always written (also when x0 already converges).  Returns 0 when
 * converged, -2 when the iteration limit is reached, -7 on breakdown. */
int or_cg(int rows, const int32_t *row_start, const int32_t *col,
          const double *val, const double *b, const double *x0, double *x,
          int max_iter, double rtol, double atol, or_cg_result *res);

/* ------------------------------------------------------------------ */
/* KD tree (cpp/kd_tree_builder.cpp, gpu_cpu_shared/kd_tree*.cuh)      */
/* ------------------------------------------------------------------ */
/* Tree node: two 32-bit words, the layout of KDNode
 * (gpu_cpu_shared/kd_tree_common.cuh:63-154):
 *   word0 = flags: low 2 bits axis (0 x, 1 y) or 2 for a leaf; high 30 bits
 *           right-child index (internal) or triangle count (leaf);
 *   word1 = float split point bits (internal) or uint32 leaf offset. */
typedef struct {
    int32_t node_count;
    uint32_t *nodes;          /* 2*node_count words */
    int64_t leaf_index_count;
    int32_t *leaf_indices;
    float bbox[4];            /* min x, min y, max x, max y */
    int32_t max_depth;        /* depth used by the build */
} or_kdtree;

/* Grid bbox as FemGrid2D::loadJSON computes it (cpp/grid.cpp:10,30) with the
 * BBox2D::expand defect (gpu_cpu_shared/misc_common.cuh:59-65): min is the
 * true minimum, max is the LAST node's coordinates. */
void or_grid_bbox(int node_count, const float *xy, float bbox[4]);

/* KDTreeBuilder::build (cpp/kd_tree_builder.cpp:81-148).  max_depth < 0:
 * min(29, round(8 + 1.3f*log(E))); min_leaf_size default 16 (:7-10). */
int or_kd_build(int node_count, const float *xy, int elem_count,
                const int32_t *elems, int max_depth, int min_leaf_size,
                or_kdtree *out);
void or_kd_free(or_kdtree *t);

/* isPointInTriagle (gpu_cpu_shared/kd_tree_common.cuh:18-61). */
int or_point_in_triangle(float px, float py, float ax, float ay, float bx,
                         float by, float cx, float cy, float *xi, float *eta);

/* KDTree::findElement (gpu_cpu_shared/kd_tree.cuh:40-141) with the nearest
 * neighbour unwinding and nearestNeghbourProcessLeaf (:149-169).
 * float_offset != 0 reproduces KDNode::getTriangleOffset returning float
 * (kd_tree_common.cuh:91,148-150): the leaf offset is rounded to float.
 * Returns the element index or -1 (then *nearest holds the closest node). */
int or_kd_find_element(const uint32_t *nodes, const int32_t *leaf_indices,
                       const float *xy, const int32_t *elems, float px,
                       float py, float *xi, float *eta, int *nearest,
                       int float_offset);

/* Batched locate for tests: one query per point. */
void or_kd_find_many(const uint32_t *nodes, const int32_t *leaf_indices,
                     const float *xy, const int32_t *elems, int n,
                     const float *pts, int32_t *element, float *xi,
                     float *eta, int32_t *nearest, int float_offset);

/* advect kernel (gpu/advection.cu:14-57), FP32 throughout. */
void or_advect_f32(const uint32_t *nodes, const int32_t *leaf_indices,
                   const float *xy, const int32_t *elems, int n,
                   const float *u, const float *v, float dt, float *u_out,
                   float *v_out, int float_offset);

/* Same point location in FP32 (start = pos - (float)vel*(float)dt), P2 shape
 * values and the interpolation sum in FP64 from the FP32 (xi, eta). */
void or_advect_f64(const uint32_t *nodes, const int32_t *leaf_indices,
                   const float *xy, const int32_t *elems, int n,
                   const double *u, const double *v, double dt, double *u_out,
                   double *v_out, int float_offset);

/* ------------------------------------------------------------------ */
/* FEM assembly (include/assembly.h), re-typed to double               */
/* ------------------------------------------------------------------ */
typedef struct {
    int rows, cols;
    int64_t nnz;
    int32_t *row_start;
    int32_t *col;
    double *val;
} or_csr;

typedef struct {
    or_csr velocity_mass;        /* V x V, no boundary elimination */
    or_csr diffusion;            /* V x V, M + dt*nu*K, velocity BC eliminated */
    or_csr velocity_weights;     /* V x V, velocityDirichletWeights (rows = BC nodes) */
    or_csr pressure_stiffness;   /* P x P, pressure BC eliminated */
    or_csr pressure_weights;     /* P x P, pressureDirichletWeights */
    or_csr velocity_divergence;  /* P x 2V, scaled by -1/dt */
    or_csr pressure_divergence;  /* 2V x P, scaled by -dt */
} or_matrices;

int or_assemble(int node_count, int pressure_count, const float *xy,
                int elem_count, const int32_t *elems, int vel_bc_count,
                const int32_t *vel_bc_nodes, int p_bc_count,
                const int32_t *p_bc_nodes, double viscosity, double dt,
                or_matrices *out);
void or_matrices_free(or_matrices *m);
void or_csr_free(or_csr *m);

/* ------------------------------------------------------------------ */
/* Time step (include/assembly.h:1477-1557, 489-712, 1559-1579)        */
/* ------------------------------------------------------------------ */
typedef struct {
    int cg_iterations[5];   /* diffusion u, diffusion v, pressure, mass u, mass v */
    double cg_residual[5];  /* true relative residual ||b-Ax||/||b|| at exit */
    int status;
} or_step_stats;

/* One update step of semiLagrangianSolve on (u, v, p) in place.  Boundary
 * values are given per Dirichlet node (already evaluated).  The velocity
 * Dirichlet lists are processed in the given order (:503-524, :698-708). */
int or_step(const or_matrices *m, const uint32_t *kd_nodes,
            const int32_t *leaf_indices, const float *xy,
            const int32_t *elems, int node_count, int pressure_count,
            int vel_bc_count, const int32_t *vel_bc_nodes,
            const double *vel_bc_u, const double *vel_bc_v, int p_bc_count,
            const int32_t *p_bc_nodes, const double *p_bc_values, double dt,
            double rtol, int max_iter, double *u, double *v, double *p,
            or_step_stats *stats);

/* Planted mistakes for the tests of or_step's composition: each value makes
 * or_step commit one plausible error, so a test can show that its checks
 * fail on it (tests only; 0 restores the faithful step).  The setting is
 * process-wide and not thread-safe. */
enum {
    OR_FAULT_NONE = 0,
    OR_FAULT_DIV_SIGN,        /* pressure rhs from +(1/dt) B u instead of -(1/dt) B u */
    OR_FAULT_GRAD_SIGN,       /* velocity correction from +dt G p instead of -dt G p */
    OR_FAULT_VEL_BC_SIGN,     /* velocity Dirichlet patch adds W g instead of subtracting */
    OR_FAULT_P_BC_SIGN,       /* pressure Dirichlet patch adds W g instead of subtracting */
    OR_FAULT_DIV_SCALE,       /* divergence scaled by -dt instead of -1/dt */
    OR_FAULT_GRAD_SCALE,      /* gradient scaled by -1/dt instead of -dt */
    OR_FAULT_NO_UPDATE_ADD,   /* velocityOut += velocityIn left out */
    OR_FAULT_SKIP_ADVECTION,  /* the diffusion starts from the unadvected velocity */
    OR_FAULT_SWAP_COMPONENTS, /* the divergence reads [v; u] instead of [u; v] */
    OR_FAULT_COUNT
};
void or_set_step_fault(int fault);
int or_get_step_fault(void);

/* imposeVelocityDirichlet (include/assembly.h:1559-1579). */
void or_impose_velocity_bc(int vel_bc_count, const int32_t *vel_bc_nodes,
                           const double *vel_bc_u, const double *vel_bc_v,
                           double *u, double *v);

#ifdef __cplusplus
}
#endif
#endif
