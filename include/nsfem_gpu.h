/*
 * nsfem_gpu.h -- C ABI of the MI355X (gfx950) CG + semi-Lagrangian advection
 * library (libnsfem_hip.so).
 *
 * This is the drop-in boundary for the GPU simulation device of
 * vasil-pashov/Navier_Stokes_FEM: every entry point below names the
 * reference interface it replaces (file:line in that repository).  Plain
 * pointers and sizes only; every function returns an int status (0 = ok,
 * like EC::ErrorCode, include/error_code.h:8-36) and the message of the last
 * failure on the calling thread is available from nsfem_last_error().
 *
 * Precision: the reference is FP32 throughout (gpu_cpu_shared/misc_common.cuh:8).
 * This library computes the plain method in FP64 (vectors, matrix values,
 * dot products).  Point location is FP32 exactly as the reference computes it
 * (mesh coordinates are float), so KD-tree element indices are bit-exact with
 * the reference.  nsfem_advect() is the reference's FP32 call verbatim.
 *
 * Ownership: the caller owns every host array and may free it when the call
 * returns (data are copied).  The device handle owns every device allocation
 * made through it; nsfem_device_destroy() frees them.  Calls on one device
 * handle are serialised on the handle's single HIP stream; host-pointer calls
 * are synchronous (they return after the result has been copied back, like
 * the reference's cuMemcpyDtoH).  Matrix uploads are thread-safe (the
 * reference uploads from three TBB tasks at once, include/assembly.h:1321,
 * 1365,1405).
 */
#ifndef NSFEM_GPU_H
#define NSFEM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define NSFEM_OK 0
#define NSFEM_ERROR (-1)                 /* EC::ErrorCode(const char*) default */
#define NSFEM_MAX_ITERATIONS_REACHED (-2)/* cpp/gpu_simulation_device.cpp:382,514 */
#define NSFEM_INVALID_ARGUMENT (-3)
#define NSFEM_HIP_ERROR (-4)
#define NSFEM_OUT_OF_MEMORY (-5)
#define NSFEM_NOT_READY (-6)             /* slot / tree / buffer not uploaded */
#define NSFEM_BREAKDOWN (-7)             /* (Ap, p) <= 0: matrix not SPD */
#define NSFEM_BARRIER_TIMEOUT (-8)       /* persistent kernel barrier gave up */
#define NSFEM_STAGNATION (-9)            /* FP32 iteration stopped reducing the FP64
                                            residual (handled inside nsfem_cg_solve) */

/* Message of the last failed call on this thread ("" if none). */
const char *nsfem_last_error(void);

/* ---------------------------------------------------------------- device */
typedef struct nsfem_device nsfem_device;
typedef struct nsfem_buffer nsfem_buffer;

/* GPUSimulationDevice::SimMatrix (include/gpu_simulation_device.h:16-23).
 * Slots 5..NSFEM_MAX_SLOTS-1 are free for the caller. */
#define NSFEM_PRESSURE_STIFFNESS 0
#define NSFEM_VELOCITY_MASS 1
#define NSFEM_DIFFUSION 2
#define NSFEM_VELOCITY_DIVERGENCE 3
#define NSFEM_PRESSURE_DIVERGENCE 4
#define NSFEM_MAX_SLOTS 16

int nsfem_get_device_count(int *count);

/* GPUSimulationDevice::init(int) (cpp/gpu_simulation_device.cpp:11-20) and
 * GPUSimulationDeviceManager::init() (:524-550): the kernels are part of the
 * library's code object, there is no module to load (replaces loadModules,
 * :22-70). */
int nsfem_device_create(int device_index, nsfem_device **out);
int nsfem_device_destroy(nsfem_device *dev);
/* Name, CU count, HBM bytes, clock (kHz) of the device behind the handle. */
int nsfem_device_info(nsfem_device *dev, char *name, int name_len, int *cu_count,
                      int64_t *hbm_bytes, int *clock_khz);
/* Block until all work issued on the handle's stream has finished. */
int nsfem_device_synchronize(nsfem_device *dev);

/* ------------------------------------------------------ matrices (CSR) */
/* uploadMatrix(SimMatrix, const SMM::TripletMatrix<float>&)
 * (cpp/gpu_simulation_device.cpp:133-140 -> GPUSparseMatrix::upload,
 * cpp/gpu_sparse_matrix.cpp:7-49): triplets are sorted by (row, col) and
 * duplicates summed in input order; the device layout is the reference's
 * CSR (rowStart[rows+1], columnIndex[nnz] ascending per row, values[nnz]). */
int nsfem_upload_matrix_triplet(nsfem_device *dev, int slot, int rows, int cols,
                                int64_t count, const int32_t *row,
                                const int32_t *col, const double *value);
/* Same slot upload from a ready CSR (columns must ascend within a row, as the
 * reference asserts at cpp/gpu_sparse_matrix.cpp:29). */
int nsfem_upload_matrix_csr(nsfem_device *dev, int slot, int rows, int cols,
                            int64_t nnz, const int32_t *row_start,
                            const int32_t *col, const double *value);
int nsfem_matrix_info(nsfem_device *dev, int slot, int *rows, int *cols, int64_t *nnz);
int nsfem_free_matrix(nsfem_device *dev, int slot);

/* ------------------------------------------------------ device buffers */
/* GPU::GPUBuffer (include/gpu_host_common.h, cpp/gpu_host_common.cpp:250-333):
 * init / uploadBuffer / downloadBuffer / RAII free. */
int nsfem_buffer_create(nsfem_device *dev, int64_t byte_size, nsfem_buffer **out);
int nsfem_buffer_destroy(nsfem_buffer *buf);
int nsfem_buffer_upload(nsfem_buffer *buf, const void *src, int64_t byte_size);
int nsfem_buffer_download(const nsfem_buffer *buf, void *dst, int64_t byte_size);
int64_t nsfem_buffer_byte_size(const nsfem_buffer *buf);
void *nsfem_buffer_device_pointer(const nsfem_buffer *buf);

/* --------------------------------------------- CSR products and BLAS-1 */
/* All vectors are FP64 device buffers.  Each call is asynchronous on the
 * handle's stream (the reference's launchers are, too). */

/* spRMult: res = A*mult (gpu_simulation_device.cpp:142-166, kernel
 * gpu/matrix_math.cu:19-51).  mult must not alias res. */
int nsfem_sp_r_mult(nsfem_device *dev, int slot, const nsfem_buffer *mult,
                    nsfem_buffer *res);
/* spRMultSub: res = lhs - A*mult; res may alias lhs (:173-199,
 * gpu/matrix_math.cu:66-99). */
int nsfem_sp_r_mult_sub(nsfem_device *dev, int slot, const nsfem_buffer *lhs,
                        const nsfem_buffer *mult, nsfem_buffer *res);
/* Device time of `repeats` back-to-back spRMult launches (res = A*mult) on
 * the handle's stream, from HIP events around them: *seconds = the mean per
 * launch.  Synchronises.  No counterpart in the reference: it measures the
 * product as the solves run it, without host launch gaps in the figure. */
int nsfem_time_sp_r_mult(nsfem_device *dev, int slot, const nsfem_buffer *mult,
                         nsfem_buffer *res, int repeats, double *seconds);
/* saxpy: y += a*x (:201-222, gpu/matrix_math.cu:106-122). */
int nsfem_saxpy(nsfem_device *dev, int n, double a, const nsfem_buffer *x,
                nsfem_buffer *y);
/* saxpby: result = a*x + b*y; result may alias y (:224-249,
 * gpu/matrix_math.cu:131-155). */
int nsfem_saxpby(nsfem_device *dev, int n, double a, double b,
                 const nsfem_buffer *x, const nsfem_buffer *y,
                 nsfem_buffer *result);
/* dotProduct: result[0] += (a, b) (:251-291, gpu/matrix_math.cu:162-194).
 * Like the reference the sum is ADDED to result[0] (the caller zeroes it);
 * unlike the reference's float atomics the sum is deterministic. */
int nsfem_dot_product(nsfem_device *dev, int n, const nsfem_buffer *a,
                      const nsfem_buffer *b, nsfem_buffer *result);

/* ------------------------------------------------- conjugate gradient */
#define NSFEM_PRECISION_FP64 0          /* plain method */
#define NSFEM_PRECISION_FP32_REFINED 1  /* FP32 SpMV/vectors, FP64 refinement */

typedef struct {
    int32_t iterations;       /* CG iterations (inner FP32 ones when refined) */
    int32_t refinements;      /* FP64 residual recomputations (refined mode) */
    int32_t status;           /* same as the return value */
    int32_t columns;          /* right-hand sides solved together */
    double rhs_norm;          /* ||b||_2 */
    double initial_residual;  /* ||b - A x0||_2 */
    double final_residual;    /* true ||b - A x||_2, recomputed in FP64 at exit */
    double tolerance;         /* the absolute tolerance the solve ran to */
    double seconds;           /* device time of the solve (HIP events) */
} nsfem_cg_stats;

/* conjugateGradient(SimMatrix, const float* b, const float* x0, float* x,
 * int maxIterations, float eps, bool useMegakernel)
 * (include/gpu_simulation_device.h:62-76; megakernel path
 * cpp/gpu_simulation_device.cpp:385-515, kernel gpu/matrix_math.cu:251-287).
 * Host arrays; absolute stopping test ||r|| < eps exactly as the reference
 * (eps^2 > ||r||^2, :440-448 and gpu/matrix_math.cu:271); max_iterations -1
 * means the row count.  use_megakernel is accepted and ignored: the
 * persistent kernel always runs.  Returns NSFEM_MAX_ITERATIONS_REACHED when
 * the limit is hit; unlike the reference x is ALWAYS written (also on an
 * early exit, :446-448) and holds the last iterate. */
int nsfem_conjugate_gradient(nsfem_device *dev, int slot, const double *b,
                             const double *x0, double *x, int max_iterations,
                             double eps, int use_megakernel);

/* Device-resident solve of A x = b for `columns` right-hand sides
 * (1 or 2; column-major buffers of rows*columns doubles: all of column 0,
 * then all of column 1) with the stopping rule
 *   ||b - A x|| <= max(rtol*||b||, atol)   (||r0|| replaces ||b|| if b = 0)
 * applied per column.  x holds x0 on entry and the solution on return.
 * precision: NSFEM_PRECISION_FP64 or NSFEM_PRECISION_FP32_REFINED.
 * stats may be NULL; with columns == 2 it must hold 2 entries. */
int nsfem_cg_solve(nsfem_device *dev, int slot, const nsfem_buffer *b,
                   nsfem_buffer *x, int columns, int max_iterations, double rtol,
                   double atol, int precision, nsfem_cg_stats *stats);

/* nsfem_cg_solve without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments, stopping rule, results and kernels as nsfem_cg_solve;
 * every work vector, the grid-reduction state and the results live in one
 * static device array of the library (about 48.4 MiB, held by each GPU the
 * library runs on until the process ends).  Limit: rows * columns <=
 * 1,048,576 (1 << 20) in either precision (an FP32 + refinement solve keeps
 * room for its FP64 continuation); larger systems return
 * NSFEM_INVALID_ARGUMENT.  Every call on one GPU shares that array: calls
 * that overlap on one GPU (from several threads or several device handles
 * of the same GPU) race and must be serialised by the caller. */
int nsfem_cg_solve_noalloc(nsfem_device *dev, int slot, const nsfem_buffer *b,
                           nsfem_buffer *x, int columns, int max_iterations,
                           double rtol, double atol, int precision,
                           nsfem_cg_stats *stats);

/* ------------------------------------------------------------ KD tree */
typedef struct nsfem_kdtree nsfem_kdtree;

/* KDTreeBuilder::buildCPUOwner / build (cpp/kd_tree_builder.cpp:17-25,
 * 81-148) on the host: midpoint splits of the bbox FemGrid2D computes
 * (cpp/grid.cpp:30 with BBox2D::expand, gpu_cpu_shared/misc_common.cuh:59-65),
 * round-robin axes, straddling triangles on both sides, leaf at
 * min_leaf_size (16 when < 0) or max_depth (min(29, round(8+1.3f*ln E))
 * when < 0).  xy: 2 floats per node; elements: 6 ints per element. */
int nsfem_kdtree_build(int node_count, const float *xy, int element_count,
                       const int32_t *elements, int max_depth, int min_leaf_size,
                       nsfem_kdtree **out);
int nsfem_kdtree_destroy(nsfem_kdtree *tree);
/* Sizes, then copies of the KDNode words (2 per node, the layout of
 * gpu_cpu_shared/kd_tree_common.cuh:63-154), leaf triangle indices, bbox. */
int nsfem_kdtree_info(const nsfem_kdtree *tree, int32_t *node_count,
                      int64_t *leaf_index_count, int32_t *max_depth, float bbox[4]);
int nsfem_kdtree_copy(const nsfem_kdtree *tree, uint32_t *node_words,
                      int32_t *leaf_indices);

/* uploadKDTree(const KDTreeCPUOwner&) (cpp/gpu_simulation_device.cpp:72-77
 * -> KDTreeCPUOwner::upload, cpp/kd_tree_builder.cpp:36-65): tree nodes,
 * leaf indices, mesh nodes (float xy) and elements (6 ints).  Also sizes
 * the velocity buffers to node_count (initVelocityBuffers, :79-89). */
int nsfem_upload_kdtree(nsfem_device *dev, const nsfem_kdtree *tree,
                        int node_count, const float *xy, int element_count,
                        const int32_t *elements);
/* Same from raw arrays (a tree built elsewhere in the reference layout). */
int nsfem_upload_kdtree_arrays(nsfem_device *dev, int32_t tree_node_count,
                               const uint32_t *node_words,
                               int64_t leaf_index_count,
                               const int32_t *leaf_indices, int node_count,
                               const float *xy, int element_count,
                               const int32_t *elements);
/* initVelocityBuffers (cpp/gpu_simulation_device.cpp:79-89). */
int nsfem_init_velocity_buffers(nsfem_device *dev, int n);

/* advect(int, const float* u, const float* v, float dt, float* uOut,
 * float* vOut) (cpp/gpu_simulation_device.cpp:91-131, kernel
 * gpu/advection.cu:14-57): host arrays, FP32 throughout, bit-exact with
 * the reference for leaf offsets < 2^24. */
int nsfem_advect(nsfem_device *dev, int n, const float *u, const float *v,
                 float dt, float *u_out, float *v_out);
/* Device-resident FP64 form: FP32 point location (start point from
 * (float)u, (float)v, (float)dt), P2 interpolation in FP64. */
int nsfem_advect_f64(nsfem_device *dev, int n, const nsfem_buffer *u,
                     const nsfem_buffer *v, double dt, nsfem_buffer *u_out,
                     nsfem_buffer *v_out);
/* Point location only (for tests and tools): element (or -1), xi, eta and,
 * when not found, the nearest node -- KDTree::findElement
 * (gpu_cpu_shared/kd_tree.cuh:40-141). Host arrays of n points. */
int nsfem_locate(nsfem_device *dev, int n, const float *points, int32_t *element,
                 float *xi, float *eta, int32_t *nearest);

/* ------------------------------------------------------- FEM assembly */
/* The five device matrices of NavierStokesAssembly::assembleAllMatrices
 * (include/assembly.h:1302-1475) in FP64 from float node coordinates, plus
 * the constant Dirichlet right-hand-side corrections, on host threads
 * (nsfem_assemble) or on the GPU (nsfem_assemble_device). */
typedef struct nsfem_assembly nsfem_assembly;
#define NSFEM_ASM_VELOCITY_MASS 0
#define NSFEM_ASM_DIFFUSION 1
#define NSFEM_ASM_PRESSURE_STIFFNESS 2
#define NSFEM_ASM_VELOCITY_DIVERGENCE 3
#define NSFEM_ASM_PRESSURE_DIVERGENCE 4

int nsfem_assemble(int node_count, int pressure_count, const float *xy,
                   int element_count, const int32_t *elements,
                   int vel_bc_count, const int32_t *vel_bc_nodes,
                   const double *vel_bc_u, const double *vel_bc_v,
                   int p_bc_count, const int32_t *p_bc_nodes,
                   const double *p_bc_values, double viscosity, double dt,
                   int threads, nsfem_assembly **out);
/* The same assembly on the GPU (one thread per matrix row, contributions
 * summed in ascending element order as the reference inserts its
 * triplets): bit-identical matrices and corrections, returned as host
 * copies in the same handle type.  A node may belong to at most 16 elements
 * (NSFEM_INVALID_ARGUMENT beyond).  Simulations assemble this way on the
 * device without the host copies (nsfem_sim_options.assembly). */
int nsfem_assemble_device(nsfem_device *dev, int node_count, int pressure_count,
                          const float *xy, int element_count, const int32_t *elements,
                          int vel_bc_count, const int32_t *vel_bc_nodes,
                          const double *vel_bc_u, const double *vel_bc_v,
                          int p_bc_count, const int32_t *p_bc_nodes,
                          const double *p_bc_values, double viscosity, double dt,
                          nsfem_assembly **out);
int nsfem_assembly_destroy(nsfem_assembly *a);
int nsfem_assembly_matrix_info(const nsfem_assembly *a, int which, int *rows,
                               int *cols, int64_t *nnz);
int nsfem_assembly_matrix_copy(const nsfem_assembly *a, int which,
                               int32_t *row_start, int32_t *col, double *val);
/* Dirichlet corrections: c_u, c_v (node_count each) and c_p
 * (pressure_count): the sums the reference subtracts per boundary entry
 * (include/assembly.h:503-524, 583-603). */
int nsfem_assembly_bc_copy(const nsfem_assembly *a, double *c_u, double *c_v,
                           double *c_p);

/* --------------------------------------------- device-resident stepping */
/* NavierStokesAssembly::semiLagrangianSolve (include/assembly.h:1477-1557)
 * with every step on the GPU: advect, two diffusion solves, the pressure
 * solve and two velocity-mass solves per step, right-hand sides and
 * Dirichlet patches computed on the device. */
typedef struct nsfem_simulation nsfem_simulation;

typedef struct {
    double viscosity;
    double dt;
    double rtol;              /* per-solve relative tolerance (1e-8) */
    int32_t max_iterations;   /* per solve, -1 = rows */
    int32_t precision;        /* NSFEM_PRECISION_* */
    int32_t pair_velocity;    /* solve u and v together (2-column CG) while the
                                 last step's column iteration counts stay within
                                 2x (+8) of each other, else one after the other */
    int32_t threads;          /* host threads for assembly / KD build, 0 = all */
    int32_t assembly;         /* NSFEM_ASSEMBLY_DEVICE (0, default) or _HOST */
    int32_t velocity_jacobi;  /* refined precision: Jacobi-preconditioned CG (CG on
                                 D^-1/2 A D^-1/2, stopped on the unscaled residual
                                 like the plain solve) for the mass solves (bit 0)
                                 and the diffusion solves (bit 1); 0 = off
                                 (default) */
} nsfem_sim_options;
#define NSFEM_ASSEMBLY_DEVICE 0   /* matrices assembled on the GPU */
#define NSFEM_ASSEMBLY_HOST 1     /* assembled on host threads, then uploaded */

typedef struct {
    int32_t steps;
    int64_t cg_iterations[5]; /* diffusion u, diffusion v, pressure, mass u, mass v */
    double max_relative_residual[5];
    double seconds;           /* device time of the steps (HIP events) */
    double advect_seconds;
    double cg_seconds[5];
    int64_t cg_refinements[5];  /* FP64 reliable updates (refined precision) */
    int64_t coarse_iterations;  /* CG iterations of the coarse-grid corrections
                                   (V-cycles when multigrid levels are added) */
    double coarse_seconds;      /* their time (device time of the CG launches;
                                   wall time of the V-cycles) */
    int64_t cg_launches[5];     /* solves whose time is in cg_seconds[k]: a paired
                                   two-column solve counts under the u entry (and
                                   a batch's shared launch under every instance) */
    int64_t cg_launch_iterations[5]; /* their iterations (a paired solve: the
                                   larger column's) */
    int64_t vcycle_solves[5];   /* of cg_launches[k], the solves by V-cycle
                                   preconditioned CG (host-driven products and
                                   smoothing sweeps, no CG-kernel launch) */
    int64_t vcycle_iterations[5]; /* of cg_launch_iterations[k], theirs */
    double vcycle_seconds[5];     /* of cg_seconds[k], theirs (wall time) */
} nsfem_step_stats;

int nsfem_simulation_create(nsfem_device *dev, int node_count,
                            int pressure_count, const float *xy,
                            int element_count, const int32_t *elements,
                            int vel_bc_count, const int32_t *vel_bc_nodes,
                            const double *vel_bc_u, const double *vel_bc_v,
                            int p_bc_count, const int32_t *p_bc_nodes,
                            const double *p_bc_values,
                            const nsfem_sim_options *options,
                            nsfem_simulation **out);
int nsfem_simulation_destroy(nsfem_simulation *sim);
/* Run `steps` update steps; stats accumulate over them (may be NULL). */
int nsfem_simulation_step(nsfem_simulation *sim, int steps, nsfem_step_stats *stats);
/* Copy fields (u, v: node_count doubles; p: pressure_count); NULL skips. */
int nsfem_simulation_get_fields(nsfem_simulation *sim, double *u, double *v, double *p);
int nsfem_simulation_set_fields(nsfem_simulation *sim, const double *u,
                                const double *v, const double *p);
/* Change rtol / max_iterations / precision / pair_velocity of an existing
 * simulation (viscosity, dt and threads are fixed at creation). */
int nsfem_simulation_set_options(nsfem_simulation *sim, const nsfem_sim_options *options);
/* Back to the initial state: zero fields with the velocity Dirichlet values
 * imposed (include/assembly.h:1503-1507). */
int nsfem_simulation_reset(nsfem_simulation *sim);
/* Batches (many independent cavities or channels of any sizes): `count`
 * simulations of one device and one precision advance `steps` steps
 * together; each solve kind (u/v diffusion as 2 columns, pressure, u/v mass
 * as 2 columns) runs for all of them in one persistent launch.  stats:
 * `count` entries (nullable), accumulated; seconds and cg_seconds are the
 * whole batch's.  A batch holding a simulation with the V-cycle
 * preconditioned pressure solve or the velocity projection of the diffusion
 * guesses is rejected (NSFEM_INVALID_ARGUMENT): both are single-simulation
 * stepping only. */
int nsfem_simulation_step_batch(nsfem_simulation *const *sims, int count, int steps,
                                nsfem_step_stats *stats);
typedef struct {
    int32_t rows_per_wave;          /* wave-tile rows chosen (64, 32 or 16) */
    int32_t nnz_per_wave;           /* wave-tile nonzero cap chosen */
    int32_t block_scale;            /* blocks per instance multiplier (1, 2, 4), pressure */
    int32_t velocity_block_scale;   /* the same for the two-column velocity solves */
    double seconds_per_iteration;   /* batched pressure iteration, best candidate */
} nsfem_batch_tuning;
/* Times the candidate wave-tile shapes and block splits on the batch's
 * pressure systems and applies the fastest to all of the batch's matrices;
 * then times the block splits of the two-column diffusion systems with that
 * shape for the velocity solves. */
int nsfem_simulation_autotune_batch(nsfem_simulation *const *sims, int count,
                                    nsfem_batch_tuning *out);

/* The simulation's own matrices (ids as the device slots:
 * NSFEM_PRESSURE_STIFFNESS ... NSFEM_PRESSURE_DIVERGENCE): shape, and
 * y = A x (lhs == NULL) or y = lhs - A x on device buffers.  Inside a
 * simulation the velocity nodes are renumbered by position (sorted by y,
 * then x) for locality; pressure vectors keep the caller's numbering, while
 * velocity-side vectors of these products are in that internal order
 * (nsfem_simulation_get_fields / set_fields map the fields back). */
int nsfem_simulation_matrix_info(nsfem_simulation *sim, int matrix, int *rows, int *cols,
                                 int64_t *nnz);
int nsfem_simulation_sp_r_mult(nsfem_simulation *sim, int matrix, const nsfem_buffer *lhs,
                               const nsfem_buffer *x, nsfem_buffer *y);

/* ------------------------------------------------------------------ */
/* Learned warm start of the pressure solve.  No counterpart in the     */
/* reference (its pressure CG starts from the previous pressure,       */
/* include/assembly.h:608-615).  A 3-layer message-passing network     */
/* over the pressure-matrix graph (hidden width 32, BF16 weights and   */
/* MFMA operands, FP32 accumulation) maps the right-hand side b to an  */
/* initial guess; CG then solves to the same tolerance as without it.  */
/* ------------------------------------------------------------------ */
typedef struct nsfem_gnn nsfem_gnn;
typedef struct nsfem_gnn_graph nsfem_gnn_graph;

/* Parameters, flat FP32 in this order: W1 [8][32], c1 [32], W2 [64][32],
 * c2 [32], W3 [64][32], c3 [32], w4 [32], c4 [1]. */
#define NSFEM_GNN_PARAM_COUNT 4481

/* A network with He-initialised weights (seeded), Adam state zero. */
int nsfem_gnn_create(nsfem_device *dev, uint64_t seed, nsfem_gnn **out);
int nsfem_gnn_destroy(nsfem_gnn *gnn);
int nsfem_gnn_get_params(nsfem_gnn *gnn, float *params);
/* Replaces the parameters and resets the Adam state. */
int nsfem_gnn_set_params(nsfem_gnn *gnn, const float *params);

/* The graph of a square pressure matrix (CSR, every row with a positive
 * diagonal), the area of each node (> 0; the lumped P1 mass) and the node
 * coordinates (2 floats per node). */
int nsfem_gnn_graph_create(nsfem_device *dev, int n, int64_t nnz,
                           const int32_t *row_start, const int32_t *col,
                           const double *val, const double *area,
                           const float *xy, nsfem_gnn_graph **out);
int nsfem_gnn_graph_destroy(nsfem_gnn_graph *graph);

/* p_hat = network(b) on device buffers of graph-n doubles. */
int nsfem_gnn_predict(nsfem_gnn *gnn, nsfem_gnn_graph *graph,
                      const nsfem_buffer *b, nsfem_buffer *p_hat);
/* Loss sum (p_hat - p)^2 / sum p^2 and its gradient (host array of
 * NSFEM_GNN_PARAM_COUNT floats, nullable); parameters unchanged. */
int nsfem_gnn_gradients(nsfem_gnn *gnn, nsfem_gnn_graph *graph,
                        const nsfem_buffer *b, const nsfem_buffer *p,
                        float *grads, double *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on one (b, p) pair, the
 * gradient first scaled down to a norm of at most 4 times the running
 * average (decay 0.95) of earlier steps' norms; the loss before the step is
 * written to *loss when loss is not NULL (that synchronises; NULL keeps the
 * step asynchronous).  nsfem_gnn_set_params resets the Adam moments and
 * the running norm. */
int nsfem_gnn_train_step(nsfem_gnn *gnn, nsfem_gnn_graph *graph,
                         const nsfem_buffer *b, const nsfem_buffer *p,
                         float learning_rate, double *loss);

/* Initial guess of a simulation's pressure solve. */
#define NSFEM_WARM_PREVIOUS 0            /* previous pressure (the reference) */
#define NSFEM_WARM_ZERO 1                /* zero */
#define NSFEM_WARM_LEARNED 2             /* network(b) */
#define NSFEM_WARM_LEARNED_OR_PREVIOUS 3 /* the one with the smaller residual */
#define NSFEM_WARM_LEARNED_CORRECTION 4  /* p_prev + network(b - A p_prev) if its
                                            residual is smaller, else p_prev */
#define NSFEM_WARM_EXTRAPOLATE 5         /* 2 p_n - p_(n-1) if its residual is
                                            smaller, else p_n (no network) */
#define NSFEM_WARM_PROJECTION 6          /* A-norm projection of the solution onto
                                            the span of past pressures (Fischer);
                                            basis depth NSFEM_PROJECTION_DEPTH */
#define NSFEM_WARM_LEARNED_SEARCH 7      /* p_prev + c network(b - A p_prev), c the
                                            A-norm-optimal step along the prediction
                                            (never worse than p_prev in the A-norm) */
#define NSFEM_WARM_LEARNED_PROJECTION 8  /* the same step from the projection guess
                                            (mode 6) instead of p_prev */
int nsfem_simulation_set_warm_start(nsfem_simulation *sim, int mode, nsfem_gnn *gnn);
/* Initial guess of the diffusion solves: the advected velocity (the
 * reference's, include/assembly.h:1521-1531) plus the A-norm projection of
 * the correction onto the span of the last `depth` corrections of each
 * velocity column (Fischer's projection on the diffusion matrix); 0 = off.
 * Costs 2 x (2 depth + 3) velocity vectors of device memory.  No
 * counterpart in the reference. */
int nsfem_simulation_set_velocity_projection(nsfem_simulation *sim, int depth);
/* Initial guess of the mass solves (the velocity corrections du, M du =
 * B_p^T p): the A-norm projection of du onto the span of the last `depth`
 * solutions of each velocity column (Fischer's projection on the mass
 * matrix) instead of the reference's guess, the velocity itself
 * (include/assembly.h:642-656); 0 = off (the reference's).  The solves stop
 * at the same relative residual.  Costs 2 x (2 depth + 3) velocity vectors
 * of device memory; single-simulation stepping only.  No counterpart in the
 * reference. */
int nsfem_simulation_set_mass_projection(nsfem_simulation *sim, int depth);
/* Coarse-grid correction of the pressure solve's initial guess, applied
 * after the warm-start mode's guess x0: with the prolongation P (one row per
 * pressure node, coarse_count columns, CSR with ascending columns) and
 * A_c = P^T A P (formed here, on the host), A_c e = P^T (b - A x0) is solved
 * by CG from zero to relative residual coarse_rtol, and x0 += P e (pressure
 * Dirichlet values imposed again; a coarse solve that does not converge
 * leaves x0 unchanged).  The pressure CG then runs from it to the same
 * tolerance as from any guess.  coarse_count = 0 removes the coarse space.
 * No counterpart in the reference, whose pressure CG starts from the
 * previous pressure (include/assembly.h:608-615).
 * Setting or removing the coarse space drops its levels and switches the
 * V-cycle preconditioner off; a call that fails part way (an invalid P, a
 * zero coarse diagonal, device memory) leaves no coarse space. */
int nsfem_simulation_set_coarse_space(nsfem_simulation *sim, int coarse_count, int64_t nnz,
                                      const int32_t *row_start, const int32_t *col,
                                      const double *val, double coarse_rtol);
/* Adds a multigrid level below the deepest coarse level so far (the coarse
 * space's, then each added one): P has one row per node of the level above,
 * `count` columns (CSR, ascending columns), and the new level's matrix is
 * P^T A_above P.  With levels present the coarse system is solved, to the
 * same coarse_rtol, by V-cycles (damped Jacobi, 2 sweeps before and after
 * on each level, CG to 1e-10 on the deepest) instead of CG.  Setting the
 * coarse space again removes the levels.  No counterpart in the reference. */
int nsfem_simulation_add_coarse_level(nsfem_simulation *sim, int count, int64_t nnz,
                                      const int32_t *row_start, const int32_t *col,
                                      const double *val);
/* enable != 0: the pressure solve becomes CG preconditioned by one V-cycle
 * over the fine pressure grid, the coarse space and its levels (damped
 * Jacobi, 2 sweeps before and after on every level), in FP64 whatever the
 * precision option, to the same relative residual, from the warm-start
 * mode's guess (the coarse correction of the guess is left out: the
 * preconditioner holds it).  Needs a coarse space; removing the coarse
 * space switches it off; single-simulation stepping only (a batch holding
 * such a simulation is rejected).  enable = 0: back to the plain CG.  No
 * counterpart in the reference, whose pressure CG is unpreconditioned
 * (include/assembly.h:608-615). */
int nsfem_simulation_set_pressure_multigrid(nsfem_simulation *sim, int enable);
/* Multigrid hierarchy below the velocity (P2) space for the diffusion
 * solves.  P has one row per velocity node, in the caller's node order, and
 * coarse_count columns (CSR, ascending columns); the rows of velocity
 * Dirichlet nodes must be empty (so every coarse correction vanishes on the
 * velocity boundary); A_1 = P^T A_diffusion P is formed here, on the host.
 * Setting it switches the diffusion solves to mode 1 of
 * nsfem_simulation_set_velocity_multigrid; coarse_count = 0 removes the
 * hierarchy (and the mode).  A call that fails part way leaves none.  No
 * counterpart in the reference, whose GPU diffusion CG is unpreconditioned
 * (include/assembly.h:528-538; its CPU CG can take an IC0 preconditioner,
 * :539-551). */
int nsfem_simulation_set_velocity_coarse_space(nsfem_simulation *sim, int coarse_count,
                                               int64_t nnz, const int32_t *row_start,
                                               const int32_t *col, const double *val);
/* Adds a level below the deepest level of the velocity hierarchy: P has one
 * row per node of the level above and `count` columns; the new level's
 * matrix is P^T A_above P (a level of at most 600 nodes is solved by its
 * dense inverse, else by CG to 1e-10).  No counterpart in the reference. */
int nsfem_simulation_add_velocity_coarse_level(nsfem_simulation *sim, int count, int64_t nnz,
                                               const int32_t *row_start, const int32_t *col,
                                               const double *val);
/* Which diffusion columns (u, v) are solved by CG preconditioned with one
 * V-cycle over the velocity nodes and the hierarchy (damped Jacobi, 2 sweeps
 * before and after on every level, z zeroed on the velocity Dirichlet rows),
 * in FP64 whatever the precision option, to the same relative residual from
 * the same guess: 0 none, 1 each column the way (plain or V-cycle) whose
 * solves of that column have taken less time so far, the slower way timed
 * again every 16 steps, 2 both columns.  Needs the hierarchy for modes 1 and
 * 2; while it is on the diffusion columns are never paired, and batches
 * reject the simulation.  No counterpart in the reference. */
int nsfem_simulation_set_velocity_multigrid(nsfem_simulation *sim, int mode);
/* Copies the last step's pressure right-hand side and pressure into device
 * buffers (either may be NULL): training pairs for the network. */
int nsfem_simulation_copy_pressure_system(nsfem_simulation *sim, nsfem_buffer *rhs,
                                          nsfem_buffer *p);

#ifdef __cplusplus
}
#endif
#endif
