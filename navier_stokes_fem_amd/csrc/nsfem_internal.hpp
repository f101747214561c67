// Internal types of libnsfem_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "nsfem_gpu.h"

namespace nsfem {

// ------------------------------------------------------------------ errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);
[[noreturn]] void fail(int code, const std::string &msg);

#define NSFEM_HIP(call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            ::nsfem::fail(e_ == hipErrorOutOfMemory ? NSFEM_OUT_OF_MEMORY      \
                                                    : NSFEM_HIP_ERROR,         \
                          std::string(#call) + ": " + hipGetErrorString(e_));  \
    } while (0)

#define NSFEM_REQUIRE(cond, code, msg)                                         \
    do {                                                                       \
        if (!(cond)) ::nsfem::fail((code), (msg));                             \
    } while (0)

// Run `body` and translate exceptions into a status code + last error.
template <typename F>
int guard(F &&body) {
    try {
        body();
        return NSFEM_OK;
    } catch (const Error &e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_last_error("host out of memory");
        return NSFEM_OUT_OF_MEMORY;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        return NSFEM_ERROR;
    }
}

// ------------------------------------------------------------ device memory
template <typename T>
struct DeviceArray {
    T *ptr = nullptr;
    size_t count = 0;
    DeviceArray() = default;
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), count(o.count) {
        o.ptr = nullptr;
        o.count = 0;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            count = o.count;
            o.ptr = nullptr;
            o.count = 0;
        }
        return *this;
    }
    ~DeviceArray() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void resize(size_t n) {  // contents are not preserved
        if (n == count && ptr) return;
        release();
        if (n) NSFEM_HIP(hipMalloc(&ptr, n * sizeof(T)));
        count = n;
    }
    void ensure(size_t n) {
        if (n > count) resize(n);
    }
    size_t bytes() const { return count * sizeof(T); }
};

// --------------------------------------------------------------- CSR matrix
// Rows are cut into tiles of at most kTileRows rows and kTileNnz nonzeros
// (a single longer row is a tile of its own); tiles are the unit of the
// LDS-staged CSR product and of the persistent CG's row ownership.
constexpr int kBlock = 256;
constexpr int kTileRows = 256;
constexpr int kTileNnz = 2048;
constexpr int kWaveRows = 64;   // wave tiles of the persistent CG
constexpr int kWaveNnz = 512;   // LDS room of a wave tile's products
// default cap on a wave tile's nonzeros: a tile's nonzeros from its first
// one rounded down to a multiple of 8 then fill at most 64 groups of 8 (one
// group per lane in the group-load product)
constexpr int kStageNnz = kWaveNnz - 8;
constexpr int kMaxGridBlocks = 4096;

struct Matrix {
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    DeviceArray<int32_t> row_start, col;
    DeviceArray<double> val;
    DeviceArray<float> val32;  // FP32 copy for the refined solver (lazy)
    DeviceArray<int32_t> tiles;  // [n_tiles + 1] first row of each tile
    int n_tiles = 0;
    std::vector<int32_t> host_tiles;
    DeviceArray<int32_t> wtiles;  // [n_wtiles + 1] first row of each wave tile
    DeviceArray<int4> wmeta;      // [n_wtiles] {first row, end row, first nnz, nnz}
    int n_wtiles = 0;
    std::vector<int64_t> host_wtile_prefix;  // [n_wtiles + 1] cumulative nnz + rows
    std::vector<int32_t> host_rs;            // row_start (wave tiles are rebuilt from it)
    int wave_rows = kWaveRows, wave_nnz = kWaveNnz;  // current wave-tile shape
    int max_multirow_nz = 0;  // most nonzeros of a wave tile with more than one row
    // ownership of wave tiles by persistent blocks, cached per grid size:
    // both arrays hold kMaxGridBlocks + 1 entries from the tiling on, so a
    // change of grid size copies and never allocates
    DeviceArray<int32_t> block_tiles;
    std::vector<int32_t> host_block_tiles;
    int block_tiles_grid = 0;
    // column minus the first row of the nonzero's wave tile, as int16; built
    // with the wave tiles of a square matrix when every offset fits
    DeviceArray<int16_t> dcol;
    bool dcol_ok = false;
    // the same for the product's tiles (offsets from the tile's first row)
    DeviceArray<int16_t> bcol;
    bool bcol_ok = false;
    // group loads (8 nonzeros per lane) need the 16-bit offsets (padded by
    // 8) and every multi-row tile within kStageNnz nonzeros
    bool groups_ok = false;
    // Jacobi scaling of mixed solves: 1/sqrt(diag) per row and the FP32
    // values of D^-1/2 A D^-1/2 (padded by 8), built on first use
    DeviceArray<double> jscale;
    DeviceArray<float> jval32;
    bool jacobi_built = false, jacobi_ok = false;
    bool ready = false;
    std::mutex mtx;
};

inline void matrix_release(Matrix &m) {
    m.row_start.release(); m.col.release(); m.val.release(); m.val32.release();
    m.tiles.release(); m.wtiles.release(); m.wmeta.release(); m.block_tiles.release();
    m.dcol.release();
    m.dcol_ok = false;
    m.bcol.release();
    m.bcol_ok = false;
    m.groups_ok = false;
    m.jscale.release();
    m.jval32.release();
    m.jacobi_built = m.jacobi_ok = false;
    m.block_tiles_grid = 0;
    m.ready = false;
}

struct KdTreeGpu;
inline void kdtree_release(KdTreeGpu &t);
struct KdTreeGpu {
    DeviceArray<uint32_t> nodes;     // 2 words per node
    DeviceArray<int32_t> leaf;       // leaf triangle indices
    DeviceArray<float> xy;           // 2 per mesh node
    DeviceArray<int32_t> elements;   // 6 per element
    DeviceArray<float> tri_xy;       // 6 per element: vertex coordinates
    int32_t node_count = 0;          // tree nodes
    int64_t leaf_count = 0;
    int mesh_nodes = 0;
    int elements_count = 0;
    bool ready = false;
};
inline void kdtree_release(KdTreeGpu &t) {
    t.nodes.release(); t.leaf.release(); t.xy.release(); t.elements.release(); t.tri_xy.release();
    t.ready = false;
}

constexpr int kMaxRed = 8;

struct CgColumnResult {
    int32_t iterations;
    int32_t refinements;
    int32_t status;
    int32_t pad;
    double rhs_norm;
    double initial_residual;
    double tolerance;
    double last_residual;  // sqrt of the last FP64 (true) residual^2 computed in-kernel
};

struct Simulation;
void simulation_detach(Simulation *s);  // frees its device memory, clears its device
struct GnnNet;
struct GnnGraph;
void simulation_forget_gnn(Simulation *s, GnnNet *g);  // a network is being destroyed

struct Device {
    int index = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t props{};
    int cu_count = 0;
    Matrix slots[NSFEM_MAX_SLOTS];
    KdTreeGpu tree;
    DeviceArray<float> vel32[4];  // initVelocityBuffers: u in, v in, u out, v out
    int vel_n = 0;
    // scratch for CG / dot
    DeviceArray<unsigned char> work;      // CG vectors
    DeviceArray<unsigned char> sync_mem;  // persistent-kernel barrier state
    DeviceArray<CgColumnResult> cg_res;
    DeviceArray<double> dot_partial;
    DeviceArray<double> scalars;
    // this device's copy of the library's static scratch array (the
    // no-allocation solve), looked up once at device creation
    unsigned char *scratch = nullptr;
    int wall_khz = 100000;
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::mutex mtx;  // serialises solves on this handle
    // buffers created on this device; destroying the device releases their
    // memory and detaches them, so a later nsfem_buffer_destroy stays safe
    std::set<struct Buffer *> buffers;
    std::set<Simulation *> sims;  // detached the same way
    // batched persistent CG: per-instance vectors, barrier state, results,
    // arguments and the block -> instance map
    DeviceArray<unsigned char> batch_work, batch_sync;
    DeviceArray<CgColumnResult> batch_res;
    DeviceArray<unsigned char> batch_args;
    DeviceArray<int2> batch_map;
    DeviceArray<double> batch_scal;
    std::set<GnnNet *> gnns;
    std::set<GnnGraph *> graphs;
    std::mutex buf_mtx;
    void set_current() const { NSFEM_HIP(hipSetDevice(index)); }
};

struct Buffer {
    Device *dev = nullptr;
    DeviceArray<unsigned char> mem;
};

// ------------------------------------------------------------ entry points
// (implemented in the .hip translation units, called by the C ABI layer)
void build_tiles(Matrix &m, const std::vector<int32_t> &row_start);
// Re-cut the rows into wave tiles of <= rows (<= 64) rows and <= nnz (<= 512)
// nonzeros (the persistent CG's unit of work).
void build_wave_tiles(Matrix &m, int rows, int nnz);
void ensure_block_tiles(Device &d, Matrix &m, int grid);
void ensure_val32(Device &d, Matrix &m);
bool ensure_jacobi(Device &d, Matrix &m);

void launch_spmv(Device &d, Matrix &m, const double *x, const double *lhs, double *y,
                 int mode);
// the same in FP32 from the FP32 copy of the values (ensure_val32 first)
void launch_spmv(Device &d, Matrix &m, const float *x, const float *lhs, float *y, int mode);
void launch_axpy(Device &d, int n, double a, const double *x, double *y);
void launch_axpby(Device &d, int n, double a, double b, const double *x,
                  const double *y, double *out);
// y = (accumulate ? y : 0) + sum_{k<l} a[k] * X[k * stride ...], one FMA per
// term in k order (the bits of a fill and l launch_axpy calls) in one pass;
// y2 (optional) gets the same values
void launch_combine(Device &d, int n, int l, const double *a, const double *X, size_t stride,
                    bool accumulate, double *y, double *y2 = nullptr);
void launch_dot_accumulate(Device &d, int n, const double *a, const double *b,
                           double *result);

struct CgRequest {
    Matrix *m;
    const double *b;  // [ncol][rows]
    double *x;        // [ncol][rows]
    int ncol;
    int max_iter;
    double rtol, atol;
    int precision;
    double reliable_delta = 0.5;
    int ru_gap = 0;   // mixed: forced reliable update after this many iterations (0 = 20000)
    int grid = 0;     // 0 = automatic
    bool jacobi = false;  // mixed: CG on D^-1/2 A D^-1/2 (Jacobi-preconditioned CG)
};
// Runs the persistent CG; fills stats[ncol].  Returns the worst status.
int run_cg(Device &d, const CgRequest &req, nsfem_cg_stats *stats);
// The same solve with every work vector and the reduction state in the
// device's static scratch array: no device memory is allocated.  Systems of
// rows x columns <= kNoAllocMaxValues only (NSFEM_INVALID_ARGUMENT beyond).
constexpr int64_t kNoAllocMaxValues = int64_t(1) << 20;
int run_cg_noalloc(Device &d, const CgRequest &req, nsfem_cg_stats *stats);
// Looks up this device's address of the scratch array (device current).
void scratch_init(Device &d);
// Many independent solves (same column count and precision) in one
// persistent launch; stats[2 * i + c] for request i, column c.  block_scale
// multiplies the blocks each instance gets (autotuned).  With `finish`, an
// instance that stagnated, drifted or reached its iteration limit is
// completed alone by run_cg.  Returns the worst status; *batch_seconds =
// device time of the batched launches.
int run_cg_batch(Device &d, const std::vector<CgRequest> &reqs, nsfem_cg_stats *stats,
                 int block_scale, double *batch_seconds, bool finish = true);

void launch_advect_f32(Device &d, const KdTreeGpu &t, int n, const float *u, const float *v, float dt,
                       float *uo, float *vo);
void launch_advect_f64(Device &d, const KdTreeGpu &t, int n, const double *u, const double *v, double dt,
                       double *uo, double *vo);
void launch_locate(Device &d, const KdTreeGpu &t, int n, const float *pts, int32_t *elem, float *xi,
                   float *eta, int32_t *nearest);
void build_tri_xy(Device &d, KdTreeGpu &t);

// elementwise helpers for the step driver
void launch_scatter(Device &d, int count, const int32_t *idx, const double *val,
                    double *dst);
void launch_scatter_value(Device &d, int count, const int32_t *idx, double v,
                          double *dst);  // dst[idx[i]] = v
void launch_scatter_value(Device &d, int count, const int32_t *idx, double v, float *dst);
// The CG with FP32 directions and an FP32 preconditioner: x += a p and
// r -= a q in FP64, r32 = (float) r, *rr += (r, r) (deterministic order, as
// launch_dot_accumulate)
void launch_pcg_update(Device &d, int n, double a, const float *p, const double *q, double *x,
                       double *r, float *r32, double *rr);
// y = A x with the FP32 copy of the values and FP32 x, products and sums in
// FP64 (ensure_val32 first)
void launch_spmv_acc64(Device &d, Matrix &m, const float *x, double *y);
// *result += (a, b) accumulated in FP64 with FP32 operands widened
void launch_dot_accumulate(Device &d, int n, const double *a, const float *b, double *result);
void launch_dot_accumulate(Device &d, int n, const float *a, const float *b, double *result);
void launch_copy(Device &d, int n, const float *src, float *dst);
// p = z + b p, FP32 storage
void launch_pcg_direction(Device &d, int n, const float *z, double b, float *p);
void launch_add(Device &d, int n, const double *a, double *y);  // y += a
void launch_add(Device &d, int n, const float *a, float *y);
// x += w * dinv .* r
void launch_jacobi(Device &d, int n, double w, const double *dinv, const double *r, double *x);
void launch_jacobi(Device &d, int n, double w, const float *dinv, const float *r, float *x);
// x = w * dinv .* b (the sweep from x = 0)
void launch_jacobi0(Device &d, int n, double w, const double *dinv, const double *b, double *x);
void launch_jacobi0(Device &d, int n, double w, const float *dinv, const float *b, float *x);
// y = x + w * dinv .* (b - A x) in one product (y aliases neither x nor b)
void launch_sweep(Device &d, Matrix &m, double w, const double *dinv, const double *x,
                  const double *b, double *y);
void launch_sweep(Device &d, Matrix &m, double w, const float *dinv, const float *x,
                  const float *b, float *y);
void launch_fill(Device &d, int n, double v, float *y);
// y = a rounded to FP32 / widened to FP64
void launch_convert(Device &d, int n, const double *a, float *y);
void launch_convert(Device &d, int n, const float *a, double *y);
void launch_copy(Device &d, int n, const double *src, double *dst);
void launch_spmv_affine(Device &d, Matrix &m, const double *x, const double *c,
                        double *y);  // y = A x - c

// Host-side helpers
struct HostKdTree {
    std::vector<uint32_t> nodes;
    std::vector<int32_t> leaf;
    float bbox[4];
    int max_depth;
};
void kd_build_host(int node_count, const float *xy, int elem_count,
                   const int32_t *elems, int max_depth, int min_leaf,
                   HostKdTree &out);

struct HostCsr {
    int rows = 0, cols = 0;
    std::vector<int32_t> rs, col;
    std::vector<double> val;
};

struct HostAssembly {
    HostCsr m[5];  // NSFEM_ASM_* order
    std::vector<double> c_u, c_v, c_p;
};

void assemble_host(int V, int P, const float *xy, int E, const int32_t *el, int nvb,
                   const int32_t *vb, const double *gu, const double *gv, int npb,
                   const int32_t *pb, const double *gp, double nu, double dt, int threads,
                   HostAssembly &out);

// Device assembly (assembly.hip): the same five matrices and corrections on
// the GPU, bit-identical to assemble_host.  With vorder/vpos (new -> original
// and original -> new velocity index) the velocity side comes out renumbered.
struct DeviceCsr {
    int rows = 0, cols = 0;
    int64_t nnz = 0;
    DeviceArray<int32_t> rs, col;
    DeviceArray<double> val;
};
struct DeviceAssembly {
    DeviceCsr m[5];  // NSFEM_ASM_* order
    DeviceArray<double> c_u, c_v, c_p;
};
void assemble_device(Device &d, int V, int P, const float *xy, int E, const int32_t *el, int nvb,
                     const int32_t *vb, const double *gu, const double *gv, int npb,
                     const int32_t *pb, const double *gp, double nu, double dt,
                     const int32_t *vorder, const int32_t *vpos, DeviceAssembly &out);
// Takes over a device CSR as a matrix (tiles built, FP32 copy made).
void adopt_csr(Device &d, Matrix &m, DeviceCsr &c);

// Validate and upload a host CSR into a device slot (builds the tile table).
void upload_csr(Device &d, Matrix &m, int rows, int cols, int64_t nnz, const int32_t *rs,
                const int32_t *col, const double *val);
// Validate and upload a KD tree + mesh into `t` (the device's own tree or a
// simulation's).
void upload_tree(Device &d, KdTreeGpu &t, int32_t tree_nodes, const uint32_t *words, int64_t leaf_count,
                 const int32_t *leaf, int mesh_nodes, const float *xy, int elem_count,
                 const int32_t *elems);
void launch_fill(Device &d, int n, double v, double *dst);

// ---------------------------------------------------- learned warm start
// Pressure-matrix graph (CSR without the diagonal, symmetric weights
// -A_ij / sqrt(A_ii A_jj)), node areas and coordinates scaled to the unit box.
struct GnnGraph {
    Device *dev = nullptr;
    int n = 0;
    int64_t ne = 0;
    DeviceArray<int32_t> rs, col;
    DeviceArray<double> w, area;
    DeviceArray<float2> xy;
};

// Network parameters (FP32 master + BF16 copies), Adam state, workspace.
struct GnnNet {
    Device *dev = nullptr;
    DeviceArray<float> params, grads, m, v;
    DeviceArray<float> clip;  // running gradient norm, this step's gradient scale
    DeviceArray<uint16_t> wt1, wt2, wr2, wt3, wr3;
    int adam_t = 0;
    int cap = 0;  // workspace rows
    DeviceArray<uint16_t> X1, X2, X3, h3, dz;
    DeviceArray<float> z1, z2, z3, dX, dy;
    DeviceArray<double> d, phat, scal, part;
};

GnnNet *gnn_create(Device &d, uint64_t seed);
void gnn_set_params(GnnNet &g, const float *host);
void gnn_get(GnnNet &g, const float *dev_src, float *host);
GnnGraph *gnn_graph_create(Device &d, int n, int64_t nnz, const int32_t *rs, const int32_t *col,
                           const double *val, const double *area, const float *xy);
void gnn_predict(GnnNet &g, GnnGraph &gr, const double *b, double *p_hat);
void gnn_gradients(GnnNet &g, GnnGraph &gr, const double *b, const double *p, float *host_grads,
                   double *loss);
void gnn_train_step(GnnNet &g, GnnGraph &gr, const double *b, const double *p, float lr,
                    double *loss);
// dst = src if src has the smaller residual ||b - A x|| (no host sync);
// tmp: A.rows doubles, scal2: 2 doubles
void select_better_guess(Device &d, Matrix &A, const double *b, const double *src, double *dst,
                         double *tmp, double *scal2);
// x += c dir with c = (dir . r) / (dir . A dir), r = b - A x: the step along
// dir that minimises the A-norm error (none unless dir . A dir > 0; no host
// sync); w: A.rows doubles of scratch, scal2: 2 doubles
void line_search_step(Device &d, Matrix &A, const double *r, const double *dir, double *x,
                      double *w, double *scal2);
inline void gnn_release(GnnNet &g) {
    g.params.release(); g.grads.release(); g.m.release(); g.v.release();
    g.wt1.release(); g.wt2.release(); g.wr2.release(); g.wt3.release(); g.wr3.release();
    g.X1.release(); g.X2.release(); g.X3.release(); g.h3.release(); g.dz.release();
    g.z1.release(); g.z2.release(); g.z3.release(); g.dX.release(); g.dy.release();
    g.d.release(); g.phat.release(); g.scal.release(); g.part.release();
    g.cap = 0;
}
inline void gnn_graph_release(GnnGraph &g) {
    g.rs.release(); g.col.release(); g.w.release(); g.area.release(); g.xy.release();
}

}  // namespace nsfem

struct nsfem_device {
    nsfem::Device impl;
};
struct nsfem_buffer {
    nsfem::Buffer impl;
};
struct nsfem_kdtree {
    nsfem::HostKdTree impl;
};
struct nsfem_assembly {
    nsfem::HostAssembly impl;
};
struct nsfem_gnn {
    nsfem::GnnNet *impl;
};
struct nsfem_gnn_graph {
    nsfem::GnnGraph *impl;
};
