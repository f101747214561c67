// Device-resident time stepping: NavierStokesAssembly::semiLagrangianSolve
// (include/assembly.h:1477-1557) with diffusionSolve (:489-558) and
// applyPressure (:560-712), every operation on the GPU.
//
// Per step (the reference's order):
//   1. (ut, vt) = advect(u, v)                                   (:1521-1526)
//   2. rhs_u = M ut - c_u, rhs_u[b] = g_u;  same for v           (:501-524)
//      u = CG(D, rhs_u, x0 = ut);  v = CG(D, rhs_v, x0 = vt)     (:530-537)
//   3. rhs_p = Bdiv [u; v] - c_p, rhs_p[b] = g_p                 (:580-603)
//      p = CG(Kp, rhs_p, x0 = p)                                 (:608-615)
//   4. rhs = Pdiv p;  du = CG(M, rhs_u, x0 = u), dv likewise     (:637-657)
//   5. u += du, v += dv; u[b] = g_u, v[b] = g_v                  (:687-708)
// The reference runs steps 2 and 4 as two separate solves; with
// pair_velocity they run as one 2-column CG (same iterates per column,
// the matrix is read once for both) while the two columns need similar
// iteration counts (pair_columns).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nsfem_internal.hpp"
#include "spd_inverse.hpp"

namespace nsfem {

struct Simulation;
void simulation_reset(Simulation &s);

// A-orthonormal basis X of past solutions of one matrix, with A X, for the
// projection guess (a pressure solve, and the coarse solves of the coarse
// correction); w, dx, guess: vectors of the matrix's size; s: dot results.
struct ProjBasis {
    int n = 0, dim = 0, max = 0;
    DeviceArray<double> x, ax, w, dx, guess, s;
    void reserve(int rows, int depth) {
        n = rows;
        max = depth;
        dim = 0;
        x.resize((size_t)depth * rows);
        ax.resize((size_t)depth * rows);
        w.resize(rows);
        dx.resize(rows);
        guess.resize(rows);
        s.resize((size_t)depth + 1);
    }
    void release() {
        x.release();
        ax.release();
        w.release();
        dx.release();
        guess.release();
        s.release();
        n = dim = max = 0;
    }
};

// A level below the coarse grid of the coarse correction (multigrid coarse
// solve): A_l = P_l^T A_(l-1) P_l, the prolongation P_l from this level to the
// one above and its transpose, 1/diag(A_l), and the level's right-hand
// side, solution and residual vectors.
struct MgLevel {
    int n = 0;
    Matrix A, P, PT;
    DeviceArray<double> dinv, b, x, r;
    DeviceArray<float> dinv32, b32, x32, r32;  // the FP32 cycle's (built on first use)
};

// A level's vectors in a cycle's precision.
template <typename T>
struct LevelVectors;
template <>
struct LevelVectors<double> {
    const double *dinv;
    double *b, *x, *r;
    explicit LevelVectors(MgLevel &l) : dinv(l.dinv.ptr), b(l.b.ptr), x(l.x.ptr), r(l.r.ptr) {}
};
template <>
struct LevelVectors<float> {
    const float *dinv;
    float *b, *x, *r;
    explicit LevelVectors(MgLevel &l)
        : dinv(l.dinv32.ptr), b(l.b32.ptr), x(l.x32.ptr), r(l.r32.ptr) {}
};

// The levels below a system (lv[0] the next coarser one), the host CSR of
// the deepest level's matrix (for the next level's Galerkin product) and,
// for a deepest level of at most kMgDenseMax nodes, its dense inverse.
struct MgLevels {
    std::vector<std::unique_ptr<MgLevel>> lv;
    Matrix deep_inv;  // as CSR
    bool deep_ok = false;
    std::vector<int32_t> deep_rs, deep_col;
    std::vector<double> deep_val;
    bool fp32_ok = false;  // the levels' FP32 vectors and values are built
    void release() {
        lv.clear();
        matrix_release(deep_inv);
        deep_ok = false;
        fp32_ok = false;
        deep_rs.clear();
        deep_col.clear();
        deep_val.clear();
    }
};

struct Simulation {
    Device *dev = nullptr;
    int V = 0, P = 0;
    nsfem_sim_options opt{};
    DeviceArray<double> uv, utv, rhs, prhs, p, du, c_uv, c_p;
    DeviceArray<int32_t> vb_idx, pb_idx;
    DeviceArray<double> vb_u, vb_v, pb_val;
    int nvb = 0, npb = 0;
    // the simulation's own matrices, indexed like the device slots
    // (NSFEM_PRESSURE_STIFFNESS ... NSFEM_PRESSURE_DIVERGENCE)
    Matrix mats[5];
    KdTreeGpu tree;  // the simulation's own KD tree and mesh
    // initial guess of the pressure solve (NSFEM_WARM_*)
    int warm_mode = NSFEM_WARM_PREVIOUS;
    GnnNet *gnn = nullptr;      // not owned
    GnnGraph *graph = nullptr;  // owned, built on first use of a learned mode
    std::vector<double> p_area;  // lumped P1 area of each pressure node
    std::vector<float> p_xy;     // coordinates of the pressure nodes
    DeviceArray<double> pguess, ptmp, pscal, pw;
    DeviceArray<double> p_older, p_keep;  // p_(n-1) and p_n for the extrapolation
    bool have_older = false;
    // projection warm start of the pressure solve
    ProjBasis proj;
    // blocks per instance in batched solves (autotuned): pressure, and the
    // two-column velocity solves (their P2 matrices have ~4x the rows)
    int batch_block_scale = 1;
    int batch_block_scale_v = 1;
    // coarse-grid correction of the pressure guess: the coarse space is
    // mg.lv[0] (prolongation P (P x nc), its transpose, A_c = P^T A P, 1/diag,
    // and its vectors b (the restricted residual), x (the correction) and r);
    // a pressure-sized residual vector; the last correction's CG iterations
    // and seconds
    bool has_coarse = false;
    int nc = 0;
    double coarse_rtol = 1e-6;
    DeviceArray<double> cy;
    ProjBasis cproj;  // projection guess of the coarse solves
    // the coarse space (lv[0]) and the multigrid levels below it (none: the
    // coarse system is solved by CG)
    MgLevels mg;
    // pressure CG preconditioned by one V-cycle over the fine grid, the
    // coarse space and its levels: 1/diag of the pressure matrix, the CG
    // vectors r, z, p, q, a smoothing residual t and a dot scratch, and the
    // FP32 cycle's fine-level vectors
    bool pmg = false;
    DeviceArray<double> pdinv, pr, pz, pp, pq, pt, pdot;
    DeviceArray<float> pdinv32, pb32, pz32, pt32, pp32;
    // the FP32 cycles of the pressure and the diffusion solves (fixed
    // kernels on fixed vectors), captured in a HIP graph on their first run
    // and replayed; dropped with their hierarchies
    hipGraphExec_t pcycle_graph = nullptr, vcycle_graph = nullptr;
    // diffusion CG preconditioned by one V-cycle over the velocity nodes,
    // the velocity coarse space (vmg.lv[0]) and the levels below it: mode
    // (0 off, 1 per column by measured cost, 2 every column), 1/diag of the
    // diffusion matrix, the CG vectors r, z, p, q, the fine smoothing
    // residual t and a dot scratch; per column the measured seconds of its
    // solves each way (exponential averages; 0 = not measured) and the steps
    // since each way last ran
    MgLevels vmg;
    int vmg_mode = 0;
    DeviceArray<double> vdinv, vr, vz, vp, vq, vt, vdot;
    DeviceArray<float> vdinv32, vb32, vz32, vt32;  // the FP32 cycle's fine level
    DeviceArray<float> vp32;                       // the FP32 CG direction
    double vmg_sec[2][2] = {{0.0, 0.0}, {0.0, 0.0}};  // [column][plain, V-cycle]
    int vmg_age[2][2] = {{0, 0}, {0, 0}};
    // projection guess of the diffusion solves' corrections u - u~ (u~ = the
    // advected velocity), one basis per column; vdel: the guessed, then the
    // true, corrections
    ProjBasis vproj[2];
    // projection guess of the mass solves' velocity corrections du, one
    // basis per column (empty: the reference's guess, the velocity itself)
    ProjBasis mproj[2];
    DeviceArray<double> vdel;
    int64_t coarse_iters = 0;
    double coarse_seconds = 0.0;
    // last step's iteration counts of the u and v columns of the diffusion
    // (0) and mass (1) solves: they decide whether the next step pairs them
    int64_t vel_iters[2][2] = {{0, 0}, {0, 0}};
    // measured device seconds per iteration of the diffusion (0) and mass
    // (1) solves run one column at a time (sec1: both launches over their
    // iterations) and as one two-column launch (sec2: over its iterations);
    // 0 = not measured yet
    double pair_sec1[2] = {0.0, 0.0}, pair_sec2[2] = {0.0, 0.0};
    // steps since each way (0: apart, 1: paired) last ran, per system
    int pair_age[2][2] = {{0, 0}, {0, 0}};
    // internal order of the velocity nodes (velocity_order): vorder[new] =
    // caller's index, vpos[caller's index] = new
    std::vector<int32_t> vorder, vpos;
};

namespace {

// Last listing of a node wins, as in the reference's assignment loop.
void last_wins(int n, const int32_t *idx, const double *a, const double *b,
               std::vector<int32_t> &oi, std::vector<double> &oa, std::vector<double> &ob) {
    std::unordered_map<int32_t, int> pos;
    for (int k = 0; k < n; ++k) {
        auto it = pos.find(idx[k]);
        if (it == pos.end()) {
            pos[idx[k]] = (int)oi.size();
            oi.push_back(idx[k]);
            oa.push_back(a[k]);
            if (b) ob.push_back(b[k]);
        } else {
            oa[it->second] = a[k];
            if (b) ob[it->second] = b[k];
        }
    }
}

template <typename T>
void put(Device &d, DeviceArray<T> &dst, const T *src, size_t n) {
    dst.resize(n);
    if (n)
        NSFEM_HIP(hipMemcpyAsync(dst.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, d.stream));
}

}  // namespace

namespace {

// Velocity nodes in lattice order, sorted by (y, x).  The reference numbers
// the P2 nodes vertices first and edge midpoints after them, so a row of a
// velocity matrix reaches two distant index blocks; sorted by position,
// every row's columns lie within about two lattice rows of it (+-8,196 at the
// 2048^2 cavity): the gathers stay local and the 16-bit column offsets fit
// (the 2048^2 diffusion CG 765 -> 669 us per mixed iteration, its product
// 561 -> 497 us).  The pressure nodes keep the caller's numbering; the
// simulation maps the velocity fields at its interface.  Returns new -> old.
std::vector<int32_t> velocity_order(int V, const float *xy) {
    std::vector<int32_t> order((size_t)V);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        const float ya = xy[2 * (size_t)a + 1], yb = xy[2 * (size_t)b + 1];
        if (ya != yb) return ya < yb;
        const float xa = xy[2 * (size_t)a], xb = xy[2 * (size_t)b];
        if (xa != xb) return xa < xb;
        return a < b;
    });
    return order;
}

// Renumbers a CSR: new row r is old row row_old[r] (rows kept when empty),
// column c becomes col_new[c] (kept when empty); columns ascend again within
// each row.  Rows are split over `threads` workers.
void permute_csr(HostCsr &c, const std::vector<int32_t> &row_old,
                 const std::vector<int32_t> &col_new, int threads) {
    HostCsr o;
    o.rows = c.rows;
    o.cols = c.cols;
    o.rs.assign((size_t)c.rows + 1, 0);
    for (int r = 0; r < c.rows; ++r) {
        const int src = row_old.empty() ? r : row_old[r];
        o.rs[r + 1] = o.rs[r] + (c.rs[src + 1] - c.rs[src]);
    }
    o.col.resize(c.col.size());
    o.val.resize(c.val.size());
    const int T = std::max(1, std::min(threads, std::max(1, c.rows / 4096)));
    auto work = [&](int t) {
        std::vector<std::pair<int32_t, double>> row;
        const int r0 = (int)((int64_t)c.rows * t / T), r1 = (int)((int64_t)c.rows * (t + 1) / T);
        for (int r = r0; r < r1; ++r) {
            const int src = row_old.empty() ? r : row_old[r];
            row.clear();
            for (int32_t e = c.rs[src]; e < c.rs[src + 1]; ++e)
                row.emplace_back(col_new.empty() ? c.col[e] : col_new[c.col[e]], c.val[e]);
            std::sort(row.begin(), row.end(),
                      [](const std::pair<int32_t, double> &a, const std::pair<int32_t, double> &b) {
                          return a.first < b.first;
                      });
            int32_t k = o.rs[r];
            for (const auto &e : row) {
                o.col[k] = e.first;
                o.val[k] = e.second;
                ++k;
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    c = std::move(o);
}

}  // namespace

Simulation *simulation_create(Device &d, int V, int P, const float *xy, int E,
                              const int32_t *el, int nvb, const int32_t *vb, const double *gu,
                              const double *gv, int npb, const int32_t *pb, const double *gp,
                              const nsfem_sim_options &opt) {
    NSFEM_REQUIRE(opt.dt > 0 && opt.viscosity >= 0, NSFEM_INVALID_ARGUMENT,
                  "simulation: dt must be > 0 and viscosity >= 0");
    NSFEM_REQUIRE(opt.precision == NSFEM_PRECISION_FP64 ||
                      opt.precision == NSFEM_PRECISION_FP32_REFINED,
                  NSFEM_INVALID_ARGUMENT, "simulation: unknown precision");
    NSFEM_REQUIRE(opt.assembly == NSFEM_ASSEMBLY_DEVICE || opt.assembly == NSFEM_ASSEMBLY_HOST,
                  NSFEM_INVALID_ARGUMENT, "simulation: unknown assembly option");
    const bool on_device = opt.assembly == NSFEM_ASSEMBLY_DEVICE;
    const auto t_start = std::chrono::steady_clock::now();
    auto lap = [&t_start] {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    };
    double t_asm = 0.0, t_tree = 0.0, t_order = 0.0;
    // renumber the velocity nodes (velocity_order): matrices, right-hand-side
    // corrections, Dirichlet nodes, coordinates and elements
    const int threads = opt.threads > 0 ? opt.threads
                                        : (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<int32_t> vorder = velocity_order(V, xy), vpos((size_t)V);
    for (int i = 0; i < V; ++i) vpos[vorder[i]] = i;
    t_order = lap();
    HostAssembly A;
    DeviceAssembly DA;
    if (on_device) {
        assemble_device(d, V, P, xy, E, el, nvb, vb, gu, gv, npb, pb, gp, opt.viscosity, opt.dt,
                        vorder.data(), vpos.data(), DA);
    } else {
        assemble_host(V, P, xy, E, el, nvb, vb, gu, gv, npb, pb, gp, opt.viscosity, opt.dt,
                      opt.threads, A);
    }
    t_asm = lap();
    HostKdTree tree;
    kd_build_host(V, xy, E, el, -1, -1, tree);
    t_tree = lap();
    if (!on_device) {
        NSFEM_REQUIRE(A.m[0].rows == V && A.m[0].cols == V && A.m[1].rows == V &&
                          A.m[1].cols == V && A.m[3].rows == P && A.m[3].cols == 2 * V &&
                          A.m[4].rows == 2 * V && A.m[4].cols == P,
                      NSFEM_ERROR, "simulation: unexpected matrix shapes");
        std::vector<int32_t> order2(2 * (size_t)V), pos2(2 * (size_t)V);
        for (int i = 0; i < V; ++i) {
            order2[i] = vorder[i];
            order2[V + i] = V + vorder[i];
            pos2[i] = vpos[i];
            pos2[V + i] = V + vpos[i];
        }
        const std::vector<int32_t> none;
        permute_csr(A.m[0], vorder, vpos, threads);
        permute_csr(A.m[1], vorder, vpos, threads);
        permute_csr(A.m[3], none, pos2, threads);
        permute_csr(A.m[4], order2, none, threads);
        std::vector<double> cu((size_t)V), cv((size_t)V);
        for (int i = 0; i < V; ++i) {
            cu[i] = A.c_u[vorder[i]];
            cv[i] = A.c_v[vorder[i]];
        }
        A.c_u.swap(cu);
        A.c_v.swap(cv);
    }
    std::vector<float> xy_in((size_t)2 * V);
    for (int i = 0; i < V; ++i) {
        xy_in[2 * (size_t)i] = xy[2 * (size_t)vorder[i]];
        xy_in[2 * (size_t)i + 1] = xy[2 * (size_t)vorder[i] + 1];
    }
    std::vector<int32_t> el_in((size_t)6 * E);
    for (size_t k = 0; k < el_in.size(); ++k) el_in[k] = vpos[el[k]];
    d.set_current();
    std::unique_ptr<Simulation> owner(new Simulation());  // freed if an upload throws
    Simulation *s = owner.get();
    s->dev = &d;
    const int slot_of[5] = {NSFEM_VELOCITY_MASS, NSFEM_DIFFUSION, NSFEM_PRESSURE_STIFFNESS,
                            NSFEM_VELOCITY_DIVERGENCE, NSFEM_PRESSURE_DIVERGENCE};
    for (int k = 0; k < 5; ++k) {
        if (on_device) {
            adopt_csr(d, s->mats[slot_of[k]], DA.m[k]);
            continue;
        }
        HostCsr &c = A.m[k];
        upload_csr(d, s->mats[slot_of[k]], c.rows, c.cols, (int64_t)c.col.size(), c.rs.data(),
                   c.col.data(), c.val.data());
        HostCsr().rs.swap(c.rs);
        std::vector<int32_t>().swap(c.col);
        std::vector<double>().swap(c.val);
    }
    upload_tree(d, s->tree, (int32_t)(tree.nodes.size() / 2), tree.nodes.data(), (int64_t)tree.leaf.size(),
                tree.leaf.data(), V, xy_in.data(), E, el_in.data());

    s->V = V;
    s->P = P;
    s->opt = opt;
    s->uv.resize(2 * (size_t)V);
    s->utv.resize(2 * (size_t)V);
    s->rhs.resize(2 * (size_t)V);
    s->du.resize(2 * (size_t)V);
    s->prhs.resize(P);
    s->p.resize(P);
    if (on_device) {
        s->c_uv.resize(2 * (size_t)V);
        launch_copy(d, V, DA.c_u.ptr, s->c_uv.ptr);
        launch_copy(d, V, DA.c_v.ptr, s->c_uv.ptr + V);
        s->c_p = std::move(DA.c_p);
    } else {
        std::vector<double> cuv(2 * (size_t)V);
        std::copy(A.c_u.begin(), A.c_u.end(), cuv.begin());
        std::copy(A.c_v.begin(), A.c_v.end(), cuv.begin() + V);
        put(d, s->c_uv, cuv.data(), cuv.size());
        put(d, s->c_p, A.c_p.data(), A.c_p.size());
    }
    std::vector<int32_t> vi, pi;
    std::vector<double> vu, vv, pv, dummy;
    last_wins(nvb, vb, gu, gv, vi, vu, vv);
    for (int32_t &node : vi) node = vpos[node];
    last_wins(npb, pb, gp, nullptr, pi, pv, dummy);
    s->nvb = (int)vi.size();
    s->npb = (int)pi.size();
    put(d, s->vb_idx, vi.data(), vi.size());
    put(d, s->vb_u, vu.data(), vu.size());
    put(d, s->vb_v, vv.data(), vv.size());
    put(d, s->pb_idx, pi.data(), pi.size());
    put(d, s->pb_val, pv.data(), pv.size());
    s->p_area.assign(P, 0.0);
    for (int e = 0; e < E; ++e) {
        const int32_t *t = el + 6 * (size_t)e;
        const double ax = xy[2 * t[0]], ay = xy[2 * t[0] + 1];
        const double area = 0.5 * std::fabs((xy[2 * t[1]] - ax) * (xy[2 * t[2] + 1] - ay) -
                                            (xy[2 * t[2]] - ax) * (xy[2 * t[1] + 1] - ay));
        for (int k = 0; k < 3; ++k) s->p_area[t[k]] += area / 3.0;
    }
    s->p_xy.assign(xy, xy + 2 * (size_t)P);
    s->vorder.swap(vorder);
    s->vpos.swap(vpos);
    simulation_reset(*s);
    if (std::getenv("NSFEM_SETUP_VERBOSE"))
        std::fprintf(stderr,
                     "[nsfem] simulation set-up: velocity order %.3f s, assembly (%s) %.3f s, "
                     "KD tree %.3f s, tiles/uploads %.3f s, total %.3f s\n",
                     t_order, on_device ? "device" : "host", t_asm - t_order, t_tree - t_asm,
                     lap() - t_tree, lap());
    {
        std::lock_guard<std::mutex> lk(d.buf_mtx);
        d.sims.insert(s);
    }
    return owner.release();
}

// zero initial field with the velocity Dirichlet values (:1503-1507)
void simulation_reset(Simulation &s) {
    Device &d = *s.dev;
    d.set_current();
    s.have_older = false;
    s.proj.dim = 0;
    s.cproj.dim = 0;
    s.vproj[0].dim = s.vproj[1].dim = 0;
    s.mproj[0].dim = s.mproj[1].dim = 0;
    for (auto &row : s.vel_iters) row[0] = row[1] = 0;
    for (int k = 0; k < 2; ++k) {
        s.pair_sec1[k] = s.pair_sec2[k] = 0.0;
        s.pair_age[k][0] = s.pair_age[k][1] = 0;
    }
    launch_fill(d, 2 * s.V, 0.0, s.uv.ptr);
    launch_fill(d, s.P, 0.0, s.p.ptr);
    launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_u.ptr, s.uv.ptr);
    launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_v.ptr, s.uv.ptr + s.V);
    NSFEM_HIP(hipStreamSynchronize(d.stream));
}

namespace {

struct SolveOut {
    int64_t iters[2];
    int64_t refines[2];
    double rel[2];
    double seconds;
};

// Reliable-update policy of the mixed solver per system.  The pressure
// Poisson solve keeps an update on every halving of the residual: from a
// cold start on the 2048^2 cavity a 10x or 100x threshold lets the FP32
// recursion stall (no convergence in 40 000 iterations, against 18 831 with
// halving).  The velocity systems (diffusion, mass) are well conditioned and
// warm-started: a 100x threshold with a forced update every 1000 iterations
// cuts their updates 4-5 fold and their time by 20-40% on that cavity.  The
// mass solves (Jacobi-scaled, ~10 iterations a step from the projection
// guess) take a 10^4x threshold: at 2048^2 their FP64 residual passes fall
// from 3.5 (0.01) to 2.9 (10^-3) and 2.6 (10^-4) a solve at the same
// iteration counts; on the two-column single-reduction kernel 10^-4 cuts
// the paired launch from 12.51 to 11.97 ms (10^-5 measures the same).
// Tuning override of both velocity thresholds: NSFEM_VELOCITY_RDELTA.
void set_update_policy(CgRequest &r, int slot) {
    if (slot == NSFEM_PRESSURE_STIFFNESS) return;
    static const double env = [] {
        const char *e = std::getenv("NSFEM_VELOCITY_RDELTA");
        return e ? std::atof(e) : 0.0;
    }();
    r.reliable_delta = env > 0.0 && env < 1.0 ? env : slot == NSFEM_VELOCITY_MASS ? 1e-4 : 0.01;
    r.ru_gap = 1000;
}

void solve(Simulation &s, int slot, const double *b, double *x, int ncol, SolveOut &o) {
    Device &d = *s.dev;
    CgRequest rq{};
    rq.m = &s.mats[slot];
    rq.b = b;
    rq.x = x;
    rq.ncol = ncol;
    rq.max_iter = s.opt.max_iterations;
    rq.rtol = s.opt.rtol;
    rq.atol = 0.0;
    rq.precision = s.opt.precision;
    rq.jacobi = (slot == NSFEM_VELOCITY_MASS && (s.opt.velocity_jacobi & 1)) ||
                (slot == NSFEM_DIFFUSION && (s.opt.velocity_jacobi & 2));
    set_update_policy(rq, slot);
    nsfem_cg_stats st[2]{};
    const int status = run_cg(d, rq, st);
    if (status != NSFEM_OK) {
        static const char *names[] = {"pressure stiffness", "velocity mass", "diffusion", "", ""};
        fail(status, std::string("CG on the ") + names[slot] + " matrix did not converge (status " +
                         std::to_string(status) + ")");
    }
    for (int c = 0; c < ncol; ++c) {
        o.iters[c] = st[c].iterations;
        o.refines[c] = st[c].refinements;
        o.rel[c] = st[c].rhs_norm > 0 ? st[c].final_residual / st[c].rhs_norm : st[c].final_residual;
    }
    o.seconds = st[0].seconds;
}

// out[i] = X_i . y for the first l basis vectors (one host synchronisation)
void proj_dots(Device &d, ProjBasis &B, int l, const double *y, double *out) {
    launch_fill(d, l, 0.0, B.s.ptr);
    for (int i = 0; i < l; ++i)
        launch_dot_accumulate(d, B.n, B.x.ptr + (size_t)i * B.n, y, B.s.ptr + i);
    NSFEM_HIP(hipMemcpyAsync(out, B.s.ptr, sizeof(double) * l, hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
}

double dev_dot(Device &d, ProjBasis &B, const double *a, const double *b) {
    double h = 0.0;
    launch_fill(d, 1, 0.0, B.s.ptr);
    launch_dot_accumulate(d, B.n, a, b, B.s.ptr);
    NSFEM_HIP(hipMemcpyAsync(&h, B.s.ptr, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    return h;
}

// Projection onto the span of past solutions (P. F. Fischer, "Projection
// techniques for iterative solution of Ax = b with successive right-hand
// sides", Comput. Methods Appl. Mech. Engrg. 163 (1998) 193-204).  With an
// A-orthonormal basis X of the past corrections, x = sum_i (X_i . b) X_i is
// the best approximation of A^-1 b in the A-norm within their span, which
// holds the previous solution: x is no worse than it in the A-norm.  An
// empty basis leaves x as it is.  The guess is kept for the basis update.
void projection_guess(Device &d, ProjBasis &B, const double *b, double *x) {
    const int l = B.dim;
    if (l > 0) {
        std::vector<double> a(l);
        proj_dots(d, B, l, b, a.data());
        launch_combine(d, B.n, l, a.data(), B.x.ptr, (size_t)B.n, false, x, B.guess.ptr);
    } else {
        launch_copy(d, B.n, x, B.guess.ptr);
    }
}

// After the solve x: the correction dx = x - guess (the whole solution after
// a restart), A-orthogonalised against X by classical Gram-Schmidt, a
// second pass only when the first cancelled more than half of its A-norm
// squared ("twice is enough": after a pass that keeps more, the loss of
// A-orthogonality is of the order of the rounding), is normalised in the
// A-norm and appended; a full basis restarts from the latest solution.  A
// correction that the basis already spans (A-norm below 1e-12 of its own
// before orthogonalisation) is dropped.
void projection_update(Device &d, ProjBasis &B, Matrix &A, const double *x) {
    const int n = B.n;
    if (B.dim == B.max) B.dim = 0;
    const int l = B.dim;
    double *dx = B.dx.ptr, *w = B.w.ptr;
    if (l == 0)
        launch_copy(d, n, x, dx);
    else
        launch_axpby(d, n, 1.0, -1.0, x, B.guess.ptr, dx);
    launch_spmv(d, A, dx, nullptr, w, 0);
    const double n0 = dev_dot(d, B, dx, w);
    if (!(n0 > 0.0)) return;
    std::vector<double> c(std::max(l, 1));
    double n1 = n0, before = n0;
    for (int pass = 0; pass < 2 && l > 0; ++pass) {
        proj_dots(d, B, l, w, c.data());  // X_i . A dx = (A X_i) . dx
        for (int i = 0; i < l; ++i) c[i] = -c[i];
        launch_combine(d, n, l, c.data(), B.x.ptr, (size_t)n, true, dx);
        launch_combine(d, n, l, c.data(), B.ax.ptr, (size_t)n, true, w);
        n1 = dev_dot(d, B, dx, w);
        if (n1 > 0.5 * before) break;
        before = n1;
    }
    if (!(n1 > 1e-12 * n0)) return;
    const double inv = 1.0 / std::sqrt(n1);
    launch_axpby(d, n, inv, 0.0, dx, dx, B.x.ptr + (size_t)l * n);
    launch_axpby(d, n, inv, 0.0, w, w, B.ax.ptr + (size_t)l * n);
    B.dim = l + 1;
}

// Initial guess of the pressure solve.  The reference keeps the previous
// step's pressure (include/assembly.h:608-615); the learned modes replace it
// with the network's prediction from the right-hand side (or keep whichever
// of the two has the smaller residual), or add the network's correction,
// predicted from the previous pressure's residual, to the previous pressure.
// ------------------------------------------------------------ coarse space
namespace {

// CSR transpose of an n x nc matrix (rows of the result list the original
// rows in ascending order).
void csr_transpose(int n, int nc, const int32_t *rs, const int32_t *col, const double *val,
                   std::vector<int32_t> &trs, std::vector<int32_t> &tcol,
                   std::vector<double> &tval) {
    trs.assign((size_t)nc + 1, 0);
    for (int64_t k = 0; k < rs[n]; ++k) ++trs[(size_t)col[k] + 1];
    for (int c = 0; c < nc; ++c) trs[c + 1] += trs[c];
    tcol.resize((size_t)rs[n]);
    tval.resize((size_t)rs[n]);
    std::vector<int32_t> fill(trs.begin(), trs.end() - 1);
    for (int r = 0; r < n; ++r)
        for (int32_t k = rs[r]; k < rs[r + 1]; ++k) {
            const int32_t at = fill[col[k]]++;
            tcol[at] = r;
            tval[at] = val[k];
        }
}

// A_c = P^T A P row by row: coarse row c sums P^T[c, f] A[f, g] P[g, c'] over
// the rows f of P^T's row c, the entries g of A's row f and the entries c' of
// P's row g, in that (CSR) order, so the result does not depend on the
// thread count.  Rows are split over `threads` workers.
void galerkin_product(int n, const std::vector<int32_t> &ars, const std::vector<int32_t> &acol,
                      const std::vector<double> &aval, int nc, const int32_t *prs,
                      const int32_t *pcol, const double *pval, const std::vector<int32_t> &trs,
                      const std::vector<int32_t> &tcol, const std::vector<double> &tval,
                      int threads, std::vector<int32_t> &crs, std::vector<int32_t> &ccol,
                      std::vector<double> &cval) {
    (void)n;
    const int T = std::max(1, std::min(threads, nc));
    std::vector<std::vector<int32_t>> pc(T), pr(T);
    std::vector<std::vector<double>> pv(T);
    auto work = [&](int t) {
        const int c0 = (int)((int64_t)nc * t / T), c1 = (int)((int64_t)nc * (t + 1) / T);
        std::vector<double> acc((size_t)nc, 0.0);
        std::vector<char> seen((size_t)nc, 0);
        std::vector<int32_t> touched;
        std::vector<int32_t> &rcnt = pr[t];
        rcnt.reserve((size_t)(c1 - c0));
        for (int c = c0; c < c1; ++c) {
            touched.clear();
            for (int32_t a = trs[c]; a < trs[c + 1]; ++a) {
                const int f = tcol[a];
                const double wf = tval[a];
                for (int32_t b = ars[f]; b < ars[f + 1]; ++b) {
                    const int g = acol[b];
                    const double wg = wf * aval[b];
                    for (int32_t e = prs[g]; e < prs[g + 1]; ++e) {
                        const int cc = pcol[e];
                        if (!seen[cc]) {
                            seen[cc] = 1;
                            touched.push_back(cc);
                        }
                        acc[cc] += wg * pval[e];
                    }
                }
            }
            std::sort(touched.begin(), touched.end());
            for (int cc : touched) {
                pc[t].push_back(cc);
                pv[t].push_back(acc[cc]);
                acc[cc] = 0.0;
                seen[cc] = 0;
            }
            rcnt.push_back((int32_t)touched.size());
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    crs.assign((size_t)nc + 1, 0);
    int c = 0;
    for (int t = 0; t < T; ++t)
        for (int32_t cnt : pr[t]) {
            crs[c + 1] = crs[c] + cnt;
            ++c;
        }
    NSFEM_REQUIRE((int64_t)crs[nc] < (int64_t)INT32_MAX, NSFEM_INVALID_ARGUMENT,
                  "coarse space: the coarse matrix is too large");
    ccol.clear();
    cval.clear();
    ccol.reserve((size_t)crs[nc]);
    cval.reserve((size_t)crs[nc]);
    for (int t = 0; t < T; ++t) {
        ccol.insert(ccol.end(), pc[t].begin(), pc[t].end());
        cval.insert(cval.end(), pv[t].begin(), pv[t].end());
    }
}

void drop_graph(hipGraphExec_t &g);

void release_pressure_multigrid(Simulation &s) {
    s.pmg = false;
    drop_graph(s.pcycle_graph);
    for (DeviceArray<float> *a : {&s.pdinv32, &s.pb32, &s.pz32, &s.pt32, &s.pp32}) a->release();
    s.pdinv.release();
    s.pr.release();
    s.pz.release();
    s.pp.release();
    s.pq.release();
    s.pt.release();
    s.pdot.release();
}

// 1/diag of a square host CSR, uploaded (a zero diagonal is rejected)
void upload_inverse_diagonal(Device &d, int n, const std::vector<int32_t> &rs,
                             const std::vector<int32_t> &col, const std::vector<double> &val,
                             DeviceArray<double> &out) {
    std::vector<double> inv((size_t)n, 0.0);
    for (int r = 0; r < n; ++r) {
        for (int32_t k = rs[r]; k < rs[r + 1]; ++k)
            if (col[k] == r) inv[r] = val[k];
        NSFEM_REQUIRE(inv[r] > 0.0, NSFEM_INVALID_ARGUMENT,
                      "coarse space: a coarse matrix has a non-positive diagonal entry");
        inv[r] = 1.0 / inv[r];
    }
    out.resize((size_t)n);
    NSFEM_HIP(hipMemcpy(out.ptr, inv.data(), inv.size() * sizeof(double), hipMemcpyHostToDevice));
}

// Multigrid solve of the coarse system: V-cycles over the levels added by
// simulation_add_coarse_level, each level smoothed by kMgSweeps damped
// Jacobi sweeps (weight kMgOmega) before and after the correction from the
// level below, the deepest level solved by FP64 CG to 1e-10; cycles repeat
// until ||b - A_c e|| <= coarse_rtol ||b|| (checked after every cycle, at
// most kMgMaxCycles).  On the 512^2 coarse grid of the 2048^2 cavity a cycle
// reduces the residual about 5x, so 9 cycles replace ~2,400 CG iterations.
constexpr int kMgSweeps = 2;
constexpr double kMgOmega = 0.8;
constexpr int kMgMaxCycles = 30;

// One V-cycle on A x = b over a level (A of n rows, its 1/diag, a residual
// vector r) and the levels below it, m.lv[l] the next coarser, from x = 0
// (zero: the first sweep needs no product) or from the x given: `sweeps`
// damped Jacobi sweeps, the correction from the level below on the
// restricted residual, as many sweeps after.  Each sweep with a product is
// one fused launch (launch_sweep) from one of x and r into the other, so the
// iterate alternates between them; the sweep from zero writes r, so that the
// 2 sweeps - 1 moves after it end in x (from a given x: 2 sweeps moves).
// The deepest level
// (l = m.lv.size()) is solved by its dense inverse when it has one, else by
// FP64 CG to 1e-10 from zero.  T = float runs the whole cycle in FP32 (the
// FP32 copies of every level's values and vectors, ensure_fp32_levels; it
// needs the dense deepest level).
template <typename T>
void mg_cycle(Device &d, Matrix &A, int n, const T *dinv, T *r, MgLevels &m, size_t l,
              const T *b, T *x, bool zero = true, int sweeps = kMgSweeps,
              double omega = kMgOmega) {
    if (l == m.lv.size() && m.deep_ok) {
        launch_spmv(d, m.deep_inv, b, nullptr, x, 0);
        return;
    }
    if (l == m.lv.size() || (zero && sweeps == 0)) launch_fill(d, n, 0.0, x);
    if (l == m.lv.size()) {
        if constexpr (std::is_same<T, double>::value) {
            CgRequest rq{};
            rq.m = &A;
            rq.b = b;
            rq.x = x;
            rq.ncol = 1;
            rq.max_iter = -1;
            rq.rtol = 1e-10;
            rq.precision = NSFEM_PRECISION_FP64;
            nsfem_cg_stats st{};
            run_cg(d, rq, &st);
        } else {
            fail(NSFEM_NOT_READY, "an FP32 V-cycle needs a dense deepest level");
        }
        return;
    }
    // the iterate in cur, the other vector free; a fused sweep moves it across
    T *cur = x, *spare = r;
    auto sweep = [&] {
        launch_sweep(d, A, omega, dinv, cur, b, spare);
        std::swap(cur, spare);
    };
    if (zero && sweeps > 0) {
        std::swap(cur, spare);
        launch_jacobi0(d, n, omega, dinv, b, cur);
        for (int k = 1; k < sweeps; ++k) sweep();
    } else {
        for (int k = 0; k < sweeps; ++k) sweep();
    }
    launch_spmv(d, A, cur, b, spare, 1);
    MgLevel &below = *m.lv[l];
    const LevelVectors<T> v(below);
    launch_spmv(d, below.PT, spare, nullptr, v.b, 0);
    mg_cycle<T>(d, below.A, below.n, v.dinv, v.r, m, l + 1, v.b, v.x);
    launch_spmv(d, below.P, v.x, nullptr, spare, 0);
    launch_add(d, n, spare, cur);
    for (int k = 0; k < sweeps; ++k) sweep();
    if (cur != x) fail(NSFEM_ERROR, "V-cycle: the sweeps did not end in the iterate's vector");
}

// The FP32 copies of m's levels (values of A, P, P^T and the dense deepest
// inverse; 1/diag; the cycle's vectors), built once.
void ensure_fp32_levels(Device &d, MgLevels &m) {
    if (m.fp32_ok) return;
    for (auto &lp : m.lv) {
        MgLevel &l = *lp;
        for (Matrix *a : {&l.A, &l.P, &l.PT}) ensure_val32(d, *a);
        l.dinv32.resize((size_t)l.n);
        launch_convert(d, l.n, l.dinv.ptr, l.dinv32.ptr);
        for (DeviceArray<float> *a : {&l.b32, &l.x32, &l.r32}) a->resize((size_t)l.n);
    }
    if (m.deep_ok) ensure_val32(d, m.deep_inv);
    m.fp32_ok = true;
}

// The pressure coarse space (present while has_coarse).
MgLevel &coarse(Simulation &s) { return *s.mg.lv[0]; }

// One V-cycle from the pressure coarse level down, from zero or from the x
// given.
void vcycle(Simulation &s, const double *b, double *x, bool zero) {
    MgLevel &C = coarse(s);
    mg_cycle<double>(*s.dev, C.A, s.nc, C.dinv.ptr, C.r.ptr, s.mg, 1, b, x, zero);
}

// V-cycles on A_c e = cr from the current e; returns the status, the cycle
// count and the wall time (the cycles synchronise with the host).
int mg_solve(Simulation &s, int64_t *cycles, double *seconds) {
    Device &d = *s.dev;
    MgLevel &C = coarse(s);
    const auto t0 = std::chrono::steady_clock::now();
    const double bb = dev_dot(d, s.cproj, C.b.ptr, C.b.ptr);
    const double tol2 = s.coarse_rtol * s.coarse_rtol * bb;
    int status = NSFEM_MAX_ITERATIONS_REACHED;
    int k = 0;
    if (!(bb > 0.0)) {
        status = NSFEM_OK;  // e = projection guess of a zero right-hand side
    } else {
        while (k < kMgMaxCycles) {
            vcycle(s, C.b.ptr, C.x.ptr, false);
            ++k;
            launch_spmv(d, C.A, C.x.ptr, C.b.ptr, C.r.ptr, 1);
            const double rr = dev_dot(d, s.cproj, C.r.ptr, C.r.ptr);
            if (!std::isfinite(rr)) {
                status = NSFEM_BREAKDOWN;
                break;
            }
            if (rr <= tol2) {
                status = NSFEM_OK;
                break;
            }
        }
    }
    *cycles = k;
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return status;
}

// The coarse correction in three parts, so that a batch can solve every
// instance's coarse system in one launch: prepare restricts r0 = b - A x0
// into cr and starts ce from the projection onto past coarse solutions;
// coarse_request describes the solve A_c e = P^T r0 (to coarse_rtol);
// finish adds P e to x0 and imposes the pressure Dirichlet values again,
// unless the coarse solve did not converge (then the guess stays as it was).
void coarse_prepare(Simulation &s) {
    Device &d = *s.dev;
    Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
    MgLevel &C = coarse(s);
    launch_spmv(d, A, s.p.ptr, s.prhs.ptr, s.cy.ptr, 1);
    launch_spmv(d, C.PT, s.cy.ptr, nullptr, C.b.ptr, 0);
    launch_fill(d, s.nc, 0.0, C.x.ptr);
    projection_guess(d, s.cproj, C.b.ptr, C.x.ptr);
}

CgRequest coarse_request(Simulation &s) {
    CgRequest rq{};
    MgLevel &C = coarse(s);
    rq.m = &C.A;
    rq.b = C.b.ptr;
    rq.x = C.x.ptr;
    rq.ncol = 1;
    rq.max_iter = -1;
    rq.rtol = s.coarse_rtol;
    rq.atol = 0.0;
    rq.precision = s.opt.precision;
    return rq;
}

void coarse_finish(Simulation &s, int status, int64_t iterations, double seconds) {
    Device &d = *s.dev;
    s.coarse_iters = iterations;
    s.coarse_seconds = seconds;
    if (status != NSFEM_OK) return;
    MgLevel &C = coarse(s);
    projection_update(d, s.cproj, C.A, C.x.ptr);
    launch_spmv(d, C.P, C.x.ptr, nullptr, s.cy.ptr, 0);
    launch_add(d, s.P, s.cy.ptr, s.p.ptr);
    launch_scatter(d, s.npb, s.pb_idx.ptr, s.pb_val.ptr, s.p.ptr);
}

void coarse_correct(Simulation &s) {
    coarse_prepare(s);
    if (s.mg.lv.size() > 1) {
        int64_t cycles = 0;
        double sec = 0.0;
        const int status = mg_solve(s, &cycles, &sec);
        coarse_finish(s, status, cycles, sec);
        return;
    }
    const CgRequest rq = coarse_request(s);
    nsfem_cg_stats st{};
    const int status = run_cg(*s.dev, rq, &st);
    coarse_finish(s, status, st.iterations, st.seconds);
}

void drop_graph(hipGraphExec_t &g) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
}

// Runs `launches` on the device's stream through a graph captured on the
// first call and replayed after it: the launches must be the same kernels
// with the same arguments every call, and allocate, copy, clear memory and
// synchronise nothing (an FP32 cycle over a fixed hierarchy and fixed
// vectors: products, sweeps, transfers and the dense deepest solve).  A
// graph replaces ~100 launches of the cycle, most of them on levels so
// small that their launch latency is their cost.  Opt-in with
// NSFEM_VCYCLE_GRAPH=1 (read at each call): at 2048^2 it gained 0.6%
// (19.67 against 19.55 steps/s), and one kernel-traced run of the bench
// (rocprofv3 --kernel-trace) ended in a host-side segmentation fault inside
// a replay, in the accuracy run that follows the timed ones; the same
// command without the profiler and the replay-against-direct test pass.
// Until that is understood the cycles launch directly by default.
template <class F>
void run_captured(Device &d, hipGraphExec_t &exec, F launches) {
    const char *e = std::getenv("NSFEM_VCYCLE_GRAPH");
    if (!(e && std::atoi(e) == 1)) {
        launches();
        return;
    }
    if (!exec) {
        NSFEM_HIP(hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal));
        hipGraph_t g = nullptr;
        try {
            launches();
        } catch (...) {
            if (hipStreamEndCapture(d.stream, &g) == hipSuccess && g) (void)hipGraphDestroy(g);
            throw;
        }
        NSFEM_HIP(hipStreamEndCapture(d.stream, &g));
        const hipError_t err = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) exec = nullptr;
        NSFEM_HIP(err);
    }
    NSFEM_HIP(hipGraphLaunch(exec, d.stream));
}

// z = M r: one V-cycle from zero over the fine pressure grid (damped Jacobi
// as on the coarse levels), the coarse space and its levels.  As many sweeps
// after as before, restriction = transposed prolongation and the deepest
// level solved exactly (or by CG to 1e-10) keep M symmetric, as
// preconditioned CG needs.  z is zeroed on the pressure Dirichlet rows
// (D M D with D the projection off them stays symmetric; the solve holds
// those rows at their values).
void fine_vcycle(Simulation &s, const double *r, double *z) {
    Device &d = *s.dev;
    mg_cycle<double>(d, s.mats[NSFEM_PRESSURE_STIFFNESS], s.P, s.pdinv.ptr, s.pt.ptr, s.mg, 0, r,
                     z);
    launch_scatter_value(d, s.npb, s.pb_idx.ptr, 0.0, z);
}

// In the refined precision, when the deepest level has its dense inverse,
// the cycle runs in FP32 (pcg_run32 hands it r rounded to FP32 and uses z
// widened; the FP64 CG keeps its FP64 stopping rule).
bool pressure_cycle_fp32(const Simulation &s) {
    return s.opt.precision == NSFEM_PRECISION_FP32_REFINED && s.mg.deep_ok;
}

// The FP32 copies of the pressure hierarchy and the fine level's FP32
// vectors (r32 in, z32 out, the sweeps' second vector), built once.
void ensure_pressure_fp32(Simulation &s) {
    if (s.mg.fp32_ok && s.pdinv32.ptr) return;
    Device &d = *s.dev;
    const int n = s.P;
    ensure_fp32_levels(d, s.mg);
    ensure_val32(d, s.mats[NSFEM_PRESSURE_STIFFNESS]);
    s.pdinv32.resize((size_t)n);
    launch_convert(d, n, s.pdinv.ptr, s.pdinv32.ptr);
    for (DeviceArray<float> *a : {&s.pb32, &s.pz32, &s.pt32, &s.pp32}) a->resize((size_t)n);
}

// z = M r in FP32 on the fixed vectors r = pb32, z = pz32 (captured)
void fine_vcycle32(Simulation &s) {
    Device &d = *s.dev;
    run_captured(d, s.pcycle_graph, [&] {
        mg_cycle<float>(d, s.mats[NSFEM_PRESSURE_STIFFNESS], s.P, s.pdinv32.ptr, s.pt32.ptr, s.mg,
                        0, s.pb32.ptr, s.pz32.ptr);
        launch_scatter_value(d, s.npb, s.pb_idx.ptr, 0.0, s.pz32.ptr);
    });
}

// a . b over n entries, through a one-entry device scratch (one host
// synchronisation)
double vec_dot(Device &d, int n, const double *a, const double *b, double *scratch) {
    double h = 0.0;
    launch_fill(d, 1, 0.0, scratch);
    launch_dot_accumulate(d, n, a, b, scratch);
    NSFEM_HIP(hipMemcpyAsync(&h, scratch, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    return h;
}

// Vectors of a preconditioned CG over n rows: residual, preconditioned
// residual, direction, product, and a one-entry dot scratch.
struct PcgWork {
    double *r, *z, *p, *q, *dot;
};

// CG preconditioned by z = precond(r) (symmetric positive definite), in
// FP64, on A x = b from the x given, to ||b - A x|| <= rtol ||b||;
// convergence is decided on a recomputed b - A x, which then replaces the
// recursive residual.  Fills o's column `c` and seconds; returns the status.
template <class Pre>
int pcg_run(Device &d, Matrix &A, int n, const double *b, double *x, const PcgWork &w,
            double rtol, int64_t max_iterations, Pre precond, SolveOut &o, int c) {
    const auto t0 = std::chrono::steady_clock::now();
    double *r = w.r, *z = w.z, *p = w.p, *q = w.q;
    const double bb = vec_dot(d, n, b, b, w.dot);
    launch_spmv(d, A, x, b, r, 1);
    double rr = vec_dot(d, n, r, r, w.dot);
    const double ref = bb > 0.0 ? bb : rr;
    const double tol2 = rtol * rtol * ref;
    const int64_t max_iter = max_iterations > 0 ? max_iterations : n;
    int64_t k = 0;
    int status = NSFEM_OK;
    if (rr > tol2) {
        status = NSFEM_MAX_ITERATIONS_REACHED;
        precond(r, z);
        launch_copy(d, n, z, p);
        double rz = vec_dot(d, n, r, z, w.dot);
        while (k < max_iter) {
            launch_spmv(d, A, p, nullptr, q, 0);
            const double pq = vec_dot(d, n, p, q, w.dot);
            if (!(pq > 0.0) || !(rz > 0.0)) {
                status = NSFEM_BREAKDOWN;
                break;
            }
            const double alpha = rz / pq;
            launch_axpy(d, n, alpha, p, x);
            launch_axpy(d, n, -alpha, q, r);
            ++k;
            rr = vec_dot(d, n, r, r, w.dot);
            if (rr <= tol2) {
                launch_spmv(d, A, x, b, r, 1);
                rr = vec_dot(d, n, r, r, w.dot);
                if (rr <= tol2) {
                    status = NSFEM_OK;
                    break;
                }
            }
            precond(r, z);
            const double rz_new = vec_dot(d, n, r, z, w.dot);
            const double beta = rz_new / rz;
            rz = rz_new;
            launch_axpby(d, n, 1.0, beta, z, p, p);
        }
    }
    o.iters[c] = k;
    o.refines[c] = 0;
    o.rel[c] = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    o.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return status;
}

// Reads a device scalar accumulated by the dot launchers (one host
// synchronisation) and clears it for the next.
double take_scalar(Device &d, double *dot) {
    double h = 0.0;
    NSFEM_HIP(hipMemcpyAsync(&h, dot, sizeof(double), hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    launch_fill(d, 1, 0.0, dot);
    return h;
}

// pcg_run with an FP32 preconditioner z32 = precond(r32) and FP32
// directions: p in FP32, q = A p on the FP32 values with every product and
// sum in FP64 (an FP32 sum would cancel away smooth directions' products on
// the pressure matrix), x and r updated in FP64 in one pass that also
// writes r rounded to FP32 and returns (r, r); every dot accumulates in
// FP64.  As in the mixed persistent CG, the recursion on the FP32 values
// only steers: convergence is decided on a recomputed FP64 b - A x, which
// then replaces the recursive residual (a reliable update), as it does
// whenever the recursive residual has fallen kPcgUpdateDrop-fold since the
// last one (the FP32 directions keep conjugacy to ~1e-7, so a longer
// recursion would stall).  The stopping rule is pcg_run's.  Returns
// NSFEM_STAGNATION after kPcgRecomputeMax recomputed residuals above the
// tolerance where the recursion had met it, or after kPcgIterationsMax
// iterations (the V-cycle preconditioned solves take tens); x is then the
// last iterate, for the FP64 form to finish.
constexpr int kPcgRecomputeMax = 8;
constexpr double kPcgUpdateDrop = 1e3;
constexpr int64_t kPcgIterationsMax = 500;

template <class Pre>
int pcg_run32(Device &d, Matrix &A, int n, const double *b, double *x, const PcgWork &w,
              float *r32, float *z32, float *p, double rtol, int64_t max_iterations,
              Pre precond, SolveOut &o, int c) {
    const auto t0 = std::chrono::steady_clock::now();
    double *r = w.r, *q = w.q;
    const double bb = vec_dot(d, n, b, b, w.dot);
    launch_spmv(d, A, x, b, r, 1);
    double rr = vec_dot(d, n, r, r, w.dot);
    const double ref = bb > 0.0 ? bb : rr;
    const double tol2 = rtol * rtol * ref;
    const int64_t max_iter = max_iterations > 0 ? max_iterations : n;
    int64_t k = 0;
    int status = NSFEM_OK;
    int recomputed = 0;
    double rr_true = rr;  // the last recomputed residual
    if (rr > tol2) {
        status = NSFEM_MAX_ITERATIONS_REACHED;
        launch_fill(d, 1, 0.0, w.dot);
        launch_convert(d, n, r, r32);
        precond(r32, z32);
        launch_copy(d, n, z32, p);
        launch_dot_accumulate(d, n, r, z32, w.dot);
        double rz = take_scalar(d, w.dot);
        while (k < max_iter) {
            launch_spmv_acc64(d, A, p, q);
            launch_dot_accumulate(d, n, q, p, w.dot);
            const double pq = take_scalar(d, w.dot);
            if (!(pq > 0.0) || !(rz > 0.0)) {
                status = NSFEM_BREAKDOWN;
                break;
            }
            const double alpha = rz / pq;
            launch_pcg_update(d, n, alpha, p, q, x, r, r32, w.dot);
            rr = take_scalar(d, w.dot);
            ++k;
            const bool met = rr <= tol2;
            if (met || rr * kPcgUpdateDrop * kPcgUpdateDrop <= rr_true) {
                launch_spmv(d, A, x, b, r, 1);
                launch_dot_accumulate(d, n, r, r, w.dot);
                rr = take_scalar(d, w.dot);
                rr_true = rr;
                if (rr <= tol2) {
                    status = NSFEM_OK;
                    break;
                }
                if (met && ++recomputed >= kPcgRecomputeMax) {
                    status = NSFEM_STAGNATION;
                    break;
                }
                launch_convert(d, n, r, r32);
            }
            if (k >= kPcgIterationsMax) {
                status = NSFEM_STAGNATION;
                break;
            }
            precond(r32, z32);
            launch_dot_accumulate(d, n, r, z32, w.dot);
            const double rz_new = take_scalar(d, w.dot);
            const double beta = rz_new / rz;
            rz = rz_new;
            launch_pcg_direction(d, n, z32, beta, p);
        }
    }
    o.iters[c] = k;
    o.refines[c] = 0;
    o.rel[c] = bb > 0.0 ? std::sqrt(rr / bb) : std::sqrt(rr);
    o.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return status;
}

// The pressure solve as CG preconditioned by fine_vcycle from the guess in
// p.  The Dirichlet rows (identity rows) start at their values, so their
// residual is zero and z, p stay zero there.
void pcg_pressure(Simulation &s, SolveOut &o) {
    Device &d = *s.dev;
    launch_scatter(d, s.npb, s.pb_idx.ptr, s.pb_val.ptr, s.p.ptr);
    const PcgWork w{s.pr.ptr, s.pz.ptr, s.pp.ptr, s.pq.ptr, s.pdot.ptr};
    Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
    int status = NSFEM_OK;
    SolveOut first{};
    if (pressure_cycle_fp32(s)) {
        ensure_pressure_fp32(s);
        status = pcg_run32(d, A, s.P, s.prhs.ptr, s.p.ptr, w, s.pb32.ptr, s.pz32.ptr, s.pp32.ptr,
                           s.opt.rtol, s.opt.max_iterations,
                           [&](const float *, float *) { fine_vcycle32(s); }, o, 0);
    }
    if (!pressure_cycle_fp32(s) || status == NSFEM_STAGNATION) {
        // FP64 throughout (from the last iterate when the FP32 form stalled)
        first = o;
        status = pcg_run(d, A, s.P, s.prhs.ptr, s.p.ptr, w, s.opt.rtol, s.opt.max_iterations,
                         [&](const double *r, double *z) { fine_vcycle(s, r, z); }, o, 0);
        o.iters[0] += first.iters[0];
        o.seconds += first.seconds;
    }
    if (status != NSFEM_OK)
        fail(status, "V-cycle preconditioned CG on the pressure stiffness matrix did not "
                     "converge (status " + std::to_string(status) + ")");
}

// z = M r for the diffusion solves: one V-cycle from zero over the velocity
// nodes (damped Jacobi), the velocity coarse space and its levels, zeroed
// on the velocity Dirichlet rows (whose residual is zero: identity rows,
// their columns eliminated, the solve starting at the boundary values; the
// prolongation's rows there are empty).  Symmetric as fine_vcycle.  In the
// refined precision the cycle runs in FP32 (velocity_vcycle32 under
// pcg_run32): the products and sweeps move about half the bytes, and the
// FP64 CG around it still decides convergence on its FP64 residual.
// Jacobi weight of the sweeps on the P2 velocity nodes (the levels below
// keep kMgOmega): D^-1 A of the diffusion matrix reaches ~2.2 (at weight 0.9
// the 2048^2 cavity's v column needs 17 iterations instead of 6); measured
// there per step, weight 0.6 / 0.65 / 0.7 / 0.8 took 22.9 / 21.4 / 20.8 /
// 23.5 ms (6.3 / 5.9 / 5.7 / 6.5 iterations); three sweeps instead of two
// save an iteration but cost more (23.5 ms at 0.7).
constexpr double kVmgFineOmega = 0.7;

bool velocity_cycle_fp32(const Simulation &s) {
    return s.opt.precision == NSFEM_PRECISION_FP32_REFINED && s.vmg.deep_ok;
}

void velocity_vcycle(Simulation &s, const double *r, double *z) {
    Device &d = *s.dev;
    mg_cycle<double>(d, s.mats[NSFEM_DIFFUSION], s.V, s.vdinv.ptr, s.vt.ptr, s.vmg, 0, r, z, true,
                     kMgSweeps, kVmgFineOmega);
    launch_scatter_value(d, s.nvb, s.vb_idx.ptr, 0.0, z);
}

// The FP32 copies of the velocity hierarchy and the P2 level's FP32 vectors
// (r32 in, z32 out, the sweeps' second vector), built once.
void ensure_velocity_fp32(Simulation &s) {
    if (s.vmg.fp32_ok && s.vdinv32.ptr) return;
    Device &d = *s.dev;
    const int V = s.V;
    ensure_fp32_levels(d, s.vmg);
    ensure_val32(d, s.mats[NSFEM_DIFFUSION]);
    s.vdinv32.resize((size_t)V);
    launch_convert(d, V, s.vdinv.ptr, s.vdinv32.ptr);
    for (DeviceArray<float> *a : {&s.vb32, &s.vz32, &s.vt32, &s.vp32}) a->resize((size_t)V);
}

// z = M r in FP32 on the fixed vectors r = vb32, z = vz32 (captured)
void velocity_vcycle32(Simulation &s) {
    Device &d = *s.dev;
    run_captured(d, s.vcycle_graph, [&] {
        mg_cycle<float>(d, s.mats[NSFEM_DIFFUSION], s.V, s.vdinv32.ptr, s.vt32.ptr, s.vmg, 0,
                        s.vb32.ptr, s.vz32.ptr, true, kMgSweeps, kVmgFineOmega);
        launch_scatter_value(d, s.nvb, s.vb_idx.ptr, 0.0, s.vz32.ptr);
    });
}

// Column c (0: u, 1: v) of the diffusion solve by CG preconditioned with
// velocity_vcycle, from the guess in the column of uv.
void pcg_velocity(Simulation &s, int c, SolveOut &o) {
    Device &d = *s.dev;
    const size_t off = (size_t)c * s.V;
    double *x = s.uv.ptr + off;
    launch_scatter(d, s.nvb, s.vb_idx.ptr, c ? s.vb_v.ptr : s.vb_u.ptr, x);
    const PcgWork w{s.vr.ptr, s.vz.ptr, s.vp.ptr, s.vq.ptr, s.vdot.ptr};
    Matrix &A = s.mats[NSFEM_DIFFUSION];
    int status = NSFEM_OK;
    SolveOut first{};
    if (velocity_cycle_fp32(s)) {
        ensure_velocity_fp32(s);
        status = pcg_run32(d, A, s.V, s.rhs.ptr + off, x, w, s.vb32.ptr, s.vz32.ptr, s.vp32.ptr,
                           s.opt.rtol, s.opt.max_iterations,
                           [&](const float *, float *) { velocity_vcycle32(s); }, o, 0);
    }
    if (!velocity_cycle_fp32(s) || status == NSFEM_STAGNATION) {
        // FP64 throughout (from the last iterate when the FP32 form stalled)
        first = o;
        status = pcg_run(d, A, s.V, s.rhs.ptr + off, x, w, s.opt.rtol, s.opt.max_iterations,
                         [&](const double *r, double *z) { velocity_vcycle(s, r, z); }, o, 0);
        o.iters[0] += first.iters[0];
        o.seconds += first.seconds;
    }
    if (status != NSFEM_OK)
        fail(status, "V-cycle preconditioned CG on the diffusion matrix did not converge "
                     "(status " + std::to_string(status) + ")");
}

}  // namespace

void simulation_set_velocity_projection(Simulation &s, int depth) {
    NSFEM_REQUIRE(depth >= 0 && depth <= 256, NSFEM_INVALID_ARGUMENT,
                  "velocity projection depth must lie in [0, 256]");
    s.dev->set_current();
    if (depth == 0) {
        s.vproj[0].release();
        s.vproj[1].release();
        s.vdel.release();
        return;
    }
    s.vproj[0].reserve(s.V, depth);
    s.vproj[1].reserve(s.V, depth);
    s.vdel.resize(2 * (size_t)s.V);
}

void simulation_set_mass_projection(Simulation &s, int depth) {
    NSFEM_REQUIRE(depth >= 0 && depth <= 256, NSFEM_INVALID_ARGUMENT,
                  "mass projection depth must lie in [0, 256]");
    s.dev->set_current();
    for (ProjBasis &B : s.mproj) {
        if (depth == 0)
            B.release();
        else
            B.reserve(s.V, depth);
    }
}

// Levels of at most this many nodes are solved by a dense inverse (formed
// on the host by Cholesky, about a second at 33^2 nodes) applied as a dense
// CSR product; larger ones by CG.  A dense 33^2 level replaces two halvings
// whose two dozen launches per cycle are bound by launch latency.
constexpr int kMgDenseMax = 1100;

namespace {
// The dense inverse of m's deepest level when it has at most kMgDenseMax
// nodes and is positive definite (else the deepest level keeps CG).
void dense_inverse(Device &d, MgLevels &m) {
    matrix_release(m.deep_inv);
    m.deep_ok = false;
    const int n = m.lv.back()->n;
    if (n > kMgDenseMax) return;
    std::vector<double> inv;
    if (!spd_inverse(n, m.deep_rs.data(), m.deep_col.data(), m.deep_val.data(), inv)) return;
    std::vector<int32_t> rs((size_t)n + 1), col((size_t)n * n);
    for (int r = 0; r <= n; ++r) rs[r] = r * n;
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < n; ++c) col[(size_t)r * n + c] = c;
    upload_csr(d, m.deep_inv, n, n, (int64_t)n * n, rs.data(), col.data(), inv.data());
    m.deep_ok = true;
}

// Appends a level below m's deepest: P (`above` rows, `count` columns; the
// upload validates it) maps the new level's nodes to the level above, whose
// matrix is the host CSR (ars, acol, aval); the new matrix P^T A_above P
// (formed on `threads` host threads) becomes m's deepest.
void append_level(Device &d, MgLevels &m, int above, const std::vector<int32_t> &ars,
                  const std::vector<int32_t> &acol, const std::vector<double> &aval, int count,
                  int64_t nnz, const int32_t *rs, const int32_t *col, const double *val,
                  int threads) {
    std::unique_ptr<MgLevel> lv(new MgLevel());
    lv->n = count;
    upload_csr(d, lv->P, above, count, nnz, rs, col, val);
    std::vector<int32_t> trs, tcol, crs, ccol;
    std::vector<double> tval, cval;
    csr_transpose(above, count, rs, col, val, trs, tcol, tval);
    galerkin_product(above, ars, acol, aval, count, rs, col, val, trs, tcol, tval, threads, crs,
                     ccol, cval);
    upload_csr(d, lv->PT, count, above, nnz, trs.data(), tcol.data(), tval.data());
    upload_csr(d, lv->A, count, count, (int64_t)ccol.size(), crs.data(), ccol.data(), cval.data());
    upload_inverse_diagonal(d, count, crs, ccol, cval, lv->dinv);
    lv->b.resize(count);
    lv->x.resize(count);
    lv->r.resize(count);
    m.deep_rs.swap(crs);
    m.deep_col.swap(ccol);
    m.deep_val.swap(cval);
    m.lv.push_back(std::move(lv));
    dense_inverse(d, m);
    m.fp32_ok = false;
}

int host_threads(const Simulation &s) {
    return s.opt.threads > 0 ? s.opt.threads
                             : (int)std::max(1u, std::thread::hardware_concurrency());
}
// The host CSR (row starts, columns, values) of a simulation matrix.
void host_matrix(Matrix &A, std::vector<int32_t> &col, std::vector<double> &val) {
    col.resize((size_t)A.nnz);
    val.resize((size_t)A.nnz);
    if (!A.nnz) return;
    NSFEM_HIP(hipMemcpy(col.data(), A.col.ptr, col.size() * sizeof(int32_t),
                        hipMemcpyDeviceToHost));
    NSFEM_HIP(hipMemcpy(val.data(), A.val.ptr, val.size() * sizeof(double),
                        hipMemcpyDeviceToHost));
}

}  // namespace

namespace {
// Drops the coarse space, its levels and the V-cycle preconditioner.
void release_coarse_space(Simulation &s) {
    s.cy.release();
    s.cproj.release();
    s.mg.release();
    release_pressure_multigrid(s);
    s.has_coarse = false;
    s.nc = 0;
}

void build_coarse_space(Simulation &s, int nc, int64_t nnz, const int32_t *rs,
                        const int32_t *col, const double *val, double rtol);
}  // namespace

// Sets (nc > 0) or removes (nc <= 0) the coarse space.  A failure part way
// leaves no coarse space at all (and no V-cycle preconditioner), never a
// half-built one.
void simulation_set_coarse_space(Simulation &s, int nc, int64_t nnz, const int32_t *rs,
                                 const int32_t *col, const double *val, double rtol) {
    Device &d = *s.dev;
    d.set_current();
    if (nc > 0) {
        NSFEM_REQUIRE(nc <= s.P && nnz > 0 && rs && col && val, NSFEM_INVALID_ARGUMENT,
                      "coarse space: need 0 < coarse count <= pressure count and a non-empty P");
        NSFEM_REQUIRE(rtol > 0.0 && rtol < 1.0, NSFEM_INVALID_ARGUMENT,
                      "coarse space: the coarse tolerance must lie in (0, 1)");
    }
    release_coarse_space(s);
    if (nc <= 0) return;
    try {
        build_coarse_space(s, nc, nnz, rs, col, val, rtol);
    } catch (...) {
        release_coarse_space(s);
        throw;
    }
}

namespace {
void build_coarse_space(Simulation &s, int nc, int64_t nnz, const int32_t *rs,
                        const int32_t *col, const double *val, double rtol) {
    Device &d = *s.dev;
    Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
    std::vector<int32_t> acol;
    std::vector<double> aval;
    host_matrix(A, acol, aval);
    append_level(d, s.mg, s.P, A.host_rs, acol, aval, nc, nnz, rs, col, val, host_threads(s));
    s.cy.resize(s.P);
    s.cproj.reserve(nc, 16);
    s.nc = nc;
    s.coarse_rtol = rtol;
    s.coarse_iters = 0;
    s.coarse_seconds = 0.0;
    s.has_coarse = true;
}
}  // namespace

// Appends a multigrid level below the deepest one: P maps the new level's
// count nodes to the deepest level's (coarse grid first), A_new = P^T A_deep
// P.  With levels the coarse system is solved by V-cycles (mg_solve).
void simulation_add_coarse_level(Simulation &s, int count, int64_t nnz, const int32_t *rs,
                                 const int32_t *col, const double *val) {
    Device &d = *s.dev;
    d.set_current();
    NSFEM_REQUIRE(s.has_coarse, NSFEM_INVALID_ARGUMENT,
                  "coarse level: set a coarse space first");
    const int above = s.mg.lv.back()->n;
    NSFEM_REQUIRE(count > 0 && count < above && nnz > 0 && rs && col && val,
                  NSFEM_INVALID_ARGUMENT,
                  "coarse level: need 0 < count < the level above's size and a non-empty P");
    drop_graph(s.pcycle_graph);
    append_level(d, s.mg, above, s.mg.deep_rs, s.mg.deep_col, s.mg.deep_val, count, nnz, rs, col,
                 val, host_threads(s));
}

// Switches the pressure solve to CG preconditioned by one V-cycle over the
// fine grid, the coarse space and its levels (FP64; the coarse correction of
// the guess is then left out: the preconditioner holds it).
void simulation_set_pressure_multigrid(Simulation &s, int enable) {
    Device &d = *s.dev;
    d.set_current();
    if (!enable) {
        release_pressure_multigrid(s);
        return;
    }
    NSFEM_REQUIRE(s.has_coarse, NSFEM_INVALID_ARGUMENT,
                  "pressure multigrid: set a coarse space (and its levels) first");
    Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
    std::vector<int32_t> acol((size_t)A.nnz);
    std::vector<double> aval((size_t)A.nnz);
    if (A.nnz) {
        NSFEM_HIP(hipMemcpy(acol.data(), A.col.ptr, acol.size() * sizeof(int32_t),
                            hipMemcpyDeviceToHost));
        NSFEM_HIP(hipMemcpy(aval.data(), A.val.ptr, aval.size() * sizeof(double),
                            hipMemcpyDeviceToHost));
    }
    upload_inverse_diagonal(d, s.P, A.host_rs, acol, aval, s.pdinv);
    s.pr.resize(s.P);
    s.pz.resize(s.P);
    s.pp.resize(s.P);
    s.pq.resize(s.P);
    s.pt.resize(s.P);
    s.pdot.resize(1);
    s.pmg = true;
}

namespace {
// Drops the velocity hierarchy and the V-cycle preconditioned diffusion
// solves.
void release_velocity_multigrid(Simulation &s) {
    drop_graph(s.vcycle_graph);
    s.vmg.release();
    s.vmg_mode = 0;
    for (DeviceArray<double> *a : {&s.vdinv, &s.vr, &s.vz, &s.vp, &s.vq, &s.vt, &s.vdot})
        a->release();
    for (DeviceArray<float> *a : {&s.vdinv32, &s.vb32, &s.vz32, &s.vt32, &s.vp32}) a->release();
    for (int c = 0; c < 2; ++c)
        for (int w = 0; w < 2; ++w) {
            s.vmg_sec[c][w] = 0.0;
            s.vmg_age[c][w] = 0;
        }
}

void build_velocity_coarse_space(Simulation &s, int nc, int64_t nnz, const int32_t *rs,
                                 const int32_t *col, const double *val) {
    Device &d = *s.dev;
    const int V = s.V;
    NSFEM_REQUIRE(rs[0] == 0 && rs[V] == nnz, NSFEM_INVALID_ARGUMENT,
                  "velocity coarse space: row starts must run from 0 to nnz");
    for (int r = 0; r < V; ++r)
        NSFEM_REQUIRE(rs[r] <= rs[r + 1], NSFEM_INVALID_ARGUMENT,
                      "velocity coarse space: row starts must not decrease");
    // rows into the internal (lattice) order of the velocity nodes
    std::vector<int32_t> prs((size_t)V + 1, 0), pcol((size_t)nnz);
    std::vector<double> pval((size_t)nnz);
    for (int r = 0; r < V; ++r) {
        const int src = s.vorder.empty() ? r : s.vorder[r];
        const int32_t len = rs[src + 1] - rs[src];
        std::copy(col + rs[src], col + rs[src] + len, pcol.begin() + prs[r]);
        std::copy(val + rs[src], val + rs[src] + len, pval.begin() + prs[r]);
        prs[r + 1] = prs[r] + len;
    }
    std::vector<int32_t> vb((size_t)s.nvb);
    if (s.nvb)
        NSFEM_HIP(hipMemcpy(vb.data(), s.vb_idx.ptr, vb.size() * sizeof(int32_t),
                            hipMemcpyDeviceToHost));
    for (int32_t r : vb)
        NSFEM_REQUIRE(prs[r] == prs[r + 1], NSFEM_INVALID_ARGUMENT,
                      "velocity coarse space: the rows of velocity Dirichlet nodes must be empty");
    Matrix &A = s.mats[NSFEM_DIFFUSION];
    std::vector<int32_t> acol;
    std::vector<double> aval;
    host_matrix(A, acol, aval);
    upload_inverse_diagonal(d, V, A.host_rs, acol, aval, s.vdinv);
    append_level(d, s.vmg, V, A.host_rs, acol, aval, nc, nnz, prs.data(), pcol.data(),
                 pval.data(), host_threads(s));
    for (DeviceArray<double> *a : {&s.vr, &s.vz, &s.vp, &s.vq, &s.vt}) a->resize((size_t)V);
    s.vdot.resize(1);
    s.vmg_mode = 1;
}
}  // namespace

// Sets (nc > 0) or removes (nc <= 0) the multigrid hierarchy below the
// velocity space: P (V x nc, rows in the caller's node order) maps its
// first coarse space onto the velocity nodes, and the diffusion solves
// switch to V-cycle preconditioned CG (mode 1: per column by measured cost).
// A failure part way leaves no hierarchy.
void simulation_set_velocity_coarse_space(Simulation &s, int nc, int64_t nnz, const int32_t *rs,
                                          const int32_t *col, const double *val) {
    Device &d = *s.dev;
    d.set_current();
    if (nc > 0)
        NSFEM_REQUIRE(nc < s.V && nnz > 0 && rs && col && val, NSFEM_INVALID_ARGUMENT,
                      "velocity coarse space: need 0 < coarse count < velocity count and a "
                      "non-empty P");
    release_velocity_multigrid(s);
    if (nc <= 0) return;
    try {
        build_velocity_coarse_space(s, nc, nnz, rs, col, val);
    } catch (...) {
        release_velocity_multigrid(s);
        throw;
    }
}

// Appends a level below the deepest level of the velocity hierarchy.
void simulation_add_velocity_coarse_level(Simulation &s, int count, int64_t nnz,
                                          const int32_t *rs, const int32_t *col,
                                          const double *val) {
    Device &d = *s.dev;
    d.set_current();
    NSFEM_REQUIRE(!s.vmg.lv.empty(), NSFEM_INVALID_ARGUMENT,
                  "velocity coarse level: set a velocity coarse space first");
    const int above = s.vmg.lv.back()->n;
    NSFEM_REQUIRE(count > 0 && count < above && nnz > 0 && rs && col && val,
                  NSFEM_INVALID_ARGUMENT,
                  "velocity coarse level: need 0 < count < the level above's size and a "
                  "non-empty P");
    drop_graph(s.vcycle_graph);
    append_level(d, s.vmg, above, s.vmg.deep_rs, s.vmg.deep_col, s.vmg.deep_val, count, nnz, rs,
                 col, val, host_threads(s));
}

// Which diffusion columns run V-cycle preconditioned CG: 0 none, 1 each
// column by measured cost, 2 both.
void simulation_set_velocity_multigrid(Simulation &s, int mode) {
    NSFEM_REQUIRE(mode >= 0 && mode <= 2, NSFEM_INVALID_ARGUMENT,
                  "velocity multigrid: mode must be 0, 1 or 2");
    NSFEM_REQUIRE(mode == 0 || !s.vmg.lv.empty(), NSFEM_INVALID_ARGUMENT,
                  "velocity multigrid: set a velocity coarse space first");
    s.vmg_mode = mode;
}

void pressure_guess_base(Simulation &s);

// the mode's guess, then the coarse-grid correction when a coarse space is set
void pressure_guess(Simulation &s) {
    pressure_guess_base(s);
    s.coarse_iters = 0;
    s.coarse_seconds = 0.0;
    if (s.has_coarse && !s.pmg) coarse_correct(s);  // the preconditioner holds it
}

void pressure_guess_base(Simulation &s) {
    Device &d = *s.dev;
    switch (s.warm_mode) {
    case NSFEM_WARM_PREVIOUS:
        return;
    case NSFEM_WARM_ZERO:
        launch_fill(d, s.P, 0.0, s.p.ptr);
        return;
    case NSFEM_WARM_LEARNED:
        gnn_predict(*s.gnn, *s.graph, s.prhs.ptr, s.p.ptr);
        return;
    case NSFEM_WARM_LEARNED_OR_PREVIOUS:
        gnn_predict(*s.gnn, *s.graph, s.prhs.ptr, s.pguess.ptr);
        select_better_guess(d, s.mats[NSFEM_PRESSURE_STIFFNESS], s.prhs.ptr, s.pguess.ptr, s.p.ptr,
                            s.ptmp.ptr, s.pscal.ptr);
        return;
    case NSFEM_WARM_EXTRAPOLATE:
        // p0 = 2 p_n - p_(n-1) when it has the smaller residual; p_n is kept
        // for the next step's extrapolation
        launch_copy(d, s.P, s.p.ptr, s.p_keep.ptr);
        if (s.have_older) {
            launch_axpby(d, s.P, -1.0, 2.0, s.p_older.ptr, s.p.ptr, s.pguess.ptr);
            select_better_guess(d, s.mats[NSFEM_PRESSURE_STIFFNESS], s.prhs.ptr, s.pguess.ptr,
                                s.p.ptr, s.ptmp.ptr, s.pscal.ptr);
        }
        return;
    case NSFEM_WARM_PROJECTION:
        projection_guess(*s.dev, s.proj, s.prhs.ptr, s.p.ptr);
        return;
    case NSFEM_WARM_LEARNED_CORRECTION: {
        // r = b - A p_prev; guess = p_prev + network(r), kept only if its
        // residual is smaller than the previous pressure's
        Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
        launch_spmv(d, A, s.p.ptr, s.prhs.ptr, s.ptmp.ptr, 1);
        gnn_predict(*s.gnn, *s.graph, s.ptmp.ptr, s.pguess.ptr);
        launch_add(d, s.P, s.p.ptr, s.pguess.ptr);
        select_better_guess(d, A, s.prhs.ptr, s.pguess.ptr, s.p.ptr, s.ptmp.ptr, s.pscal.ptr);
        return;
    }
    case NSFEM_WARM_LEARNED_SEARCH:
    case NSFEM_WARM_LEARNED_PROJECTION: {
        // from the previous pressure (or its projection onto past
        // pressures): r = b - A x, then the network's correction d = net(r)
        // taken with the A-norm-optimal step, x += c d: no worse than x in
        // the A-norm whatever the prediction
        Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
        if (s.warm_mode == NSFEM_WARM_LEARNED_PROJECTION)
            projection_guess(*s.dev, s.proj, s.prhs.ptr, s.p.ptr);
        launch_spmv(d, A, s.p.ptr, s.prhs.ptr, s.ptmp.ptr, 1);
        gnn_predict(*s.gnn, *s.graph, s.ptmp.ptr, s.pguess.ptr);
        line_search_step(d, A, s.ptmp.ptr, s.pguess.ptr, s.p.ptr, s.pw.ptr, s.pscal.ptr);
        return;
    }
    }
}

// after the pressure solve: the extrapolating mode's p_n becomes p_(n-1);
// the projecting mode extends its basis
void pressure_after_solve(Simulation &s) {
    if (s.warm_mode == NSFEM_WARM_PROJECTION || s.warm_mode == NSFEM_WARM_LEARNED_PROJECTION) {
        projection_update(*s.dev, s.proj, s.mats[NSFEM_PRESSURE_STIFFNESS], s.p.ptr);
        return;
    }
    if (s.warm_mode != NSFEM_WARM_EXTRAPOLATE) return;
    std::swap(s.p_older, s.p_keep);
    s.have_older = true;
}

}  // namespace

void simulation_set_warm_start(Simulation &s, int mode, GnnNet *g) {
    NSFEM_REQUIRE(mode >= NSFEM_WARM_PREVIOUS && mode <= NSFEM_WARM_LEARNED_PROJECTION,
                  NSFEM_INVALID_ARGUMENT, "unknown warm-start mode");
    const bool learned = mode == NSFEM_WARM_LEARNED || mode == NSFEM_WARM_LEARNED_OR_PREVIOUS ||
                         mode == NSFEM_WARM_LEARNED_CORRECTION ||
                         mode == NSFEM_WARM_LEARNED_SEARCH ||
                         mode == NSFEM_WARM_LEARNED_PROJECTION;
    NSFEM_REQUIRE(!learned || (g && g->dev == s.dev), NSFEM_INVALID_ARGUMENT,
                  "a learned warm start needs a network on the simulation's device");
    Device &d = *s.dev;
    d.set_current();
    if (learned && !s.graph) {
        // the graph of the simulation's pressure matrix, read back once
        Matrix &A = s.mats[NSFEM_PRESSURE_STIFFNESS];
        std::vector<int32_t> rs(A.rows + 1), col((size_t)A.nnz);
        std::vector<double> val((size_t)A.nnz);
        NSFEM_HIP(hipMemcpy(rs.data(), A.row_start.ptr, rs.size() * sizeof(int32_t),
                            hipMemcpyDeviceToHost));
        NSFEM_HIP(hipMemcpy(col.data(), A.col.ptr, col.size() * sizeof(int32_t),
                            hipMemcpyDeviceToHost));
        NSFEM_HIP(hipMemcpy(val.data(), A.val.ptr, val.size() * sizeof(double),
                            hipMemcpyDeviceToHost));
        s.graph = gnn_graph_create(d, s.P, A.nnz, rs.data(), col.data(), val.data(),
                                   s.p_area.data(), s.p_xy.data());
    }
    if (mode == NSFEM_WARM_LEARNED_OR_PREVIOUS || mode == NSFEM_WARM_LEARNED_CORRECTION ||
        mode == NSFEM_WARM_EXTRAPOLATE || mode == NSFEM_WARM_LEARNED_SEARCH ||
        mode == NSFEM_WARM_LEARNED_PROJECTION) {
        s.pguess.resize(s.P);
        s.ptmp.resize(s.P);
        s.pscal.resize(2);
    }
    if (mode == NSFEM_WARM_LEARNED_SEARCH || mode == NSFEM_WARM_LEARNED_PROJECTION)
        s.pw.resize(s.P);
    if (mode == NSFEM_WARM_EXTRAPOLATE) {
        s.p_older.resize(s.P);
        s.p_keep.resize(s.P);
    }
    if (mode == NSFEM_WARM_PROJECTION || mode == NSFEM_WARM_LEARNED_PROJECTION) {
        // basis size: NSFEM_PROJECTION_DEPTH (default 16), 2 x depth + 3
        // pressure vectors of device memory
        static const int depth_env = [] {
            const char *e = std::getenv("NSFEM_PROJECTION_DEPTH");
            return e ? std::atoi(e) : 0;
        }();
        s.proj.reserve(s.P, depth_env > 0 ? std::min(depth_env, 256) : 16);
    }
    s.proj.dim = 0;
    s.have_older = false;
    s.gnn = learned ? g : nullptr;
    s.warm_mode = mode;
}

void simulation_forget_gnn(Simulation *s, GnnNet *g) {
    if (s->gnn == g) {
        s->gnn = nullptr;
        s->warm_mode = NSFEM_WARM_PREVIOUS;
    }
}

void simulation_copy_pressure(Simulation &s, double *rhs, double *p) {
    Device &d = *s.dev;
    d.set_current();
    if (rhs) launch_copy(d, s.P, s.prhs.ptr, rhs);
    if (p) launch_copy(d, s.P, s.p.ptr, p);
    NSFEM_HIP(hipStreamSynchronize(d.stream));
}

// Initial guess of the diffusion solves: the advected velocity u~ (the
// reference's) plus the A-norm projection of the correction u - u~ onto the
// span of past corrections, with A the diffusion matrix and the residual
// b - A u~ as the projected right-hand side.  u~ is kept in du (free until
// the velocity-correction step) for the basis update.
void diffusion_guess(Simulation &s) {
    Device &d = *s.dev;
    const int V = s.V;
    Matrix &A = s.mats[NSFEM_DIFFUSION];
    for (int c = 0; c < 2; ++c) {
        double *x = s.uv.ptr + (size_t)c * V, *del = s.vdel.ptr + (size_t)c * V;
        ProjBasis &B = s.vproj[c];
        launch_spmv(d, A, x, s.rhs.ptr + (size_t)c * V, B.w.ptr, 1);  // r0 = b - A u~
        if (B.dim == 0) launch_fill(d, V, 0.0, del);  // else the projection writes all of del
        projection_guess(d, B, B.w.ptr, del);
        launch_copy(d, V, x, s.du.ptr + (size_t)c * V);
        launch_add(d, V, del, x);
    }
}

// After the diffusion solves: the true corrections u - u~ extend the bases.
void diffusion_after_solve(Simulation &s) {
    Device &d = *s.dev;
    const int V = s.V;
    for (int c = 0; c < 2; ++c) {
        double *del = s.vdel.ptr + (size_t)c * V;
        launch_axpby(d, V, 1.0, -1.0, s.uv.ptr + (size_t)c * V, s.du.ptr + (size_t)c * V, del);
        projection_update(d, s.vproj[c], s.mats[NSFEM_DIFFUSION], del);
    }
}

// Pairs the u and v columns of a velocity system (one 2-column solve: the
// matrix is streamed once for both) unless the last step's column counts
// were unbalanced: the 2-column CG runs both columns to the slower one's
// count, so when one needs more than twice the other's iterations (+8) the
// columns are solved one after the other.  On the 2048^2 cavity the
// diffusion columns take ~28 and ~124 iterations (separately 107 ms a step
// against 120 paired), the mass columns ~28 and ~24 (paired 34 ms against 45).
// Whether the next step solves a velocity system's u and v columns as one
// two-column CG.  Once both ways have been timed on this simulation, by
// cost: pairing takes max(a, b) two-column iterations, the separate solves
// a + b one-column ones, a and b the last step's counts (on the 2048^2
// cavity's P2 matrices a two-column iteration costs 1.95 one-column ones,
// so only nearly equal counts pay; on small systems, bound by launch and
// synchronisation latency, pairing pays more often).  Before that, counts
// within 2x (+8) of each other pair, and a system measured only paired is
// run apart once to time it.  The way not taken runs again once it has not
// run for kPairRetry steps: a first measurement can hold one-time costs
// (the first launch of a kernel form, lazily built matrix copies), and the
// counts drift as the flow develops.
constexpr int kPairRetry = 16;

bool pair_columns(Simulation &s, int sys) {
    if (!s.opt.pair_velocity) return false;
    const int64_t a = s.vel_iters[sys][0], b = s.vel_iters[sys][1];
    if (a <= 0 || b <= 0) return true;
    const double t1 = s.pair_sec1[sys], t2 = s.pair_sec2[sys];
    if (t1 > 0.0 && t2 > 0.0) {
        const bool pair = t2 * (double)std::max(a, b) <= t1 * (double)(a + b);
        return s.pair_age[sys][pair ? 0 : 1] >= kPairRetry ? !pair : pair;
    }
    const bool close = std::max(a, b) <= 2 * std::min(a, b) + 8;
    return close && !(t2 > 0.0 && t1 == 0.0);
}

// Records the timing of a velocity system's step for pair_columns: the
// least seconds per iteration seen each way (a first launch of a kernel
// form can carry its one-time loading in its event time; the cost per
// iteration of a form on a matrix is otherwise steady).
void record_pairing(Simulation &s, int sys, bool paired, const SolveOut &u, const SolveOut &v) {
    s.pair_age[sys][paired ? 1 : 0] = 0;
    ++s.pair_age[sys][paired ? 0 : 1];
    double &t = paired ? s.pair_sec2[sys] : s.pair_sec1[sys];
    const int64_t its = paired ? std::max(u.iters[0], u.iters[1]) : u.iters[0] + v.iters[0];
    if (its <= 0) return;
    const double sec = (paired ? u.seconds : u.seconds + v.seconds) / (double)its;
    t = t > 0.0 ? std::min(t, sec) : sec;
}

// Whether column c of the diffusion solve runs V-cycle preconditioned CG
// this step: never without a hierarchy or in mode 0, always in mode 2; in
// mode 1 the way whose solves of this column took less time so far
// (exponential averages; the V-cycle first, then the plain CG once so that
// both are measured), except that the slower way runs again once it has not
// run for kVmgRetry steps (the columns' needs drift as the flow develops).
constexpr int kVmgRetry = 16;

bool velocity_pcg(const Simulation &s, int c) {
    if (s.vmg_mode == 0 || s.vmg.lv.empty()) return false;
    if (s.vmg_mode == 2) return true;
    const double plain = s.vmg_sec[c][0], mg = s.vmg_sec[c][1];
    if (mg == 0.0) return true;
    if (plain == 0.0) return false;
    const int faster = mg <= plain ? 1 : 0;
    return (s.vmg_age[c][1 - faster] >= kVmgRetry ? 1 - faster : faster) == 1;
}

void record_velocity_way(Simulation &s, int c, bool mg, double seconds) {
    const int w = mg ? 1 : 0;
    double &t = s.vmg_sec[c][w];
    // a way that had not run for kVmgRetry steps is measured afresh
    t = t > 0.0 && s.vmg_age[c][w] < kVmgRetry ? 0.5 * (t + seconds) : seconds;
    s.vmg_age[c][w] = 0;
    ++s.vmg_age[c][1 - w];
}

void simulation_step(Simulation &s, int steps, nsfem_step_stats *stats) {
    Device &d = *s.dev;
    d.set_current();
    const int V = s.V, P = s.P;
    double *u = s.uv.ptr, *v = s.uv.ptr + V;
    double *ut = s.utv.ptr, *vt = s.utv.ptr + V;
    Matrix &M = s.mats[NSFEM_VELOCITY_MASS];
    Matrix &BD = s.mats[NSFEM_VELOCITY_DIVERGENCE];
    Matrix &PD = s.mats[NSFEM_PRESSURE_DIVERGENCE];
    hipEvent_t e0, e1, ea;
    NSFEM_HIP(hipEventCreate(&e0));
    NSFEM_HIP(hipEventCreate(&e1));
    NSFEM_HIP(hipEventCreate(&ea));
    for (int step = 0; step < steps; ++step) {
        NSFEM_HIP(hipEventRecord(e0, d.stream));
        // 1. advection
        launch_advect_f64(d, s.tree, V, u, v, s.opt.dt, ut, vt);
        NSFEM_HIP(hipEventRecord(ea, d.stream));
        // 2. diffusion
        launch_spmv(d, M, ut, s.c_uv.ptr, s.rhs.ptr, 2);
        launch_spmv(d, M, vt, s.c_uv.ptr + V, s.rhs.ptr + V, 2);
        launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_u.ptr, s.rhs.ptr);
        launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_v.ptr, s.rhs.ptr + V);
        launch_copy(d, 2 * V, s.utv.ptr, s.uv.ptr);
        SolveOut o[5]{};
        if (s.vproj[0].max > 0) diffusion_guess(s);
        const bool vmg_on = s.vmg_mode != 0 && !s.vmg.lv.empty();
        const bool pcg[2] = {velocity_pcg(s, 0), velocity_pcg(s, 1)};
        const bool pair_d = !vmg_on && pair_columns(s, 0);
        if (pair_d) {
            solve(s, NSFEM_DIFFUSION, s.rhs.ptr, s.uv.ptr, 2, o[0]);
            o[1].iters[0] = o[0].iters[1];
            o[1].refines[0] = o[0].refines[1];
            o[1].rel[0] = o[0].rel[1];
        } else {
            for (int c = 0; c < 2; ++c) {
                if (pcg[c])
                    pcg_velocity(s, c, o[c]);
                else
                    solve(s, NSFEM_DIFFUSION, s.rhs.ptr + (size_t)c * V, c ? v : u, 1, o[c]);
            }
        }
        s.vel_iters[0][0] = o[0].iters[0];
        s.vel_iters[0][1] = o[1].iters[0];
        if (vmg_on) {
            for (int c = 0; c < 2; ++c) record_velocity_way(s, c, pcg[c], o[c].seconds);
        } else {
            record_pairing(s, 0, pair_d, o[0], o[1]);
        }
        if (s.vproj[0].max > 0) diffusion_after_solve(s);
        // 3. pressure
        launch_spmv(d, BD, s.uv.ptr, s.c_p.ptr, s.prhs.ptr, 2);
        launch_scatter(d, s.npb, s.pb_idx.ptr, s.pb_val.ptr, s.prhs.ptr);
        pressure_guess(s);
        const int64_t coarse_iters = s.coarse_iters;
        const double coarse_seconds = s.coarse_seconds;
        if (s.pmg && s.has_coarse)
            pcg_pressure(s, o[2]);
        else
            solve(s, NSFEM_PRESSURE_STIFFNESS, s.prhs.ptr, s.p.ptr, 1, o[2]);
        pressure_after_solve(s);
        // 4. velocity correction
        launch_spmv(d, PD, s.p.ptr, nullptr, s.rhs.ptr, 0);
        for (int c = 0; c < 2; ++c) {
            ProjBasis &B = s.mproj[c];
            double *x0 = s.du.ptr + (size_t)c * V;
            // the velocity input, unless a non-empty basis writes all of x0
            if (B.max == 0 || B.dim == 0) launch_copy(d, V, s.uv.ptr + (size_t)c * V, x0);
            if (B.max > 0) projection_guess(d, B, s.rhs.ptr + (size_t)c * V, x0);
        }
        const bool pair_m = pair_columns(s, 1);
        if (pair_m) {
            solve(s, NSFEM_VELOCITY_MASS, s.rhs.ptr, s.du.ptr, 2, o[3]);
            o[4].iters[0] = o[3].iters[1];
            o[4].refines[0] = o[3].refines[1];
            o[4].rel[0] = o[3].rel[1];
        } else {
            solve(s, NSFEM_VELOCITY_MASS, s.rhs.ptr, s.du.ptr, 1, o[3]);
            solve(s, NSFEM_VELOCITY_MASS, s.rhs.ptr + V, s.du.ptr + V, 1, o[4]);
        }
        s.vel_iters[1][0] = o[3].iters[0];
        s.vel_iters[1][1] = o[4].iters[0];
        if (s.mproj[0].max > 0)
            for (int c = 0; c < 2; ++c) projection_update(d, s.mproj[c], M, s.du.ptr + (size_t)c * V);
        record_pairing(s, 1, pair_m, o[3], o[4]);
        // 5. update and Dirichlet values
        launch_add(d, 2 * V, s.du.ptr, s.uv.ptr);
        launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_u.ptr, u);
        launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_v.ptr, v);
        NSFEM_HIP(hipEventRecord(e1, d.stream));
        NSFEM_HIP(hipEventSynchronize(e1));
        if (stats) {
            float ms = 0.f, ma = 0.f;
            NSFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
            NSFEM_HIP(hipEventElapsedTime(&ma, e0, ea));
            stats->steps += 1;
            stats->seconds += ms * 1e-3;
            stats->advect_seconds += ma * 1e-3;
            for (int k = 0; k < 5; ++k) {
                stats->cg_iterations[k] += o[k].iters[0];
                stats->cg_refinements[k] += o[k].refines[0];
                stats->max_relative_residual[k] =
                    std::max(stats->max_relative_residual[k], o[k].rel[0]);
            }
            const double t[5] = {o[0].seconds, pair_d ? 0.0 : o[1].seconds, o[2].seconds,
                                 o[3].seconds, pair_m ? 0.0 : o[4].seconds};
            for (int k = 0; k < 5; ++k) stats->cg_seconds[k] += t[k];
            const int64_t li[5] = {
                pair_d ? std::max(o[0].iters[0], o[0].iters[1]) : o[0].iters[0],
                pair_d ? 0 : o[1].iters[0], o[2].iters[0],
                pair_m ? std::max(o[3].iters[0], o[3].iters[1]) : o[3].iters[0],
                pair_m ? 0 : o[4].iters[0]};
            const int64_t ln[5] = {1, pair_d ? 0 : 1, 1, 1, pair_m ? 0 : 1};
            for (int k = 0; k < 5; ++k) {
                stats->cg_launches[k] += ln[k];
                stats->cg_launch_iterations[k] += li[k];
            }
            const bool vc[3] = {pcg[0], pcg[1], s.pmg && s.has_coarse};
            for (int k = 0; k < 3; ++k) {
                if (!vc[k]) continue;
                stats->vcycle_solves[k] += 1;
                stats->vcycle_iterations[k] += o[k].iters[0];
                stats->vcycle_seconds[k] += o[k].seconds;
            }
            stats->coarse_iterations += coarse_iters;
            stats->coarse_seconds += coarse_seconds;
        }
    }
    (void)P;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(ea);
}

// ------------------------------------------------------------ batches
// Several simulations stepped together: the per-instance stages (advection,
// right-hand sides, boundary values) are launched instance by instance, and
// each of the three solve kinds (u/v diffusion as a 2-column solve, pressure,
// u/v mass as a 2-column solve) runs for every instance in one persistent
// launch (run_cg_batch).
void simulation_step_batch(const std::vector<Simulation *> &S, int steps,
                           nsfem_step_stats *stats) {
    NSFEM_REQUIRE(!S.empty(), NSFEM_INVALID_ARGUMENT, "empty batch");
    Device &d = *S[0]->dev;
    for (Simulation *s : S) {
        NSFEM_REQUIRE(s->dev == &d, NSFEM_INVALID_ARGUMENT,
                      "a batch needs simulations of one device");
        NSFEM_REQUIRE(s->opt.precision == S[0]->opt.precision, NSFEM_INVALID_ARGUMENT,
                      "a batch needs one precision");
        NSFEM_REQUIRE(!s->pmg, NSFEM_INVALID_ARGUMENT,
                      "the V-cycle preconditioned pressure solve steps one simulation at a time");
        NSFEM_REQUIRE(s->vmg_mode == 0, NSFEM_INVALID_ARGUMENT,
                      "the V-cycle preconditioned diffusion solves step one simulation at a time");
        NSFEM_REQUIRE(s->vproj[0].max == 0, NSFEM_INVALID_ARGUMENT,
                      "the velocity projection of the diffusion guesses steps one simulation "
                      "at a time");
        NSFEM_REQUIRE(s->mproj[0].max == 0, NSFEM_INVALID_ARGUMENT,
                      "the projection of the mass guesses steps one simulation at a time");
    }
    d.set_current();
    const size_t I = S.size();
    const int scale = S[0]->batch_block_scale;
    const int scale_v = S[0]->batch_block_scale_v;
    std::vector<nsfem_cg_stats> st(2 * I);
    std::vector<CgRequest> rq(I);
    auto request = [&](Simulation &s, int slot, const double *b, double *x, int ncol) {
        CgRequest r{};
        r.m = &s.mats[slot];
        r.b = b;
        r.x = x;
        r.ncol = ncol;
        r.max_iter = s.opt.max_iterations;
        r.rtol = s.opt.rtol;
        r.atol = 0.0;
        r.precision = s.opt.precision;
        set_update_policy(r, slot);
        return r;
    };
    auto check = [&](int status, const char *what) {
        if (status != NSFEM_OK)
            fail(status, std::string("batched CG on the ") + what + " matrices did not converge");
    };
    hipEvent_t e0, e1;
    NSFEM_HIP(hipEventCreate(&e0));
    NSFEM_HIP(hipEventCreate(&e1));
    for (int step = 0; step < steps; ++step) {
        NSFEM_HIP(hipEventRecord(e0, d.stream));
        double sec[3] = {0, 0, 0};
        for (Simulation *sp : S) {
            Simulation &s = *sp;
            const int V = s.V;
            Matrix &M = s.mats[NSFEM_VELOCITY_MASS];
            launch_advect_f64(d, s.tree, V, s.uv.ptr, s.uv.ptr + V, s.opt.dt, s.utv.ptr,
                              s.utv.ptr + V);
            launch_spmv(d, M, s.utv.ptr, s.c_uv.ptr, s.rhs.ptr, 2);
            launch_spmv(d, M, s.utv.ptr + V, s.c_uv.ptr + V, s.rhs.ptr + V, 2);
            launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_u.ptr, s.rhs.ptr);
            launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_v.ptr, s.rhs.ptr + V);
            launch_copy(d, 2 * V, s.utv.ptr, s.uv.ptr);
        }
        for (size_t i = 0; i < I; ++i)
            rq[i] = request(*S[i], NSFEM_DIFFUSION, S[i]->rhs.ptr, S[i]->uv.ptr, 2);
        std::vector<nsfem_cg_stats> sd(2 * I);
        check(run_cg_batch(d, rq, sd.data(), scale_v, &sec[0]), "diffusion");
        // pressure guesses; the instances' coarse corrections share one
        // batched launch
        std::vector<size_t> with_coarse;
        std::vector<CgRequest> crq;
        for (size_t i = 0; i < I; ++i) {
            Simulation &s = *S[i];
            launch_spmv(d, s.mats[NSFEM_VELOCITY_DIVERGENCE], s.uv.ptr, s.c_p.ptr, s.prhs.ptr, 2);
            launch_scatter(d, s.npb, s.pb_idx.ptr, s.pb_val.ptr, s.prhs.ptr);
            pressure_guess_base(s);
            s.coarse_iters = 0;
            s.coarse_seconds = 0.0;
            if (s.has_coarse && s.mg.lv.size() > 1) {
                coarse_correct(s);  // multigrid: its own V-cycles
            } else if (s.has_coarse) {
                coarse_prepare(s);
                with_coarse.push_back(i);
                crq.push_back(coarse_request(s));
            }
        }
        if (!crq.empty()) {
            std::vector<nsfem_cg_stats> cst(2 * crq.size());
            double csec = 0.0;
            run_cg_batch(d, crq, cst.data(), scale, &csec);
            for (size_t k = 0; k < with_coarse.size(); ++k)
                coarse_finish(*S[with_coarse[k]], cst[2 * k].status, cst[2 * k].iterations, csec);
        }
        if (stats)
            for (size_t i = 0; i < I; ++i) {
                stats[i].coarse_iterations += S[i]->coarse_iters;
                stats[i].coarse_seconds += S[i]->coarse_seconds;
            }
        for (size_t i = 0; i < I; ++i)
            rq[i] = request(*S[i], NSFEM_PRESSURE_STIFFNESS, S[i]->prhs.ptr, S[i]->p.ptr, 1);
        std::vector<nsfem_cg_stats> sp_(2 * I);
        check(run_cg_batch(d, rq, sp_.data(), scale, &sec[1]), "pressure");
        for (Simulation *sp : S) pressure_after_solve(*sp);
        for (Simulation *sp : S) {
            Simulation &s = *sp;
            launch_spmv(d, s.mats[NSFEM_PRESSURE_DIVERGENCE], s.p.ptr, nullptr, s.rhs.ptr, 0);
            launch_copy(d, 2 * s.V, s.uv.ptr, s.du.ptr);
        }
        for (size_t i = 0; i < I; ++i)
            rq[i] = request(*S[i], NSFEM_VELOCITY_MASS, S[i]->rhs.ptr, S[i]->du.ptr, 2);
        std::vector<nsfem_cg_stats> sm(2 * I);
        check(run_cg_batch(d, rq, sm.data(), scale_v, &sec[2]), "velocity mass");
        for (Simulation *sp : S) {
            Simulation &s = *sp;
            launch_add(d, 2 * s.V, s.du.ptr, s.uv.ptr);
            launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_u.ptr, s.uv.ptr);
            launch_scatter(d, s.nvb, s.vb_idx.ptr, s.vb_v.ptr, s.uv.ptr + s.V);
        }
        NSFEM_HIP(hipEventRecord(e1, d.stream));
        NSFEM_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NSFEM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (stats) {
            for (size_t i = 0; i < I; ++i) {
                nsfem_step_stats &o = stats[i];
                o.steps += 1;
                o.seconds += ms * 1e-3;  // the whole batch's step
                const nsfem_cg_stats *kinds[5] = {&sd[2 * i], &sd[2 * i + 1], &sp_[2 * i],
                                                  &sm[2 * i], &sm[2 * i + 1]};
                for (int k = 0; k < 5; ++k) {
                    const nsfem_cg_stats &c = *kinds[k];
                    o.cg_iterations[k] += c.iterations;
                    o.cg_refinements[k] += c.refinements;
                    const double rel = c.rhs_norm > 0 ? c.final_residual / c.rhs_norm
                                                      : c.final_residual;
                    o.max_relative_residual[k] = std::max(o.max_relative_residual[k], rel);
                }
                o.cg_seconds[0] += sec[0];
                o.cg_seconds[2] += sec[1];
                o.cg_seconds[3] += sec[2];
                const int64_t li[5] = {std::max<int64_t>(sd[2 * i].iterations,
                                                         sd[2 * i + 1].iterations),
                                       0, sp_[2 * i].iterations,
                                       std::max<int64_t>(sm[2 * i].iterations,
                                                         sm[2 * i + 1].iterations),
                                       0};
                const int64_t ln[5] = {1, 0, 1, 1, 0};
                for (int k = 0; k < 5; ++k) {
                    o.cg_launches[k] += ln[k];
                    o.cg_launch_iterations[k] += li[k];
                }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

// Picks the wave-tile shape and the blocks per instance of batched solves:
// every candidate runs `iters` iterations of the batched pressure solve on a
// fixed right-hand side (b = 1, x0 = 0), and the fastest is applied to all
// matrices of the batch.
void simulation_autotune_batch(const std::vector<Simulation *> &S, nsfem_batch_tuning *out) {
    NSFEM_REQUIRE(!S.empty(), NSFEM_INVALID_ARGUMENT, "empty batch");
    Device &d = *S[0]->dev;
    for (Simulation *s : S)
        NSFEM_REQUIRE(s->dev == &d, NSFEM_INVALID_ARGUMENT, "a batch needs simulations of one device");
    d.set_current();
    const size_t I = S.size();
    size_t total = 0;
    std::vector<size_t> off(I);
    for (size_t i = 0; i < I; ++i) {
        off[i] = total;
        total += (size_t)S[i]->P;
    }
    size_t total_v = 0;
    std::vector<size_t> off_v(I);
    for (size_t i = 0; i < I; ++i) {
        off_v[i] = total_v;
        total_v += 2 * (size_t)S[i]->V;
    }
    DeviceArray<double> b, x;
    b.resize(std::max(total, total_v));
    x.resize(std::max(total, total_v));
    launch_fill(d, (int)std::max(total, total_v), 1.0, b.ptr);
    const int iters = 40;
    struct Shape { int rows, nnz; };
    const Shape shapes[3] = {{64, 512}, {32, 256}, {16, 128}};
    const int scales[3] = {1, 2, 4};
    double best = 1e300;
    Shape best_shape = shapes[0];
    int best_scale = 1;
    std::vector<nsfem_cg_stats> st(2 * I);
    for (const Shape &sh : shapes) {
        for (Simulation *s : S)
            for (Matrix &m : s->mats)
                if (m.ready && m.rows == m.cols) build_wave_tiles(m, sh.rows, sh.nnz);
        for (int sc : scales) {
            std::vector<CgRequest> rq(I);
            for (size_t i = 0; i < I; ++i) {
                CgRequest r{};
                r.m = &S[i]->mats[NSFEM_PRESSURE_STIFFNESS];
                r.b = b.ptr + off[i];
                r.x = x.ptr + off[i];
                r.ncol = 1;
                r.max_iter = iters;
                r.rtol = 0.0;
                r.atol = 0.0;
                r.precision = S[0]->opt.precision;
                rq[i] = r;
            }
            launch_fill(d, (int)total, 0.0, x.ptr);
            double sec = 0.0;
            run_cg_batch(d, rq, st.data(), sc, &sec, false);
            if (std::getenv("NSFEM_CG_VERBOSE"))
                std::fprintf(stderr, "[nsfem] batch autotune rows=%d nnz=%d scale=%d: %.3f ms\n",
                             sh.rows, sh.nnz, sc, sec * 1e3);
            if (sec < best) {
                best = sec;
                best_shape = sh;
                best_scale = sc;
            }
        }
    }
    for (Simulation *s : S) {
        for (Matrix &m : s->mats)
            if (m.ready && m.rows == m.cols) build_wave_tiles(m, best_shape.rows, best_shape.nnz);
        s->batch_block_scale = best_scale;
    }
    // the two-column velocity solves, with the chosen tile shape: their
    // matrices have ~4x the rows of the pressure matrix, so the share of
    // blocks that suits the pressure can overfill the device for them
    double best_v = 1e300;
    int best_scale_v = 1;
    for (int sc : scales) {
        std::vector<CgRequest> rq(I);
        for (size_t i = 0; i < I; ++i) {
            CgRequest r{};
            r.m = &S[i]->mats[NSFEM_DIFFUSION];
            r.b = b.ptr + off_v[i];
            r.x = x.ptr + off_v[i];
            r.ncol = 2;
            r.max_iter = iters;
            r.rtol = 0.0;
            r.atol = 0.0;
            r.precision = S[0]->opt.precision;
            set_update_policy(r, NSFEM_DIFFUSION);
            rq[i] = r;
        }
        launch_fill(d, (int)total_v, 0.0, x.ptr);
        double sec = 0.0;
        run_cg_batch(d, rq, st.data(), sc, &sec, false);
        if (std::getenv("NSFEM_CG_VERBOSE"))
            std::fprintf(stderr, "[nsfem] batch autotune velocity scale=%d: %.3f ms\n", sc, sec * 1e3);
        if (sec < best_v) {
            best_v = sec;
            best_scale_v = sc;
        }
    }
    for (Simulation *s : S) s->batch_block_scale_v = best_scale_v;
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    if (out) {
        out->rows_per_wave = best_shape.rows;
        out->nnz_per_wave = best_shape.nnz;
        out->block_scale = best_scale;
        out->velocity_block_scale = best_scale_v;
        out->seconds_per_iteration = best / iters;
    }
}

// velocity fields in the caller's node order (the device holds them in the
// internal order of velocity_order)
void simulation_get(Simulation &s, double *u, double *v, double *p) {
    Device &d = *s.dev;
    d.set_current();
    const size_t V = (size_t)s.V;
    std::vector<double> tmp;
    if (u || v) {
        tmp.resize(2 * V);
        NSFEM_HIP(hipMemcpyAsync(tmp.data(), s.uv.ptr, sizeof(double) * 2 * V,
                                 hipMemcpyDeviceToHost, d.stream));
    }
    if (p) NSFEM_HIP(hipMemcpyAsync(p, s.p.ptr, sizeof(double) * s.P, hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    for (size_t i = 0; i < V && (u || v); ++i) {
        if (u) u[i] = tmp[s.vpos[i]];
        if (v) v[i] = tmp[V + s.vpos[i]];
    }
}

void simulation_set(Simulation &s, const double *u, const double *v, const double *p) {
    Device &d = *s.dev;
    d.set_current();
    const size_t V = (size_t)s.V;
    std::vector<double> tmp;
    if (u || v) {
        tmp.resize(2 * V);
        NSFEM_HIP(hipMemcpyAsync(tmp.data(), s.uv.ptr, sizeof(double) * 2 * V,
                                 hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        for (size_t i = 0; i < V; ++i) {
            if (u) tmp[s.vpos[i]] = u[i];
            if (v) tmp[V + s.vpos[i]] = v[i];
        }
        NSFEM_HIP(hipMemcpyAsync(s.uv.ptr, tmp.data(), sizeof(double) * 2 * V,
                                 hipMemcpyHostToDevice, d.stream));
    }
    if (p) NSFEM_HIP(hipMemcpyAsync(s.p.ptr, p, sizeof(double) * s.P, hipMemcpyHostToDevice, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
}

void simulation_detach(Simulation *s) {
    for (Matrix &m : s->mats) matrix_release(m);
    kdtree_release(s->tree);
    if (s->graph) gnn_graph_release(*s->graph);
    s->gnn = nullptr;
    s->warm_mode = NSFEM_WARM_PREVIOUS;
    s->pguess.release();
    s->p_older.release();
    s->p_keep.release();
    s->proj.release();
    s->cproj.release();
    s->vproj[0].release();
    s->vproj[1].release();
    s->mproj[0].release();
    s->mproj[1].release();
    s->vdel.release();
    release_velocity_multigrid(*s);
    s->mg.release();
    release_pressure_multigrid(*s);
    s->cy.release();
    s->has_coarse = false;
    s->ptmp.release();
    s->pscal.release();
    s->uv.release();
    s->utv.release();
    s->rhs.release();
    s->prhs.release();
    s->p.release();
    s->du.release();
    s->c_uv.release();
    s->c_p.release();
    s->vb_idx.release();
    s->pb_idx.release();
    s->vb_u.release();
    s->vb_v.release();
    s->pb_val.release();
    s->dev = nullptr;
}

void simulation_destroy(Simulation *s) {
    if (s && s->dev) {
        std::lock_guard<std::mutex> lk(s->dev->buf_mtx);
        s->dev->sims.erase(s);
    }
    delete s->graph;
    delete s;
}

}  // namespace nsfem

struct nsfem_simulation {
    nsfem::Simulation *impl;
};

using namespace nsfem;

extern "C" {

int nsfem_simulation_create(nsfem_device *h, int V, int P, const float *xy, int E,
                            const int32_t *el, int nvb, const int32_t *vb, const double *gu,
                            const double *gv, int npb, const int32_t *pb, const double *gp,
                            const nsfem_sim_options *opt, nsfem_simulation **out) {
    return guard([&] {
        NSFEM_REQUIRE(h && out && opt && xy && el, NSFEM_INVALID_ARGUMENT, "null argument");
        NSFEM_REQUIRE((nvb == 0 || (vb && gu && gv)) && (npb == 0 || (pb && gp)),
                      NSFEM_INVALID_ARGUMENT, "null boundary arrays");
        std::lock_guard<std::mutex> lk(h->impl.mtx);
        auto *w = new nsfem_simulation();
        try {
            w->impl = simulation_create(h->impl, V, P, xy, E, el, nvb, vb, gu, gv, npb, pb, gp,
                                        *opt);
        } catch (...) {
            delete w;
            throw;
        }
        *out = w;
    });
}

int nsfem_simulation_destroy(nsfem_simulation *s) {
    return guard([&] {
        if (!s) return;
        if (s->impl && s->impl->dev) {
            (void)hipSetDevice(s->impl->dev->index);
            (void)hipStreamSynchronize(s->impl->dev->stream);
        }
        simulation_destroy(s->impl);
        delete s;
    });
}

int nsfem_simulation_step(nsfem_simulation *s, int steps, nsfem_step_stats *stats) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && steps >= 0, NSFEM_INVALID_ARGUMENT, "bad simulation call");
        NSFEM_REQUIRE(s->impl->dev, NSFEM_INVALID_ARGUMENT, "simulation's device was destroyed");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_step(*s->impl, steps, stats);
    });
}

int nsfem_simulation_get_fields(nsfem_simulation *s, double *u, double *v, double *p) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        simulation_get(*s->impl, u, v, p);
    });
}

int nsfem_simulation_set_options(nsfem_simulation *s, const nsfem_sim_options *opt) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && opt, NSFEM_INVALID_ARGUMENT, "null argument");
        NSFEM_REQUIRE(opt->precision == NSFEM_PRECISION_FP64 ||
                          opt->precision == NSFEM_PRECISION_FP32_REFINED,
                      NSFEM_INVALID_ARGUMENT, "unknown precision");
        NSFEM_REQUIRE(opt->rtol >= 0, NSFEM_INVALID_ARGUMENT, "negative rtol");
        s->impl->opt.rtol = opt->rtol;
        s->impl->opt.max_iterations = opt->max_iterations;
        s->impl->opt.precision = opt->precision;
        s->impl->opt.pair_velocity = opt->pair_velocity;
        s->impl->opt.velocity_jacobi = opt->velocity_jacobi;
        // the solves' cost per iteration may change: time both ways again
        for (int k = 0; k < 2; ++k) {
            s->impl->pair_sec1[k] = s->impl->pair_sec2[k] = 0.0;
            s->impl->pair_age[k][0] = s->impl->pair_age[k][1] = 0;
        }
        for (auto &t : s->impl->vmg_sec) t[0] = t[1] = 0.0;
    });
}

int nsfem_simulation_reset(nsfem_simulation *s) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_reset(*s->impl);
    });
}

int nsfem_simulation_set_warm_start(nsfem_simulation *s, int mode, nsfem_gnn *g) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        NSFEM_REQUIRE(!g || g->impl, NSFEM_INVALID_ARGUMENT, "network of a destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_warm_start(*s->impl, mode, g ? g->impl : nullptr);
    });
}

int nsfem_simulation_set_coarse_space(nsfem_simulation *s, int coarse_count, int64_t nnz,
                                      const int32_t *row_start, const int32_t *col,
                                      const double *val, double coarse_rtol) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_coarse_space(*s->impl, coarse_count, nnz, row_start, col, val,
                                    coarse_rtol);
    });
}

int nsfem_simulation_add_coarse_level(nsfem_simulation *s, int count, int64_t nnz,
                                      const int32_t *row_start, const int32_t *col,
                                      const double *val) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_add_coarse_level(*s->impl, count, nnz, row_start, col, val);
    });
}

int nsfem_simulation_set_velocity_coarse_space(nsfem_simulation *s, int coarse_count,
                                               int64_t nnz, const int32_t *row_start,
                                               const int32_t *col, const double *val) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_velocity_coarse_space(*s->impl, coarse_count, nnz, row_start, col, val);
    });
}

int nsfem_simulation_add_velocity_coarse_level(nsfem_simulation *s, int count, int64_t nnz,
                                               const int32_t *row_start, const int32_t *col,
                                               const double *val) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_add_velocity_coarse_level(*s->impl, count, nnz, row_start, col, val);
    });
}

int nsfem_simulation_set_velocity_multigrid(nsfem_simulation *s, int mode) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_velocity_multigrid(*s->impl, mode);
    });
}

int nsfem_simulation_set_pressure_multigrid(nsfem_simulation *s, int enable) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_pressure_multigrid(*s->impl, enable);
    });
}

int nsfem_simulation_set_velocity_projection(nsfem_simulation *s, int depth) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_velocity_projection(*s->impl, depth);
    });
}

int nsfem_simulation_set_mass_projection(nsfem_simulation *s, int depth) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_set_mass_projection(*s->impl, depth);
    });
}

int nsfem_simulation_copy_pressure_system(nsfem_simulation *s, nsfem_buffer *rhs,
                                          nsfem_buffer *p) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        const int64_t need = (int64_t)s->impl->P * (int64_t)sizeof(double);
        for (nsfem_buffer *b : {rhs, p})
            NSFEM_REQUIRE(!b || (b->impl.dev == s->impl->dev && (int64_t)b->impl.mem.bytes() >= need),
                          NSFEM_INVALID_ARGUMENT,
                          "buffer too small or on another device (needs pressure-count doubles)");
        std::lock_guard<std::mutex> lk(s->impl->dev->mtx);
        simulation_copy_pressure(*s->impl, rhs ? (double *)rhs->impl.mem.ptr : nullptr,
                                 p ? (double *)p->impl.mem.ptr : nullptr);
    });
}

static std::vector<Simulation *> batch_of(nsfem_simulation *const *sims, int count) {
    NSFEM_REQUIRE(sims && count > 0, NSFEM_INVALID_ARGUMENT, "empty batch");
    std::vector<Simulation *> S;
    for (int i = 0; i < count; ++i) {
        NSFEM_REQUIRE(sims[i] && sims[i]->impl && sims[i]->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device in the batch");
        for (Simulation *o : S)
            NSFEM_REQUIRE(o != sims[i]->impl, NSFEM_INVALID_ARGUMENT,
                          "a simulation appears twice in the batch");
        S.push_back(sims[i]->impl);
    }
    return S;
}

int nsfem_simulation_step_batch(nsfem_simulation *const *sims, int count, int steps,
                                nsfem_step_stats *stats) {
    return guard([&] {
        std::vector<Simulation *> S = batch_of(sims, count);
        NSFEM_REQUIRE(steps >= 0, NSFEM_INVALID_ARGUMENT, "negative step count");
        std::lock_guard<std::mutex> lk(S[0]->dev->mtx);
        simulation_step_batch(S, steps, stats);
    });
}

int nsfem_simulation_autotune_batch(nsfem_simulation *const *sims, int count,
                                    nsfem_batch_tuning *out) {
    return guard([&] {
        std::vector<Simulation *> S = batch_of(sims, count);
        std::lock_guard<std::mutex> lk(S[0]->dev->mtx);
        simulation_autotune_batch(S, out);
    });
}

int nsfem_simulation_matrix_info(nsfem_simulation *s, int slot, int *rows, int *cols,
                                 int64_t *nnz) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        NSFEM_REQUIRE(slot >= 0 && slot < 5, NSFEM_INVALID_ARGUMENT, "matrix id out of range");
        const Matrix &m = s->impl->mats[slot];
        if (rows) *rows = m.rows;
        if (cols) *cols = m.cols;
        if (nnz) *nnz = m.nnz;
    });
}

int nsfem_simulation_sp_r_mult(nsfem_simulation *s, int slot, const nsfem_buffer *lhs,
                               const nsfem_buffer *x, nsfem_buffer *y) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        NSFEM_REQUIRE(slot >= 0 && slot < 5, NSFEM_INVALID_ARGUMENT, "matrix id out of range");
        Simulation &sim = *s->impl;
        Matrix &m = sim.mats[slot];
        auto ok = [&](const nsfem_buffer *b, int64_t n) {
            return b && b->impl.dev == sim.dev && (int64_t)b->impl.mem.bytes() >= n * 8;
        };
        NSFEM_REQUIRE(ok(x, m.cols) && ok(y, m.rows) && (!lhs || ok(lhs, m.rows)),
                      NSFEM_INVALID_ARGUMENT,
                      "buffers must be on the simulation's device and hold cols (x) / rows doubles");
        NSFEM_REQUIRE(x->impl.mem.ptr != y->impl.mem.ptr, NSFEM_INVALID_ARGUMENT,
                      "x must not alias y");
        std::lock_guard<std::mutex> lk(sim.dev->mtx);
        sim.dev->set_current();
        // stream-ordered like nsfem_sp_r_mult: no host synchronisation
        launch_spmv(*sim.dev, m, (const double *)x->impl.mem.ptr,
                    lhs ? (const double *)lhs->impl.mem.ptr : nullptr, (double *)y->impl.mem.ptr,
                    lhs ? 1 : 0);
    });
}

int nsfem_simulation_set_fields(nsfem_simulation *s, const double *u, const double *v,
                                const double *p) {
    return guard([&] {
        NSFEM_REQUIRE(s && s->impl && s->impl->dev, NSFEM_INVALID_ARGUMENT,
                      "null simulation or destroyed device");
        simulation_set(*s->impl, u, v, p);
    });
}

}  // extern "C"
