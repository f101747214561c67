// Host FEM assembly of the five device matrices of
// NavierStokesAssembly::assembleAllMatrices (include/assembly.h:1302-1475),
// re-typed to FP64 (node coordinates stay float, as in the mesh format).
//
// The reference builds triplets element by element and lets the sparse
// library sum duplicates.  Here every row is assembled independently from a
// node -> (element, local index) adjacency, in ascending element order, so
// each nonzero receives its element contributions in the same order as the
// reference's triplet insertion order; rows are split over host threads.
//
// Local matrices follow the reference functors exactly:
//   mass      LocalMassFunctor        (:842-882) shapeSquared * |J|
//   stiffness LocalStiffnessFunctor   (:725-791) delPSq with B / |J|
//   diffusion mass + stiffness*dt*viscosity  (:1346-1357)
//   divergence LocalDivergenceFunctor (:793-840), scaled by -1/dt (velocity
//              divergence) and -dt (pressure divergence) (:1449,1461)
// with the 8-point triangle rule of TriangleIntegrator (:259-304).
// Dirichlet elimination is assembleBCMatrix (:994-1032): boundary rows keep
// only a unit diagonal, boundary columns move to the right-hand side.  The
// moved columns are applied to the boundary values once, here, giving the
// constant corrections c the reference subtracts every step (:503-524,
// :583-603); a node listed in several boundaries contributes once per
// listing, as in the reference's loop.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "fem_tables.hpp"
#include "nsfem_internal.hpp"

namespace nsfem {

namespace {
using fem::Geo;
using fem::Tables;
using fem::geometry;
using fem::stiff;

template <typename F>
void parallel_rows(int rows, int threads, F &&f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::max(1, std::min(threads, rows / 4096 + 1));
    if (threads == 1) {
        f(0, rows);
        return;
    }
    std::vector<std::thread> pool;
    const int chunk = (rows + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int lo = t * chunk, hi = std::min(rows, lo + chunk);
        if (lo >= hi) break;
        pool.emplace_back([&, lo, hi] { f(lo, hi); });
    }
    for (auto &th : pool) th.join();
}

// Sorted unique columns of one row into `out`; returns count.
inline int unique_sorted(int32_t *buf, int n) {
    std::sort(buf, buf + n);
    return (int)(std::unique(buf, buf + n) - buf);
}

inline int find_col(const int32_t *cols, int n, int32_t c) {
    return (int)(std::lower_bound(cols, cols + n, c) - cols);
}

}  // namespace

void assemble_host(int V, int P, const float *xy, int E, const int32_t *el, int nvb,
                   const int32_t *vb, const double *gu, const double *gv, int npb,
                   const int32_t *pb, const double *gp, double nu, double dt, int threads,
                   HostAssembly &out) {
    NSFEM_REQUIRE(V > 0 && P > 0 && P <= V && E > 0, NSFEM_INVALID_ARGUMENT,
                  "assembly: bad mesh sizes");
    NSFEM_REQUIRE(dt > 0, NSFEM_INVALID_ARGUMENT, "assembly: dt must be positive");
    for (int64_t k = 0; k < 6ll * E; ++k)
        NSFEM_REQUIRE(el[k] >= 0 && el[k] < V, NSFEM_INVALID_ARGUMENT,
                      "assembly: element node index out of range");
    for (int64_t e = 0; e < E; ++e)
        for (int k = 0; k < 6; ++k)
            NSFEM_REQUIRE(k < 3 ? el[6 * e + k] < P : el[6 * e + k] >= P, NSFEM_INVALID_ARGUMENT,
                          "assembly: vertices must be nodes 0..P-1 and midpoints P..V-1");
    for (int k = 0; k < nvb; ++k)
        NSFEM_REQUIRE(vb[k] >= 0 && vb[k] < V, NSFEM_INVALID_ARGUMENT,
                      "assembly: velocity Dirichlet node out of range");
    for (int k = 0; k < npb; ++k)
        NSFEM_REQUIRE(pb[k] >= 0 && pb[k] < P, NSFEM_INVALID_ARGUMENT,
                      "assembly: pressure Dirichlet node out of range");
    static const Tables T;
    // boundary masks and per-node sums of listed values
    std::vector<char> vmask(V, 0), pmask(P, 0);
    std::vector<double> Gu(V, 0.0), Gv(V, 0.0), Gp(P, 0.0);
    for (int k = 0; k < nvb; ++k) {
        vmask[vb[k]] = 1;
        Gu[vb[k]] += gu[k];
        Gv[vb[k]] += gv[k];
    }
    for (int k = 0; k < npb; ++k) {
        pmask[pb[k]] = 1;
        Gp[pb[k]] += gp[k];
    }
    // node -> (element, local) adjacency, ascending element order
    std::vector<int64_t> adj_start((size_t)V + 1, 0);
    for (int64_t k = 0; k < 6ll * E; ++k) adj_start[el[k] + 1]++;
    for (int i = 0; i < V; ++i) adj_start[i + 1] += adj_start[i];
    std::vector<int32_t> adj((size_t)adj_start[V]);
    {
        std::vector<int64_t> pos(adj_start.begin(), adj_start.end() - 1);
        for (int e = 0; e < E; ++e)
            for (int k = 0; k < 6; ++k) adj[pos[el[6 * e + k]]++] = e * 8 + k;
    }
    out.c_u.assign(V, 0.0);
    out.c_v.assign(V, 0.0);
    out.c_p.assign(P, 0.0);

    // ---- pass 1: row lengths
    HostCsr &M = out.m[NSFEM_ASM_VELOCITY_MASS], &D = out.m[NSFEM_ASM_DIFFUSION],
            &K = out.m[NSFEM_ASM_PRESSURE_STIFFNESS], &BD = out.m[NSFEM_ASM_VELOCITY_DIVERGENCE],
            &PD = out.m[NSFEM_ASM_PRESSURE_DIVERGENCE];
    M.rows = M.cols = V;
    D.rows = D.cols = V;
    K.rows = K.cols = P;
    BD.rows = P;
    BD.cols = 2 * V;
    PD.rows = 2 * V;
    PD.cols = P;
    std::vector<int32_t> lenM(V), lenD(V), lenK(P), lenPD(V);
    parallel_rows(V, threads, [&](int lo, int hi) {
        std::vector<int32_t> buf;
        for (int r = lo; r < hi; ++r) {
            buf.clear();
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3;
                for (int k = 0; k < 6; ++k) buf.push_back(el[6 * e + k]);
            }
            const int n = unique_sorted(buf.data(), (int)buf.size());
            lenM[r] = n;
            if (vmask[r]) {
                lenD[r] = 1;
            } else {
                int m = 0;
                for (int i = 0; i < n; ++i) m += !vmask[buf[i]];
                lenD[r] = m;
            }
            // pressure divergence rows r and r+V: P1 vertices of adjacent elements
            buf.clear();
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3;
                for (int k = 0; k < 3; ++k) buf.push_back(el[6 * e + k]);
            }
            lenPD[r] = unique_sorted(buf.data(), (int)buf.size());
            if (r < P) {
                if (pmask[r]) {
                    lenK[r] = 1;
                } else {
                    int m = 0;
                    for (int i = 0; i < lenPD[r]; ++i) m += !pmask[buf[i]];
                    lenK[r] = m;
                }
            }
        }
    });
    auto prefix = [](HostCsr &C, const std::vector<int32_t> &len, int rows, int mult_rows) {
        C.rs.assign((size_t)rows + 1, 0);
        for (int r = 0; r < rows; ++r) C.rs[r + 1] = C.rs[r] + len[r % mult_rows];
    };
    prefix(M, lenM, V, V);
    prefix(D, lenD, V, V);
    prefix(K, lenK, P, P);
    prefix(PD, lenPD, 2 * V, V);
    BD.rs.assign((size_t)P + 1, 0);
    for (int r = 0; r < P; ++r) BD.rs[r + 1] = BD.rs[r] + 2 * lenM[r];
    for (HostCsr *C : {&M, &D, &K, &BD, &PD}) {
        NSFEM_REQUIRE(C->rs.back() < (int64_t)INT32_MAX, NSFEM_INVALID_ARGUMENT,
                      "assembly: matrix exceeds 2^31 nonzeros");
        C->col.resize(C->rs.back());
        C->val.assign(C->rs.back(), 0.0);
    }

    // ---- pass 2: columns and values
    const double dtinv = 1.0 / dt;
    parallel_rows(V, threads, [&](int lo, int hi) {
        std::vector<int32_t> buf;
        for (int r = lo; r < hi; ++r) {
            buf.clear();
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3;
                for (int k = 0; k < 6; ++k) buf.push_back(el[6 * e + k]);
            }
            const int n = unique_sorted(buf.data(), (int)buf.size());
            int32_t *mc = &M.col[M.rs[r]];
            double *mv = &M.val[M.rs[r]];
            std::copy(buf.begin(), buf.begin() + n, mc);
            int32_t *dc = &D.col[D.rs[r]];
            double *dv = &D.val[D.rs[r]];
            int nd = 0;
            if (vmask[r]) {
                dc[0] = r;
                dv[0] = 1.0;
                nd = 1;
            } else {
                for (int i = 0; i < n; ++i)
                    if (!vmask[buf[i]]) dc[nd++] = buf[i];
            }
            const bool rowb = vmask[r] != 0;
            double cu = 0.0, cv = 0.0;
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3, kr = adj[a] & 7;
                const int32_t *ee = el + 6 * e;
                const Geo g = geometry(xy, ee);
                for (int kc = 0; kc < 6; ++kc) {
                    const int c = ee[kc];
                    const double mloc = T.mass2[kr][kc] * g.jabs;
                    mv[find_col(mc, n, c)] += mloc;
                    if (!rowb) {
                        const double dloc = mloc + stiff(&T.delsq2[0][0], 6, g, kr, kc) * dt * nu;
                        if (!vmask[c]) {
                            dv[find_col(dc, nd, c)] += dloc;
                        } else {
                            cu += dloc * Gu[c];
                            cv += dloc * Gv[c];
                        }
                    }
                }
            }
            out.c_u[r] = cu;
            out.c_v[r] = cv;

            // pressure divergence rows r (x) and r+V (y)
            buf.clear();
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3;
                for (int k = 0; k < 3; ++k) buf.push_back(el[6 * e + k]);
            }
            const int np = unique_sorted(buf.data(), (int)buf.size());
            int32_t *pxc = &PD.col[PD.rs[r]];
            double *pxv = &PD.val[PD.rs[r]];
            int32_t *pyc = &PD.col[PD.rs[r + V]];
            double *pyv = &PD.val[PD.rs[r + V]];
            std::copy(buf.begin(), buf.begin() + np, pxc);
            std::copy(buf.begin(), buf.begin() + np, pyc);
            for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                const int e = adj[a] >> 3, kr = adj[a] & 7;
                const Geo g = geometry(xy, el + 6 * e);
                for (int kc = 0; kc < 3; ++kc) {
                    const int c = el[6 * e + kc];
                    const int pos = find_col(pxc, np, c);
                    pxv[pos] += g.B[0][0] * T.comb_pd[kr][kc] + g.B[0][1] * T.comb_pd[kr][kc + 3];
                    pyv[pos] += g.B[1][0] * T.comb_pd[kr][kc] + g.B[1][1] * T.comb_pd[kr][kc + 3];
                }
            }
            for (int i = 0; i < np; ++i) {
                pxv[i] *= -dt;
                pyv[i] *= -dt;
            }

            if (r < P) {
                // pressure stiffness (P1), Dirichlet eliminated
                int32_t *kc_ = &K.col[K.rs[r]];
                double *kv = &K.val[K.rs[r]];
                int nk = 0;
                if (pmask[r]) {
                    kc_[0] = r;
                    kv[0] = 1.0;
                    nk = 1;
                } else {
                    for (int i = 0; i < np; ++i)
                        if (!pmask[buf[i]]) kc_[nk++] = buf[i];
                }
                // velocity divergence row r: columns (P2 nodes) then (+V)
                int32_t *bc = &BD.col[BD.rs[r]];
                double *bv = &BD.val[BD.rs[r]];
                for (int i = 0; i < n; ++i) {
                    bc[i] = mc[i];
                    bc[n + i] = mc[i] + V;
                }
                double cp = 0.0;
                for (int64_t a = adj_start[r]; a < adj_start[r + 1]; ++a) {
                    const int e = adj[a] >> 3, kr = adj[a] & 7;
                    const int32_t *ee = el + 6 * e;
                    const Geo g = geometry(xy, ee);
                    if (!pmask[r]) {
                        for (int kc = 0; kc < 3; ++kc) {
                            const int c = ee[kc];
                            const double v = stiff(&T.delsq1[0][0], 3, g, kr, kc);
                            if (!pmask[c]) kv[find_col(kc_, nk, c)] += v;
                            else cp += v * Gp[c];
                        }
                    }
                    for (int kc = 0; kc < 6; ++kc) {
                        const int pos = find_col(mc, n, ee[kc]);
                        bv[pos] += g.B[0][0] * T.comb_vd[kr][kc] + g.B[0][1] * T.comb_vd[kr][kc + 6];
                        bv[n + pos] +=
                            g.B[1][0] * T.comb_vd[kr][kc] + g.B[1][1] * T.comb_vd[kr][kc + 6];
                    }
                }
                for (int i = 0; i < 2 * n; ++i) bv[i] *= -dtinv;
                out.c_p[r] = cp;
            }
        }
    });
}

}  // namespace nsfem
