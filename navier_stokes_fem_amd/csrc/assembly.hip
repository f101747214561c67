// Device FEM assembly of the five matrices of
// NavierStokesAssembly::assembleAllMatrices (include/assembly.h:1302-1475)
// and of the constant Dirichlet corrections, in FP64 on the GPU.
//
// The reference builds triplets element by element and lets its sparse
// library sum the duplicates (mass :1321-1336, diffusion with
// assembleBCMatrix :994-1032 and :1346-1357, pressure stiffness :1393-1420,
// divergence assembleDivergenceMatrix :1627-1662 scaled at :1449, :1461).
// Here every matrix row is one thread:
//   1. node -> (element, local index) adjacency: counted and filled with
//      atomics, then each node's list sorted (ascending element order);
//   2. pass 1: each row's columns -- the sorted distinct nodes of its
//      adjacent elements, taken from a per-thread candidate list in LDS by
//      repeated minimum selection -- counted; an exclusive scan gives the
//      row starts;
//   3. pass 2: columns written and every nonzero summed over the row's
//      element contributions in ascending element order, which is the order
//      in which the reference inserts the triplets of that entry.
// Local matrices come from the same tables and geometry code as the host
// assembly (fem_tables.hpp) and this translation unit is built without FMA
// contraction, so every value is bit-identical to the host assembly's and
// to the oracle's.
//
// Velocity nodes may be renumbered on the way (vpos: original -> new index,
// vorder: new -> original): rows and columns of the velocity side come out
// in the new numbering with ascending columns, exactly what permuting the
// assembled matrices would give, so a simulation never assembles on the host.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <unordered_map>
#include <vector>

#include "fem_tables.hpp"
#include "nsfem_internal.hpp"

namespace nsfem {

namespace {

constexpr int kAsmBlock = 128;
constexpr int kMaxAdj = 16;                // elements around one node
constexpr int kMaxCand = 6 * kMaxAdj;      // candidate columns of one row

struct DevTables {
    double mass2[36];
    double delsq2[144];
    double delsq1[36];
    double comb_vd[36];
    double comb_pd[36];
};

struct AsmArgs {
    int V, P, E;
    const float *xy;       // original numbering
    const int32_t *el;     // original numbering (vertices are pressure indices)
    const int32_t *adj_start, *adj;  // original node -> e * 8 + local
    const int32_t *vorder, *vpos;    // null: identity
    const unsigned char *vmask;      // velocity Dirichlet mask, new numbering
    const double *gu, *gv;           // summed listed values, new numbering
    const unsigned char *pmask;      // pressure Dirichlet mask
    const double *gp;
    const DevTables *T;
    double nu, dt;
    int *error;
    // pass 1
    int64_t *lenM, *lenD, *lenPD, *lenK, *lenBD;
    // pass 2
    const int64_t *rsM, *rsD, *rsPD, *rsK, *rsBD;
    int32_t *colM, *colD, *colPD, *colK, *colBD;
    double *valM, *valD, *valPD, *valK, *valBD;
    double *cu, *cv, *cp;
};

__device__ __forceinline__ int vel_id(const AsmArgs &A, int orig) {
    return A.vpos ? A.vpos[orig] : orig;
}

// ------------------------------------------------------------- adjacency
__global__ void k_adj_count(int64_t n, const int32_t *__restrict__ el, int32_t *cnt) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[el[i]], 1);
}

__global__ void k_adj_fill(int64_t n, const int32_t *__restrict__ el,
                           const int32_t *__restrict__ start, int32_t *cursor,
                           int32_t *__restrict__ adj) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int node = el[i];
        const int pos = start[node] + atomicAdd(&cursor[node], 1);
        adj[pos] = (int32_t)((i / 6) * 8 + i % 6);
    }
}

// each node's list in ascending (element, local) order: insertion sort of a
// handful of entries
__global__ void k_adj_sort(int n, const int32_t *__restrict__ start, int32_t *__restrict__ adj,
                           int *error) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int a0 = start[i], a1 = start[i + 1];
    if (a1 - a0 > kMaxAdj) atomicOr(error, 1);
    for (int a = a0 + 1; a < a1; ++a) {
        const int32_t v = adj[a];
        int b = a - 1;
        while (b >= a0 && adj[b] > v) {
            adj[b + 1] = adj[b];
            --b;
        }
        adj[b + 1] = v;
    }
}

// smallest candidate above `last` (INT_MAX if none); cand is this thread's
// strided LDS column
__device__ __forceinline__ int next_col(const int *cand, int C, int last) {
    int m = INT_MAX;
    for (int j = 0; j < C; ++j) {
        const int c = cand[j * kAsmBlock];
        if (c > last && c < m) m = c;
    }
    return m;
}

// ---------------------------------------------------------- velocity rows
// Rows of the velocity mass matrix M, the diffusion matrix D (Dirichlet rows
// and columns eliminated, the moved columns applied to the boundary values:
// c_u, c_v) and the pressure divergence rows r and r + V.
template <bool FILL>
__global__ __launch_bounds__(kAsmBlock) void k_velocity_rows(const AsmArgs A) {
    __shared__ int lds[kMaxCand * kAsmBlock];
    const int rn = blockIdx.x * kAsmBlock + threadIdx.x;
    if (rn >= A.V) return;
    int *cand = lds + threadIdx.x;
    const int o = A.vorder ? A.vorder[rn] : rn;
    const int a0 = A.adj_start[o], a1 = A.adj_start[o + 1];
    const int nadj = min(a1 - a0, kMaxAdj);
    const bool rowb = A.vmask[rn] != 0;
    const fem::Tables &T = *reinterpret_cast<const fem::Tables *>(A.T);

    // ---- velocity-side columns (6 nodes per adjacent element)
    const int C = 6 * nadj;
    for (int a = 0; a < nadj; ++a) {
        const int e = A.adj[a0 + a] >> 3;
        for (int k = 0; k < 6; ++k) cand[(6 * a + k) * kAsmBlock] = vel_id(A, A.el[6 * e + k]);
    }
    int nM = 0, nD = 0;
    int64_t pM = 0, pD = 0;
    if (FILL) {
        pM = A.rsM[rn];
        pD = A.rsD[rn];
        if (rowb) {
            A.colD[pD] = rn;
            A.valD[pD] = 1.0;
        }
    }
    for (int last = -1;;) {
        const int c = next_col(cand, C, last);
        if (c == INT_MAX) break;
        last = c;
        const bool colb = A.vmask[c] != 0;
        if (FILL) {
            double m = 0.0, dsum = 0.0;
            for (int j = 0; j < C; ++j) {
                if (cand[j * kAsmBlock] != c) continue;
                const int32_t w = A.adj[a0 + j / 6];
                const int e = w >> 3, kr = w & 7, kc = j % 6;
                const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
                const double mloc = T.mass2[kr][kc] * g.jabs;
                m += mloc;
                if (!rowb && !colb) {
                    const double dloc = mloc + fem::stiff(&T.delsq2[0][0], 6, g, kr, kc) * A.dt * A.nu;
                    dsum += dloc;
                }
            }
            A.colM[pM + nM] = c;
            A.valM[pM + nM] = m;
            if (!rowb && !colb) {
                A.colD[pD + nD] = c;
                A.valD[pD + nD] = dsum;
            }
        }
        ++nM;
        if (!rowb && !colb) ++nD;
    }
    if (rowb) nD = 1;
    if (FILL) {
        // boundary columns applied to the boundary values, in (element,
        // local) order
        double cu = 0.0, cv = 0.0;
        if (!rowb) {
            for (int j = 0; j < C; ++j) {
                const int c = cand[j * kAsmBlock];
                if (!A.vmask[c]) continue;
                const int32_t w = A.adj[a0 + j / 6];
                const int e = w >> 3, kr = w & 7, kc = j % 6;
                const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
                const double mloc = T.mass2[kr][kc] * g.jabs;
                const double dloc = mloc + fem::stiff(&T.delsq2[0][0], 6, g, kr, kc) * A.dt * A.nu;
                cu += dloc * A.gu[c];
                cv += dloc * A.gv[c];
            }
        }
        A.cu[rn] = cu;
        A.cv[rn] = cv;
    } else {
        A.lenM[rn] = nM;
        A.lenD[rn] = nD;
    }

    // ---- pressure divergence rows rn (x) and rn + V (y): the pressure
    // nodes (vertices) of the adjacent elements
    const int Cp = 3 * nadj;
    for (int a = 0; a < nadj; ++a) {
        const int e = A.adj[a0 + a] >> 3;
        for (int k = 0; k < 3; ++k) cand[(3 * a + k) * kAsmBlock] = A.el[6 * e + k];
    }
    int nP = 0;
    int64_t px = 0, py = 0;
    if (FILL) {
        px = A.rsPD[rn];
        py = A.rsPD[rn + A.V];
    }
    for (int last = -1;;) {
        const int c = next_col(cand, Cp, last);
        if (c == INT_MAX) break;
        last = c;
        if (FILL) {
            double sx = 0.0, sy = 0.0;
            for (int j = 0; j < Cp; ++j) {
                if (cand[j * kAsmBlock] != c) continue;
                const int32_t w = A.adj[a0 + j / 3];
                const int e = w >> 3, kr = w & 7, kc = j % 3;
                const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
                sx += g.B[0][0] * T.comb_pd[kr][kc] + g.B[0][1] * T.comb_pd[kr][kc + 3];
                sy += g.B[1][0] * T.comb_pd[kr][kc] + g.B[1][1] * T.comb_pd[kr][kc + 3];
            }
            A.colPD[px + nP] = c;
            A.valPD[px + nP] = sx * -A.dt;
            A.colPD[py + nP] = c;
            A.valPD[py + nP] = sy * -A.dt;
        }
        ++nP;
    }
    if (!FILL) A.lenPD[rn] = nP;
}

// ---------------------------------------------------------- pressure rows
// Rows of the pressure stiffness matrix K (Dirichlet eliminated, c_p) and of
// the velocity divergence matrix (columns: the velocity nodes of the
// adjacent elements, then the same + V).
template <bool FILL>
__global__ __launch_bounds__(kAsmBlock) void k_pressure_rows(const AsmArgs A) {
    __shared__ int lds[kMaxCand * kAsmBlock];
    const int r = blockIdx.x * kAsmBlock + threadIdx.x;
    if (r >= A.P) return;
    int *cand = lds + threadIdx.x;
    const int a0 = A.adj_start[r], a1 = A.adj_start[r + 1];
    const int nadj = min(a1 - a0, kMaxAdj);
    const bool rowb = A.pmask[r] != 0;
    const fem::Tables &T = *reinterpret_cast<const fem::Tables *>(A.T);

    // ---- K: pressure nodes of the adjacent elements
    const int Cp = 3 * nadj;
    for (int a = 0; a < nadj; ++a) {
        const int e = A.adj[a0 + a] >> 3;
        for (int k = 0; k < 3; ++k) cand[(3 * a + k) * kAsmBlock] = A.el[6 * e + k];
    }
    int nK = 0;
    int64_t pK = 0;
    if (FILL) {
        pK = A.rsK[r];
        if (rowb) {
            A.colK[pK] = r;
            A.valK[pK] = 1.0;
        }
    }
    if (!rowb) {
        for (int last = -1;;) {
            const int c = next_col(cand, Cp, last);
            if (c == INT_MAX) break;
            last = c;
            if (A.pmask[c]) continue;
            if (FILL) {
                double s = 0.0;
                for (int j = 0; j < Cp; ++j) {
                    if (cand[j * kAsmBlock] != c) continue;
                    const int32_t w = A.adj[a0 + j / 3];
                    const int e = w >> 3, kr = w & 7, kc = j % 3;
                    const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
                    s += fem::stiff(&T.delsq1[0][0], 3, g, kr, kc);
                }
                A.colK[pK + nK] = c;
                A.valK[pK + nK] = s;
            }
            ++nK;
        }
    } else {
        nK = 1;
    }
    if (FILL) {
        double cp = 0.0;
        if (!rowb) {
            for (int j = 0; j < Cp; ++j) {
                const int c = cand[j * kAsmBlock];
                if (!A.pmask[c]) continue;
                const int32_t w = A.adj[a0 + j / 3];
                const int e = w >> 3, kr = w & 7, kc = j % 3;
                const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
                cp += fem::stiff(&T.delsq1[0][0], 3, g, kr, kc) * A.gp[c];
            }
        }
        A.cp[r] = cp;
    } else {
        A.lenK[r] = nK;
    }

    // ---- velocity divergence row r: velocity nodes (new numbering)
    const int C = 6 * nadj;
    for (int a = 0; a < nadj; ++a) {
        const int e = A.adj[a0 + a] >> 3;
        for (int k = 0; k < 6; ++k) cand[(6 * a + k) * kAsmBlock] = vel_id(A, A.el[6 * e + k]);
    }
    int n = 0;
    for (int last = -1;;) {
        const int c = next_col(cand, C, last);
        if (c == INT_MAX) break;
        last = c;
        ++n;
    }
    if (!FILL) {
        A.lenBD[r] = 2 * n;
        return;
    }
    const int64_t pB = A.rsBD[r];
    const double dtinv = 1.0 / A.dt;
    int i = 0;
    for (int last = -1;;) {
        const int c = next_col(cand, C, last);
        if (c == INT_MAX) break;
        last = c;
        double sx = 0.0, sy = 0.0;
        for (int j = 0; j < C; ++j) {
            if (cand[j * kAsmBlock] != c) continue;
            const int32_t w = A.adj[a0 + j / 6];
            const int e = w >> 3, kr = w & 7, kc = j % 6;
            const fem::Geo g = fem::geometry(A.xy, A.el + 6 * e);
            sx += g.B[0][0] * T.comb_vd[kr][kc] + g.B[0][1] * T.comb_vd[kr][kc + 6];
            sy += g.B[1][0] * T.comb_vd[kr][kc] + g.B[1][1] * T.comb_vd[kr][kc + 6];
        }
        A.colBD[pB + i] = c;
        A.valBD[pB + i] = sx * -dtinv;
        A.colBD[pB + n + i] = c + A.V;
        A.valBD[pB + n + i] = sy * -dtinv;
        ++i;
    }
}

__global__ void k_fill_bytes(int64_t n, unsigned char v, unsigned char *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = v;
}

__global__ void k_scatter_bc(int count, const int32_t *__restrict__ idx,
                             const double *__restrict__ a, const double *__restrict__ b,
                             unsigned char *__restrict__ mask, double *__restrict__ da,
                             double *__restrict__ db) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int node = idx[i];
    mask[node] = 1;
    da[node] = a[i];
    if (db) db[node] = b[i];
}

__global__ void k_narrow_starts(int64_t n, const int64_t *__restrict__ a, int32_t *__restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (int32_t)a[i];
}

inline int blocks_for(int64_t n, int bs) {
    return (int)std::max<int64_t>(1, std::min<int64_t>((n + bs - 1) / bs, 1 << 20));
}

template <typename T>
void upload_vec(Device &d, DeviceArray<T> &dst, const T *src, size_t n) {
    dst.resize(std::max<size_t>(n, 1));
    if (n)
        NSFEM_HIP(hipMemcpyAsync(dst.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, d.stream));
}

// exclusive scan of rows lengths (len[rows] = 0) into 64-bit starts
void scan_rows(Device &d, int64_t *len, int64_t *starts, int rows, DeviceArray<unsigned char> &tmp) {
    size_t bytes = 0;
    NSFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, len, starts, rows + 1, d.stream));
    tmp.ensure(std::max<size_t>(bytes, 1));
    NSFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, len, starts, rows + 1, d.stream));
}

}  // namespace

void assemble_device(Device &d, int V, int P, const float *xy, int E, const int32_t *el, int nvb,
                     const int32_t *vb, const double *gu, const double *gv, int npb,
                     const int32_t *pb, const double *gp, double nu, double dt,
                     const int32_t *vorder, const int32_t *vpos, DeviceAssembly &out) {
    NSFEM_REQUIRE(V > 0 && P > 0 && P <= V && E > 0, NSFEM_INVALID_ARGUMENT,
                  "assembly: bad mesh sizes");
    NSFEM_REQUIRE(dt > 0, NSFEM_INVALID_ARGUMENT, "assembly: dt must be positive");
    NSFEM_REQUIRE(6ll * E < (int64_t)INT32_MAX / 2, NSFEM_INVALID_ARGUMENT,
                  "assembly: too many elements");
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
    d.set_current();
    // boundary values summed per node in listing order (a node listed twice
    // contributes twice, as in the reference's loop), then scattered into
    // dense per-node arrays on the device (velocity side in new numbering)
    std::vector<int32_t> bn, pn;
    std::vector<double> bu, bv, bp;
    {
        std::unordered_map<int32_t, size_t> at;
        for (int k = 0; k < nvb; ++k) {
            const int32_t node = vpos ? vpos[vb[k]] : vb[k];
            auto it = at.find(node);
            if (it == at.end()) {
                at[node] = bn.size();
                bn.push_back(node);
                bu.push_back(0.0 + gu[k]);
                bv.push_back(0.0 + gv[k]);
            } else {
                bu[it->second] += gu[k];
                bv[it->second] += gv[k];
            }
        }
        at.clear();
        for (int k = 0; k < npb; ++k) {
            auto it = at.find(pb[k]);
            if (it == at.end()) {
                at[pb[k]] = pn.size();
                pn.push_back(pb[k]);
                bp.push_back(0.0 + gp[k]);
            } else {
                bp[it->second] += gp[k];
            }
        }
    }
    DeviceArray<float> dxy;
    DeviceArray<int32_t> del, dvorder, dvpos, adj_start, adj, cursor, dbn, dpn;
    DeviceArray<unsigned char> vmask, pmask, tmp;
    DeviceArray<double> dgu, dgv, dgp, dbu, dbv, dbp;
    DeviceArray<int> err;
    DeviceArray<DevTables> tables;
    upload_vec(d, dxy, xy, 2 * (size_t)V);
    upload_vec(d, del, el, 6 * (size_t)E);
    if (vorder) {
        upload_vec(d, dvorder, vorder, (size_t)V);
        upload_vec(d, dvpos, vpos, (size_t)V);
    }
    {
        static const fem::Tables T;
        static_assert(sizeof(DevTables) == sizeof(fem::Tables), "table layout");
        upload_vec(d, tables, reinterpret_cast<const DevTables *>(&T), 1);
    }
    vmask.resize(V);
    pmask.resize(P);
    dgu.resize(V);
    dgv.resize(V);
    dgp.resize(P);
    err.resize(1);
    NSFEM_HIP(hipMemsetAsync(err.ptr, 0, sizeof(int), d.stream));
    k_fill_bytes<<<blocks_for(V, 256), 256, 0, d.stream>>>(V, 0, vmask.ptr);
    k_fill_bytes<<<blocks_for(P, 256), 256, 0, d.stream>>>(P, 0, pmask.ptr);
    launch_fill(d, V, 0.0, dgu.ptr);
    launch_fill(d, V, 0.0, dgv.ptr);
    launch_fill(d, P, 0.0, dgp.ptr);
    upload_vec(d, dbn, bn.data(), bn.size());
    upload_vec(d, dbu, bu.data(), bu.size());
    upload_vec(d, dbv, bv.data(), bv.size());
    upload_vec(d, dpn, pn.data(), pn.size());
    upload_vec(d, dbp, bp.data(), bp.size());
    if (!bn.empty())
        k_scatter_bc<<<blocks_for((int64_t)bn.size(), 256), 256, 0, d.stream>>>(
            (int)bn.size(), dbn.ptr, dbu.ptr, dbv.ptr, vmask.ptr, dgu.ptr, dgv.ptr);
    if (!pn.empty())
        k_scatter_bc<<<blocks_for((int64_t)pn.size(), 256), 256, 0, d.stream>>>(
            (int)pn.size(), dpn.ptr, dbp.ptr, nullptr, pmask.ptr, dgp.ptr, nullptr);
    NSFEM_HIP(hipGetLastError());

    // ---- adjacency
    const int64_t n6 = 6 * (int64_t)E;
    DeviceArray<int32_t> cnt;
    cnt.resize((size_t)V + 1);
    NSFEM_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(int32_t) * ((size_t)V + 1), d.stream));
    k_adj_count<<<blocks_for(n6, 256), 256, 0, d.stream>>>(n6, del.ptr, cnt.ptr);
    adj_start.resize((size_t)V + 1);
    {
        size_t bytes = 0;
        NSFEM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, cnt.ptr, adj_start.ptr, V + 1,
                                                   d.stream));
        tmp.ensure(std::max<size_t>(bytes, 1));
        NSFEM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.ptr, bytes, cnt.ptr, adj_start.ptr, V + 1,
                                                   d.stream));
    }
    adj.resize((size_t)n6);
    NSFEM_HIP(hipMemsetAsync(cnt.ptr, 0, sizeof(int32_t) * ((size_t)V + 1), d.stream));
    k_adj_fill<<<blocks_for(n6, 256), 256, 0, d.stream>>>(n6, del.ptr, adj_start.ptr, cnt.ptr,
                                                          adj.ptr);
    k_adj_sort<<<blocks_for(V, 256), 256, 0, d.stream>>>(V, adj_start.ptr, adj.ptr, err.ptr);
    NSFEM_HIP(hipGetLastError());
    cnt.release();

    // ---- pass 1: row lengths
    DeviceArray<int64_t> lenM, lenD, lenPD, lenK, lenBD;
    lenM.resize((size_t)V + 1);
    lenD.resize((size_t)V + 1);
    lenPD.resize(2 * (size_t)V + 1);
    lenK.resize((size_t)P + 1);
    lenBD.resize((size_t)P + 1);
    AsmArgs A{};
    A.V = V;
    A.P = P;
    A.E = E;
    A.xy = dxy.ptr;
    A.el = del.ptr;
    A.adj_start = adj_start.ptr;
    A.adj = adj.ptr;
    A.vorder = vorder ? dvorder.ptr : nullptr;
    A.vpos = vorder ? dvpos.ptr : nullptr;
    A.vmask = vmask.ptr;
    A.gu = dgu.ptr;
    A.gv = dgv.ptr;
    A.pmask = pmask.ptr;
    A.gp = dgp.ptr;
    A.T = tables.ptr;
    A.nu = nu;
    A.dt = dt;
    A.error = err.ptr;
    A.lenM = lenM.ptr;
    A.lenD = lenD.ptr;
    A.lenPD = lenPD.ptr;
    A.lenK = lenK.ptr;
    A.lenBD = lenBD.ptr;
    k_velocity_rows<false><<<(V + kAsmBlock - 1) / kAsmBlock, kAsmBlock, 0, d.stream>>>(A);
    k_pressure_rows<false><<<(P + kAsmBlock - 1) / kAsmBlock, kAsmBlock, 0, d.stream>>>(A);
    NSFEM_HIP(hipGetLastError());
    {
        int h_err = 0;
        NSFEM_HIP(hipMemcpyAsync(&h_err, err.ptr, sizeof(int), hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        NSFEM_REQUIRE(h_err == 0, NSFEM_INVALID_ARGUMENT,
                      "device assembly: a node belongs to more than " + std::to_string(kMaxAdj) +
                          " elements (use the host assembly)");
    }
    // the y rows of the pressure divergence repeat the x rows' lengths
    NSFEM_HIP(hipMemcpyAsync(lenPD.ptr + V, lenPD.ptr, sizeof(int64_t) * V,
                             hipMemcpyDeviceToDevice, d.stream));
    int64_t *lens[5] = {lenM.ptr, lenD.ptr, lenK.ptr, lenBD.ptr, lenPD.ptr};
    const int rows[5] = {V, V, P, P, 2 * V};
    const int cols[5] = {V, V, P, 2 * V, P};
    DeviceArray<int64_t> starts[5];
    int64_t nnz[5];
    for (int k = 0; k < 5; ++k) {
        NSFEM_HIP(hipMemsetAsync(lens[k] + rows[k], 0, sizeof(int64_t), d.stream));
        starts[k].resize((size_t)rows[k] + 1);
        scan_rows(d, lens[k], starts[k].ptr, rows[k], tmp);
        NSFEM_HIP(hipMemcpyAsync(&nnz[k], starts[k].ptr + rows[k], sizeof(int64_t),
                                 hipMemcpyDeviceToHost, d.stream));
    }
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    for (int k = 0; k < 5; ++k)
        NSFEM_REQUIRE(nnz[k] < (int64_t)INT32_MAX, NSFEM_INVALID_ARGUMENT,
                      "assembly: matrix exceeds 2^31 nonzeros");
    lenM.release();
    lenD.release();
    lenPD.release();
    lenK.release();
    lenBD.release();

    // ---- pass 2: columns and values (outputs in the slot order of
    // DeviceAssembly: M, D, K, velocity divergence, pressure divergence)
    for (int k = 0; k < 5; ++k) {
        DeviceCsr &c = out.m[k];
        c.rows = rows[k];
        c.cols = cols[k];
        c.nnz = nnz[k];
        c.col.resize((size_t)std::max<int64_t>(nnz[k], 1));
        c.val.resize((size_t)std::max<int64_t>(nnz[k], 1));
        c.rs.resize((size_t)rows[k] + 1);
        k_narrow_starts<<<blocks_for(rows[k] + 1, 256), 256, 0, d.stream>>>(rows[k] + 1,
                                                                            starts[k].ptr, c.rs.ptr);
        if (nnz[k] == 0) {  // entry 0 of an empty matrix stays a valid column
            NSFEM_HIP(hipMemsetAsync(c.col.ptr, 0, sizeof(int32_t), d.stream));
            NSFEM_HIP(hipMemsetAsync(c.val.ptr, 0, sizeof(double), d.stream));
        }
    }
    out.c_u.resize(V);
    out.c_v.resize(V);
    out.c_p.resize(P);
    A.rsM = starts[0].ptr;
    A.rsD = starts[1].ptr;
    A.rsK = starts[2].ptr;
    A.rsBD = starts[3].ptr;
    A.rsPD = starts[4].ptr;
    A.colM = out.m[0].col.ptr;
    A.valM = out.m[0].val.ptr;
    A.colD = out.m[1].col.ptr;
    A.valD = out.m[1].val.ptr;
    A.colK = out.m[2].col.ptr;
    A.valK = out.m[2].val.ptr;
    A.colBD = out.m[3].col.ptr;
    A.valBD = out.m[3].val.ptr;
    A.colPD = out.m[4].col.ptr;
    A.valPD = out.m[4].val.ptr;
    A.cu = out.c_u.ptr;
    A.cv = out.c_v.ptr;
    A.cp = out.c_p.ptr;
    k_velocity_rows<true><<<(V + kAsmBlock - 1) / kAsmBlock, kAsmBlock, 0, d.stream>>>(A);
    k_pressure_rows<true><<<(P + kAsmBlock - 1) / kAsmBlock, kAsmBlock, 0, d.stream>>>(A);
    NSFEM_HIP(hipGetLastError());
    NSFEM_HIP(hipStreamSynchronize(d.stream));
}

}  // namespace nsfem
