// CSR products, BLAS-1 and the persistent conjugate-gradient kernel for
// gfx950.  Replaces gpu/matrix_math.cu of the reference (spRMult :19-51,
// spRMultSub :66-99, saxpy :106-122, saxpby :131-155, dotProduct :162-194,
// GridSync :209-249, conjugateGradientMegakernel :251-287) and the host
// orchestration of cpp/gpu_simulation_device.cpp:142-515.
//
// Layout in HBM: the reference CSR (int32 row_start[n+1], int32 col[nnz]
// ascending per row, values[nnz]) plus a tile table cutting the rows into
// runs of <= kTileRows rows and <= kTileNnz nonzeros.  A tile is staged
// through LDS: one coalesced pass over col/val computes every product
// val*x[col] into LDS, then one thread per row adds its products in row
// order.  That keeps the column and value streams fully coalesced for the
// 7 (P1) to 19 (P2) nonzeros per row of these matrices, and it adds the
// products of a row in the same order as the reference's row loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <type_traits>

#include "nsfem_internal.hpp"

namespace nsfem {

// ----------------------------------------------------------------- helpers
namespace {

constexpr int kWaves = kBlock / 64;
// The product loads a tile's columns and values in groups of 4 nonzeros
// starting at a multiple of 4 (16-byte column and 2 x 16-byte value loads
// per lane); tiles hold at most kTileNnz - 3 nonzeros so that the aligned
// span of any tile fits kTileNnz products, kGroups groups per thread.
constexpr int kSpmvTileNnz = kTileNnz - 3;
typedef short i16x4 __attribute__((ext_vector_type(4)));
constexpr int kGroups = kTileNnz / (4 * kBlock);
static_assert(kGroups * 4 * kBlock == kTileNnz, "tile must be whole groups per thread");
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Workgroups are dispatched round-robin over the 8 XCDs; this maps the
// hardware block id to a logical id so that each XCD (and its L2) works on
// one contiguous range of rows.
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
    constexpr int nx = 8;
    const int per = nb / nx, rem = nb % nx;
    const int xcd = b % nx, local = b / nx;
    return xcd < rem ? xcd * (per + 1) + local
                     : rem * (per + 1) + (xcd - rem) * per + local;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic block sum of K doubles; every thread receives the totals.
template <int K, int BS>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds) {
    constexpr int NW = BS / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * NW + w] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NW; ++i) s += lds[k * NW + i];
        v[k] = s;
    }
    __syncthreads();
}

// Matrix streams (column indices, values) are read once per product: load
// them non-temporally so they do not evict the vector entries the gathers
// reuse from L2.
#ifndef NSFEM_STREAM_LOADS
#define NSFEM_STREAM_LOADS 1
#endif
template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
#if NSFEM_STREAM_LOADS
    return __builtin_nontemporal_load(p);
#else
    return *p;
#endif
}

// 16-bit column offsets unless NSFEM_CG_IDX16=0 (comparison runs)
bool want_idx16() {
    static const bool on = [] {
        const char *e = std::getenv("NSFEM_CG_IDX16");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c);
template <>
__device__ __forceinline__ double fma_t<double>(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}

}  // namespace

// ------------------------------------------------------------ tile tables
namespace {
// Greedy cut of the rows into runs of <= max_rows rows and <= max_nnz
// nonzeros; a single longer row forms a run of its own.
std::vector<int32_t> cut_rows(const std::vector<int32_t> &rs, int rows, int max_rows,
                              int max_nnz) {
    std::vector<int32_t> t;
    t.reserve((size_t)rows / std::max(1, max_rows / 2) + 2);
    int r = 0;
    t.push_back(0);
    while (r < rows) {
        int end = r;
        int64_t nz = 0;
        while (end < rows && end - r < max_rows) {
            const int64_t rn = rs[end + 1] - rs[end];
            if (end > r && nz + rn > max_nnz) break;
            nz += rn;
            ++end;
            if (nz > max_nnz) break;
        }
        t.push_back(end);
        r = end;
    }
    return t;
}
}  // namespace

namespace {
// dcol[k] = col[k] - first row of k's wave tile, one wave per tile; bad[0]
// is raised when an offset does not fit 16 bits.
__global__ void k_col_delta(int n_wtiles, const int4 *__restrict__ meta,
                            const int32_t *__restrict__ col, int16_t *__restrict__ dcol,
                            int *bad) {
    const int t = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (t >= n_wtiles) return;
    const int4 m = meta[t];
    int out = 0;
    for (int k = threadIdx.x & 63; k < m.w; k += 64) {
        const int dv = col[m.z + k] - m.x;
        out |= (dv < -32768 || dv > 32767) ? 1 : 0;
        dcol[m.z + k] = (int16_t)dv;
    }
    if (out) __hip_atomic_fetch_or(bad, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Kernels stream 16-bit column offsets instead of 32-bit columns when every
// offset from its tile's first row fits (banded numberings: the offsets span
// about twice the mesh width).  Returns whether they do; `out` is released
// if not.
bool col_delta(const Matrix &m, int n_tiles, const int4 *meta, DeviceArray<int16_t> &out) {
    out.release();
    if (m.nnz <= 0 || m.rows != m.cols || n_tiles <= 0) return false;
    out.resize((size_t)m.nnz + 8);  // 8 zero offsets of padding for group loads
    NSFEM_HIP(hipMemset(out.ptr + m.nnz, 0, 8 * sizeof(int16_t)));
    DeviceArray<int> bad;
    bad.resize(1);
    NSFEM_HIP(hipMemset(bad.ptr, 0, sizeof(int)));
    constexpr int kWavesPerBlock = 4;
    k_col_delta<<<(n_tiles + kWavesPerBlock - 1) / kWavesPerBlock, 64 * kWavesPerBlock>>>(
        n_tiles, meta, m.col.ptr, out.ptr, bad.ptr);
    NSFEM_HIP(hipGetLastError());
    int h = 1;
    NSFEM_HIP(hipMemcpy(&h, bad.ptr, sizeof(int), hipMemcpyDeviceToHost));
    if (h != 0) out.release();
    return h == 0;
}

// wave tiles: the persistent CG's offsets
void build_col_delta(Matrix &m) { m.dcol_ok = col_delta(m, m.n_wtiles, m.wmeta.ptr, m.dcol); }

// Group loads (8 nonzeros per lane from a multiple of 8) need the 16-bit
// offsets (padded by 8) and every multi-row tile within kStageNnz nonzeros.
void set_group_loads(Matrix &m) {
    m.groups_ok = m.dcol_ok && m.nnz > 0 && m.max_multirow_nz <= kStageNnz;
}

}  // namespace

void build_tiles(Matrix &m, const std::vector<int32_t> &rs) {
    std::vector<int32_t> t = cut_rows(rs, m.rows, kTileRows, kSpmvTileNnz);
    m.n_tiles = (int)t.size() - 1;
    m.host_tiles = t;
    m.tiles.resize(t.size());
    NSFEM_HIP(hipMemcpy(m.tiles.ptr, t.data(), t.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    m.host_rs = rs;
    {
        std::vector<int4> meta((size_t)std::max(m.n_tiles, 1));
        for (int i = 0; i < m.n_tiles; ++i)
            meta[i] = make_int4(t[i], t[i + 1], rs[t[i]], rs[t[i + 1]] - rs[t[i]]);
        DeviceArray<int4> dmeta;
        dmeta.resize(meta.size());
        NSFEM_HIP(hipMemcpy(dmeta.ptr, meta.data(), meta.size() * sizeof(int4),
                            hipMemcpyHostToDevice));
        m.bcol_ok = col_delta(m, m.n_tiles, dmeta.ptr, m.bcol);
    }
    build_wave_tiles(m, kWaveRows, kStageNnz);
}

void build_wave_tiles(Matrix &m, int rows, int nnz) {
    NSFEM_REQUIRE(rows >= 1 && rows <= kWaveRows && nnz >= 64 && nnz <= kWaveNnz,
                  NSFEM_INVALID_ARGUMENT, "wave tile shape out of range");
    const std::vector<int32_t> &rs = m.host_rs;
    m.wave_rows = rows;
    m.wave_nnz = nnz;
    std::vector<int32_t> w = cut_rows(rs, m.rows, rows, nnz);
    m.n_wtiles = (int)w.size() - 1;
    m.host_wtile_prefix.assign(w.size(), 0);
    for (int i = 0; i < m.n_wtiles; ++i)
        m.host_wtile_prefix[i + 1] =
            m.host_wtile_prefix[i] + (rs[w[i + 1]] - rs[w[i]]) + (w[i + 1] - w[i]);
    m.wtiles.resize(w.size());
    NSFEM_HIP(hipMemcpy(m.wtiles.ptr, w.data(), w.size() * sizeof(int32_t),
                        hipMemcpyHostToDevice));
    std::vector<int4> meta((size_t)std::max(m.n_wtiles, 1));
    m.max_multirow_nz = 0;
    for (int i = 0; i < m.n_wtiles; ++i) {
        meta[i] = make_int4(w[i], w[i + 1], rs[w[i]], rs[w[i + 1]] - rs[w[i]]);
        if (w[i + 1] - w[i] > 1) m.max_multirow_nz = std::max(m.max_multirow_nz, meta[i].w);
    }
    m.wmeta.resize(meta.size());
    NSFEM_HIP(hipMemcpy(m.wmeta.ptr, meta.data(), meta.size() * sizeof(int4),
                        hipMemcpyHostToDevice));
    m.block_tiles.resize(kMaxGridBlocks + 1);
    m.host_block_tiles.reserve(kMaxGridBlocks + 1);
    m.block_tiles_grid = 0;
    build_col_delta(m);
    set_group_loads(m);
}

// Persistent blocks own contiguous wave-tile ranges balanced by (nnz + rows).
// The table lives in arrays sized for the largest grid at tiling time, so
// this only computes and copies.
void ensure_block_tiles(Device &d, Matrix &m, int grid) {
    NSFEM_REQUIRE(grid >= 1 && grid <= kMaxGridBlocks, NSFEM_ERROR, "grid size out of range");
    if (m.block_tiles_grid == grid) return;
    std::vector<int32_t> &bt = m.host_block_tiles;
    bt.resize((size_t)grid + 1);  // within the reserved capacity
    const int64_t total = m.host_wtile_prefix[m.n_wtiles];
    int t = 0;
    for (int b = 0; b <= grid; ++b) {
        const int64_t target = total * b / grid;
        while (t < m.n_wtiles && m.host_wtile_prefix[t] < target) ++t;
        bt[b] = t;
    }
    bt[0] = 0;
    bt[grid] = m.n_wtiles;
    NSFEM_HIP(hipMemcpyAsync(m.block_tiles.ptr, bt.data(), bt.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    m.block_tiles_grid = grid;
}

namespace {
__global__ void k_to_float(int64_t n, const double *__restrict__ a, float *__restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
}  // namespace

// The FP32 copy of the values, with 8 zero entries of padding (the group-load
// product loads 8 values per lane from a multiple of 8).
void ensure_val32(Device &d, Matrix &m) {
    if (m.val32.ptr && m.val32.count == (size_t)m.nnz + 8) return;
    m.val32.resize((size_t)m.nnz + 8);
    NSFEM_HIP(hipMemsetAsync(m.val32.ptr + m.nnz, 0, 8 * sizeof(float), d.stream));
    if (m.nnz) {
        const int grid = (int)std::min<int64_t>((m.nnz + 255) / 256, 65536);
        k_to_float<<<grid, 256, 0, d.stream>>>(m.nnz, m.val.ptr, m.val32.ptr);
        NSFEM_HIP(hipGetLastError());
    }
}

// --------------------------------------------------------- Jacobi scaling
namespace {

// scale[r] = 1 / sqrt(a_rr) (the row's diagonal entries summed); a row
// without a positive diagonal sets *bad (and keeps scale 1)
__global__ void k_jacobi_scale(int n, const int32_t *__restrict__ rs,
                               const int32_t *__restrict__ col, const double *__restrict__ val,
                               double *__restrict__ scale, int32_t *__restrict__ bad) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double d = 0.0;
    for (int32_t k = rs[r]; k < rs[r + 1]; ++k)
        if (col[k] == r) d += val[k];
    if (d > 0.0) {
        scale[r] = 1.0 / sqrt(d);
    } else {
        scale[r] = 1.0;
        *bad = 1;
    }
}

// out[k] = (float)(scale[row] * a_k * scale[col_k]): D^-1/2 A D^-1/2 in FP32
__global__ void k_jacobi_values(int n, const int32_t *__restrict__ rs,
                                const int32_t *__restrict__ col, const double *__restrict__ val,
                                const double *__restrict__ scale, float *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const double sr = scale[r];
    for (int32_t k = rs[r]; k < rs[r + 1]; ++k) out[k] = (float)(sr * val[k] * scale[col[k]]);
}
}  // namespace

// The Jacobi-scaled FP32 values and the row scales of a square matrix
// (built once; 8 zero entries of padding as val32).  False when a row has
// no positive diagonal entry: the solve then runs unscaled.
bool ensure_jacobi(Device &d, Matrix &m) {
    if (m.jacobi_built) return m.jacobi_ok;
    m.jacobi_built = true;
    m.jacobi_ok = false;
    if (m.rows != m.cols || m.rows == 0) return false;
    m.jscale.resize((size_t)m.rows);
    DeviceArray<int32_t> bad;
    bad.resize(1);
    NSFEM_HIP(hipMemsetAsync(bad.ptr, 0, sizeof(int32_t), d.stream));
    const int grid = (m.rows + 255) / 256;
    k_jacobi_scale<<<grid, 256, 0, d.stream>>>(m.rows, m.row_start.ptr, m.col.ptr, m.val.ptr,
                                               m.jscale.ptr, bad.ptr);
    NSFEM_HIP(hipGetLastError());
    int32_t h = 0;
    NSFEM_HIP(hipMemcpyAsync(&h, bad.ptr, sizeof(h), hipMemcpyDeviceToHost, d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    if (h) {
        m.jscale.release();
        return false;
    }
    m.jval32.resize((size_t)m.nnz + 8);
    NSFEM_HIP(hipMemsetAsync(m.jval32.ptr + m.nnz, 0, 8 * sizeof(float), d.stream));
    k_jacobi_values<<<grid, 256, 0, d.stream>>>(m.rows, m.row_start.ptr, m.col.ptr, m.val.ptr,
                                                m.jscale.ptr, m.jval32.ptr);
    NSFEM_HIP(hipGetLastError());
    m.jacobi_ok = true;
    return true;
}

// ------------------------------------------------------------- CSR product
namespace {

// MODE 0: y = A x.  MODE 1: y = lhs - A x (y may alias lhs).
// MODE 2: y = A x - lhs.  MODE 3: y = x + omega dinv (lhs - A x), a damped
// Jacobi sweep into another vector (y aliases neither x nor lhs; the same
// roundings as the residual product followed by k_jacobi).
// IT = int16_t reads the columns as 16-bit offsets from the tile's first row
// (bcol, square matrices whose offsets fit): 2 bytes less per nonzero.
// T = double: the FP64 values; T = float: the FP32 copy (val32) with FP32
// x.  S: the type of the products, sums, lhs and y: T (FP32 throughout for
// the multigrid smoothing of mixed-precision solves), or double with FP32
// values and x (each FP32 x FP32 product exact in FP64, summed in FP64: the
// product of a CG's FP32 direction, free of the cancellation an FP32 sum
// suffers on smooth vectors).
template <int MODE, typename IT, typename T, typename S = T>
__global__ __launch_bounds__(kBlock) void k_spmv(int n_tiles, int n_rows, int64_t nnz,
                                                 const int32_t *__restrict__ tiles,
                                                 const int32_t *__restrict__ rs,
                                                 const int32_t *__restrict__ col,
                                                 const int16_t *__restrict__ bcol,
                                                 const T *__restrict__ val,
                                                 const T *__restrict__ x,
                                                 const S *lhs, S *y,
                                                 const S *__restrict__ dinv, S omega) {
    constexpr bool F64 = std::is_same<T, double>::value;
    constexpr bool S64 = std::is_same<S, double>::value;
    __shared__ __align__(16) S prod[kTileNnz];
    __shared__ S red[kWaves];
    const int t = xcd_swizzle(blockIdx.x, gridDim.x);
    if (t >= n_tiles) return;
    const int tid = threadIdx.x;
    const int r0 = tiles[t], r1 = tiles[t + 1];
    const int k0 = rs[r0];
    const int nz = rs[r1] - k0;
    if (nz <= kSpmvTileNnz) {
        // groups of 4 nonzeros from ka = k0 rounded down to a multiple of 4;
        // entries of the neighbouring tiles in the first and last group are
        // multiplied too (their columns are valid) and never summed, entries
        // past the end of the matrix are column 0, value 0
        const int ka = k0 & ~3;
        const int ng = (k0 + nz - ka + 3) >> 2;
        i32x4 c[kGroups];
        T v[kGroups][4];
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int grp = tid + g * kBlock;
            c[g] = i32x4{0, 0, 0, 0};
#pragma unroll
            for (int e = 0; e < 4; ++e) v[g][e] = T(0);
            if (grp < ng) {
                const int64_t base = ka + 4 * (int64_t)grp;
                if (base + 4 <= nnz) {
                    if constexpr (std::is_same<IT, int16_t>::value) {
                        // entries of a neighbouring tile decode relative to
                        // the wrong row: clamped into the vector, never summed
                        const i16x4 d = ld_stream(reinterpret_cast<const i16x4 *>(bcol + base));
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            c[g][e] = min(max(r0 + (int)d[e], 0), n_rows - 1);
                    } else {
                        c[g] = ld_stream(reinterpret_cast<const i32x4 *>(col + base));
                    }
                    if constexpr (F64) {
                        const f64x2 a = ld_stream(reinterpret_cast<const f64x2 *>(val + base));
                        const f64x2 b = ld_stream(reinterpret_cast<const f64x2 *>(val + base + 2));
                        v[g][0] = a.x;
                        v[g][1] = a.y;
                        v[g][2] = b.x;
                        v[g][3] = b.y;
                    } else {
                        const f32x4_t a = ld_stream(reinterpret_cast<const f32x4_t *>(val + base));
                        v[g][0] = a.x;
                        v[g][1] = a.y;
                        v[g][2] = a.z;
                        v[g][3] = a.w;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const bool in = base + e < nnz;
                        c[g][e] = in ? col[base + e] : 0;
                        v[g][e] = in ? val[base + e] : T(0);
                    }
                }
            }
        }
        // columns transposed through the wave's own slice of prod (256
        // products; the 256 columns take 1 KB of it), so that each gather
        // instruction covers 64 consecutive nonzeros
        const int lane = tid & 63;
#pragma unroll
        for (int g = 0; g < kGroups; ++g) {
            const int chunk = 4 * (g * kBlock + (tid & ~63));  // first element of the wave's slice
            int *cs = reinterpret_cast<int *>(prod + chunk);
            if (chunk < 4 * ng) {  // wave-uniform
                *reinterpret_cast<i32x4 *>(cs + 4 * lane) = c[g];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                int cc[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) cc[j] = cs[lane + 64 * j];
                T xv[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) xv[j] = x[cc[j]];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int j = 0; j < 4; ++j) prod[chunk + lane + 64 * j] = (S)xv[j];
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
                __builtin_amdgcn_wave_barrier();
                if constexpr (S64) {
                    f64x2 p0 = *reinterpret_cast<const f64x2 *>(prod + chunk + 4 * lane);
                    f64x2 p1 = *reinterpret_cast<const f64x2 *>(prod + chunk + 4 * lane + 2);
                    p0.x = (double)v[g][0] * p0.x;
                    p0.y = (double)v[g][1] * p0.y;
                    p1.x = (double)v[g][2] * p1.x;
                    p1.y = (double)v[g][3] * p1.y;
                    *reinterpret_cast<f64x2 *>(prod + chunk + 4 * lane) = p0;
                    *reinterpret_cast<f64x2 *>(prod + chunk + 4 * lane + 2) = p1;
                } else {
                    f32x4_t p = *reinterpret_cast<const f32x4_t *>(prod + chunk + 4 * lane);
                    p.x = v[g][0] * p.x;
                    p.y = v[g][1] * p.y;
                    p.z = v[g][2] * p.z;
                    p.w = v[g][3] * p.w;
                    *reinterpret_cast<f32x4_t *>(prod + chunk + 4 * lane) = p;
                }
            }
        }
        __syncthreads();
        const int row = r0 + tid;
        if (row < r1) {
            const int a = rs[row] - ka, b = rs[row + 1] - ka;
            S s = S(0);
            for (int k = a; k < b; ++k) s += prod[k];
            if (MODE == 0) y[row] = s;
            if (MODE == 1) y[row] = lhs[row] - s;
            if (MODE == 2) y[row] = s - lhs[row];
            if (MODE == 3) {
                const S r = lhs[row] - s;
                y[row] = __builtin_fma(omega * dinv[row], r, (S)x[row]);
            }
        }
    } else {  // one row longer than a tile
        S s = S(0);
        for (int k = tid; k < nz; k += kBlock) s += (S)val[k0 + k] * (S)x[col[k0 + k]];
        s = wave_sum(s);
        if ((tid & 63) == 0) red[tid >> 6] = s;
        __syncthreads();
        if (tid == 0) {
            S tot = S(0);
            for (int i = 0; i < kWaves; ++i) tot += red[i];
            if (MODE == 0) y[r0] = tot;
            if (MODE == 1) y[r0] = lhs[r0] - tot;
            if (MODE == 2) y[r0] = tot - lhs[r0];
            if (MODE == 3) {
                const S r = lhs[r0] - tot;
                y[r0] = __builtin_fma(omega * dinv[r0], r, (S)x[r0]);
            }
        }
    }
}

__global__ void k_axpy(int n, double a, const double *__restrict__ x, double *y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = __builtin_fma(a, x[i], y[i]);
}

// y = y0 + a_0 X_0 + ... + a_{l-1} X_{l-1} (y0 = 0, or y's own values when
// `accumulate`), X_k = X + k * stride, one fused multiply-add per term in k
// order: the bits of a fill (or nothing) followed by l k_axpy launches, in
// one pass over the vectors.  y2, when given, receives the same values.
constexpr int kCombineMax = 8;
struct CombineCoef {
    double a[kCombineMax];
};

__global__ void k_combine(int n, int l, CombineCoef c, const double *__restrict__ X,
                          size_t stride, int accumulate, double *__restrict__ y,
                          double *__restrict__ y2) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double t = accumulate ? y[i] : 0.0;
        for (int k = 0; k < l; ++k) t = __builtin_fma(c.a[k], X[(size_t)k * stride + i], t);
        y[i] = t;
        if (y2) y2[i] = t;
    }
}

__global__ void k_axpby(int n, double a, double b, const double *x, const double *y,
                        double *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = __builtin_fma(a, x[i], b * y[i]);
}

// Deterministic dot: block k sums the contiguous chunk k, a single block adds
// the chunk sums in order.
__global__ __launch_bounds__(kBlock) void k_dot_partial(int n, int chunk,
                                                        const double *__restrict__ a,
                                                        const double *__restrict__ b,
                                                        double *__restrict__ partial) {
    __shared__ double lds[kWaves];
    const int lo = blockIdx.x * chunk;
    const int hi = min(n, lo + chunk);
    double s[1] = {0.0};
    for (int i = lo + threadIdx.x; i < hi; i += kBlock) s[0] = __builtin_fma(a[i], b[i], s[0]);
    block_sum<1, kBlock>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void k_dot_final(int nb, const double *__restrict__ partial,
                                                      double *result) {
    __shared__ double lds[kWaves];
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < nb; i += kBlock) s[0] += partial[i];
    block_sum<1, kBlock>(s, lds);
    if (threadIdx.x == 0) result[0] += s[0];
}

// The update of a CG with FP32 directions p (q = A p in FP64) and an FP32
// preconditioner: x += a p and r -= a q in FP64, r32 = r rounded to FP32
// (the preconditioner's input), and the partial sums of (r, r) per
// contiguous chunk in k_dot_partial's order (k_dot_final adds them).
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, int chunk, double a,
                                                       const float *__restrict__ p,
                                                       const double *__restrict__ q,
                                                       double *__restrict__ x,
                                                       double *__restrict__ r,
                                                       float *__restrict__ r32,
                                                       double *__restrict__ partial) {
    __shared__ double lds[kWaves];
    const int lo = blockIdx.x * chunk;
    const int hi = min(n, lo + chunk);
    double s[1] = {0.0};
    for (int i = lo + threadIdx.x; i < hi; i += kBlock) {
        x[i] = __builtin_fma(a, (double)p[i], x[i]);
        const double ri = __builtin_fma(-a, q[i], r[i]);
        r[i] = ri;
        r32[i] = (float)ri;
        s[0] = __builtin_fma(ri, ri, s[0]);
    }
    block_sum<1, kBlock>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Partial sums of (a, b), each operand FP64 or FP32 (widened), per chunk,
// in k_dot_partial's order.
template <typename A, typename B>
__global__ __launch_bounds__(kBlock) void k_dot_any_partial(int n, int chunk,
                                                            const A *__restrict__ a,
                                                            const B *__restrict__ b,
                                                            double *__restrict__ partial) {
    __shared__ double lds[kWaves];
    const int lo = blockIdx.x * chunk;
    const int hi = min(n, lo + chunk);
    double s[1] = {0.0};
    for (int i = lo + threadIdx.x; i < hi; i += kBlock)
        s[0] = __builtin_fma((double)a[i], (double)b[i], s[0]);
    block_sum<1, kBlock>(s, lds);
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// p = z + b p in FP32 storage (the CG's new direction; the sum in FP64)
__global__ void k_pcg_direction(int n, const float *__restrict__ z, double b,
                                float *__restrict__ p) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (float)__builtin_fma(b, (double)p[i], (double)z[i]);
}

__global__ void k_scatter(int count, const int32_t *__restrict__ idx,
                          const double *__restrict__ val, double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[idx[i]] = val[i];
}

template <typename T>
__global__ void k_scatter_value(int count, const int32_t *__restrict__ idx, T v,
                                T *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[idx[i]] = v;
}

template <typename T>
__global__ void k_add(int n, const T *__restrict__ a, T *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a[i] + y[i];
}

// x += w * dinv .* r (a damped Jacobi update from the residual r)
template <typename T>
__global__ void k_jacobi(int n, T w, const T *__restrict__ dinv, const T *__restrict__ r,
                         T *__restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = __builtin_fma(w * dinv[i], r[i], x[i]);
}

// x = w * dinv .* b (the damped Jacobi sweep from x = 0)
template <typename T>
__global__ void k_jacobi0(int n, T w, const T *__restrict__ dinv, const T *__restrict__ b,
                          T *__restrict__ x) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (w * dinv[i]) * b[i];
}

// y = a rounded to (or widened from) y's type
template <typename A, typename Y>
__global__ void k_convert(int n, const A *__restrict__ a, Y *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (Y)a[i];
}

template <typename T>
__global__ void k_fill(int n, T v, T *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = v;
}

__global__ void k_copy(int n, const double *__restrict__ a, double *__restrict__ y) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        y[i] = a[i];
}

// One element per thread up to 2^28 elements: a full grid streams at
// 6.1 TB/s (copy) on MI355X where a grid-stride loop over 8192 blocks
// reaches 4.5.
inline int ew_grid(int n) { return std::max(1, std::min((n + kBlock - 1) / kBlock, 1 << 20)); }

}  // namespace

namespace {
template <typename T>
const T *values(Matrix &m);
template <>
const double *values<double>(Matrix &m) {
    return m.val.ptr;
}
template <>
const float *values<float>(Matrix &m) {
    return m.val32.ptr;
}

template <int MODE, typename IT, typename T, typename S>
void spmv_launch(Device &d, Matrix &m, const T *x, const S *lhs, S *y, const S *dinv,
                 S omega) {
    k_spmv<MODE, IT, T, S><<<m.n_tiles, kBlock, 0, d.stream>>>(
        m.n_tiles, m.rows, m.nnz, m.tiles.ptr, m.row_start.ptr, m.col.ptr, m.bcol.ptr,
        values<T>(m), x, lhs, y, dinv, omega);
}

template <int MODE, typename T, typename S>
void spmv_launch(Device &d, Matrix &m, const T *x, const S *lhs, S *y,
                 const S *dinv = nullptr, S omega = S(0)) {
    if (m.bcol_ok && want_idx16())
        spmv_launch<MODE, int16_t, T, S>(d, m, x, lhs, y, dinv, omega);
    else
        spmv_launch<MODE, int32_t, T, S>(d, m, x, lhs, y, dinv, omega);
}

template <typename T>
void spmv_any(Device &d, Matrix &m, const T *x, const T *lhs, T *y, int mode) {
    if (m.rows == 0) return;
    if (mode == 0)
        spmv_launch<0>(d, m, x, lhs, y);
    else if (mode == 1)
        spmv_launch<1>(d, m, x, lhs, y);
    else
        spmv_launch<2>(d, m, x, lhs, y);
    NSFEM_HIP(hipGetLastError());
}
}  // namespace

void launch_spmv(Device &d, Matrix &m, const double *x, const double *lhs, double *y,
                 int mode) {
    spmv_any(d, m, x, lhs, y, mode);
}

// The FP32 product reads the FP32 copy of the values (val32, built by
// ensure_val32 before the first call).
void launch_spmv(Device &d, Matrix &m, const float *x, const float *lhs, float *y, int mode) {
    NSFEM_REQUIRE(m.rows == 0 || m.val32.ptr, NSFEM_NOT_READY,
                  "FP32 product without the FP32 copy of the values");
    spmv_any(d, m, x, lhs, y, mode);
}

// y = A x with FP32 values and x, the products and sums in FP64
void launch_spmv_acc64(Device &d, Matrix &m, const float *x, double *y) {
    if (m.rows == 0) return;
    NSFEM_REQUIRE(m.val32.ptr, NSFEM_NOT_READY, "FP32 product without the FP32 copy of the values");
    spmv_launch<0, float, double>(d, m, x, (const double *)nullptr, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_sweep(Device &d, Matrix &m, double w, const double *dinv, const double *x,
                  const double *b, double *y) {
    if (m.rows == 0) return;
    spmv_launch<3>(d, m, x, b, y, dinv, w);
    NSFEM_HIP(hipGetLastError());
}

void launch_sweep(Device &d, Matrix &m, double w, const float *dinv, const float *x,
                  const float *b, float *y) {
    if (m.rows == 0) return;
    NSFEM_REQUIRE(m.val32.ptr, NSFEM_NOT_READY, "FP32 sweep without the FP32 copy of the values");
    spmv_launch<3>(d, m, x, b, y, dinv, (float)w);
    NSFEM_HIP(hipGetLastError());
}

void launch_spmv_affine(Device &d, Matrix &m, const double *x, const double *c, double *y) {
    launch_spmv(d, m, x, c, y, 2);
}

void launch_axpy(Device &d, int n, double a, const double *x, double *y) {
    if (n <= 0) return;
    k_axpy<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, x, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_combine(Device &d, int n, int l, const double *a, const double *X, size_t stride,
                    bool accumulate, double *y, double *y2) {
    if (n <= 0) return;
    if (l <= 0) {
        if (!accumulate) launch_fill(d, n, 0.0, y);
        if (y2) launch_copy(d, n, y, y2);
        return;
    }
    // terms beyond kCombineMax continue from the stored partial sum (an FP64
    // store and load: the same bits as one longer chain)
    for (int k0 = 0; k0 < l; k0 += kCombineMax) {
        const int m = std::min(kCombineMax, l - k0);
        CombineCoef c{};
        for (int k = 0; k < m; ++k) c.a[k] = a[k0 + k];
        const bool last = k0 + m == l;
        k_combine<<<ew_grid(n), kBlock, 0, d.stream>>>(n, m, c, X + (size_t)k0 * stride, stride,
                                                        (accumulate || k0 > 0) ? 1 : 0, y,
                                                        last ? y2 : nullptr);
        NSFEM_HIP(hipGetLastError());
    }
}

void launch_axpby(Device &d, int n, double a, double b, const double *x, const double *y,
                  double *out) {
    if (n <= 0) return;
    k_axpby<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, b, x, y, out);
    NSFEM_HIP(hipGetLastError());
}

namespace {
constexpr int kDotBlocks = 1024;  // per-block partials of the dot product

// partial: kDotBlocks doubles of scratch
void dot_accumulate(Device &d, int n, const double *a, const double *b, double *result,
                    double *partial) {
    const int nb = std::max(1, std::min((n + 4095) / 4096, kDotBlocks));
    const int chunk = std::max(1, (n + nb - 1) / nb);
    k_dot_partial<<<nb, kBlock, 0, d.stream>>>(n, chunk, a, b, partial);
    k_dot_final<<<1, kBlock, 0, d.stream>>>(nb, partial, result);
    NSFEM_HIP(hipGetLastError());
}
}  // namespace

void launch_dot_accumulate(Device &d, int n, const double *a, const double *b,
                           double *result) {
    d.dot_partial.ensure(kDotBlocks);
    dot_accumulate(d, n, a, b, result, d.dot_partial.ptr);
}

namespace {
// blocks and chunk of the deterministic dot over n entries
void dot_shape(int n, int &nb, int &chunk) {
    nb = std::max(1, std::min((n + 4095) / 4096, kDotBlocks));
    chunk = std::max(1, (n + nb - 1) / nb);
}
}  // namespace

void launch_pcg_update(Device &d, int n, double a, const float *p, const double *q, double *x,
                       double *r, float *r32, double *rr) {
    if (n <= 0) return;
    d.dot_partial.ensure(kDotBlocks);
    int nb, chunk;
    dot_shape(n, nb, chunk);
    k_pcg_update<<<nb, kBlock, 0, d.stream>>>(n, chunk, a, p, q, x, r, r32, d.dot_partial.ptr);
    k_dot_final<<<1, kBlock, 0, d.stream>>>(nb, d.dot_partial.ptr, rr);
    NSFEM_HIP(hipGetLastError());
}

namespace {
template <typename A, typename B>
void dot_any(Device &d, int n, const A *a, const B *b, double *result) {
    if (n <= 0) return;
    d.dot_partial.ensure(kDotBlocks);
    int nb, chunk;
    dot_shape(n, nb, chunk);
    k_dot_any_partial<<<nb, kBlock, 0, d.stream>>>(n, chunk, a, b, d.dot_partial.ptr);
    k_dot_final<<<1, kBlock, 0, d.stream>>>(nb, d.dot_partial.ptr, result);
    NSFEM_HIP(hipGetLastError());
}
}  // namespace

void launch_dot_accumulate(Device &d, int n, const double *a, const float *b, double *result) {
    dot_any(d, n, a, b, result);
}

void launch_dot_accumulate(Device &d, int n, const float *a, const float *b, double *result) {
    dot_any(d, n, a, b, result);
}

void launch_copy(Device &d, int n, const float *a, float *y) {
    if (n <= 0) return;
    k_convert<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_pcg_direction(Device &d, int n, const float *z, double b, float *p) {
    if (n <= 0) return;
    k_pcg_direction<<<ew_grid(n), kBlock, 0, d.stream>>>(n, z, b, p);
    NSFEM_HIP(hipGetLastError());
}

void launch_scatter(Device &d, int count, const int32_t *idx, const double *val, double *dst) {
    if (count <= 0) return;
    k_scatter<<<(count + kBlock - 1) / kBlock, kBlock, 0, d.stream>>>(count, idx, val, dst);
    NSFEM_HIP(hipGetLastError());
}

void launch_scatter_value(Device &d, int count, const int32_t *idx, double v, double *dst) {
    if (count <= 0) return;
    k_scatter_value<<<(count + kBlock - 1) / kBlock, kBlock, 0, d.stream>>>(count, idx, v, dst);
    NSFEM_HIP(hipGetLastError());
}

void launch_scatter_value(Device &d, int count, const int32_t *idx, double v, float *dst) {
    if (count <= 0) return;
    k_scatter_value<<<(count + kBlock - 1) / kBlock, kBlock, 0, d.stream>>>(count, idx, (float)v,
                                                                            dst);
    NSFEM_HIP(hipGetLastError());
}

void launch_add(Device &d, int n, const double *a, double *y) {
    if (n <= 0) return;
    k_add<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_add(Device &d, int n, const float *a, float *y) {
    if (n <= 0) return;
    k_add<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_jacobi(Device &d, int n, double w, const double *dinv, const double *r, double *x) {
    if (n <= 0) return;
    k_jacobi<<<ew_grid(n), kBlock, 0, d.stream>>>(n, w, dinv, r, x);
    NSFEM_HIP(hipGetLastError());
}

void launch_jacobi(Device &d, int n, double w, const float *dinv, const float *r, float *x) {
    if (n <= 0) return;
    k_jacobi<<<ew_grid(n), kBlock, 0, d.stream>>>(n, (float)w, dinv, r, x);
    NSFEM_HIP(hipGetLastError());
}

void launch_jacobi0(Device &d, int n, double w, const double *dinv, const double *b, double *x) {
    if (n <= 0) return;
    k_jacobi0<<<ew_grid(n), kBlock, 0, d.stream>>>(n, w, dinv, b, x);
    NSFEM_HIP(hipGetLastError());
}

void launch_jacobi0(Device &d, int n, double w, const float *dinv, const float *b, float *x) {
    if (n <= 0) return;
    k_jacobi0<<<ew_grid(n), kBlock, 0, d.stream>>>(n, (float)w, dinv, b, x);
    NSFEM_HIP(hipGetLastError());
}

void launch_fill(Device &d, int n, double v, double *y) {
    if (n <= 0) return;
    k_fill<<<ew_grid(n), kBlock, 0, d.stream>>>(n, v, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_fill(Device &d, int n, double v, float *y) {
    if (n <= 0) return;
    k_fill<<<ew_grid(n), kBlock, 0, d.stream>>>(n, (float)v, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_convert(Device &d, int n, const double *a, float *y) {
    if (n <= 0) return;
    k_convert<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_convert(Device &d, int n, const float *a, double *y) {
    if (n <= 0) return;
    k_convert<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

void launch_copy(Device &d, int n, const double *a, double *y) {
    if (n <= 0) return;
    k_copy<<<ew_grid(n), kBlock, 0, d.stream>>>(n, a, y);
    NSFEM_HIP(hipGetLastError());
}

// =====================================================================
// Persistent conjugate gradient
// =====================================================================
//
// One cooperative launch runs a whole solve.  Each block owns a contiguous
// run of tiles (rows); per iteration there are two grid-wide reductions:
//
//   phase A  p_new = r + beta*p_old (owner rows, written),
//            q = A p_new, gathering r and p_old (p_new of a neighbour is
//            recomputed from them, bit-identically, so no extra barrier),
//            x += alpha_prev * p_old (the x update of the previous step,
//            deferred to here), delta += (p_new, q)        -> reduce delta
//   phase B  alpha = gamma/delta; r -= alpha q; gamma' += (r, r) -> reduce
//
// This is textbook CG (thesis Alg. CG; gpu/matrix_math.cu:256-286) with the
// reference's 4 grid barriers per iteration fused down to 2.
//
// Mixed precision (NSFEM_PRECISION_FP32_REFINED): the matrix values and the
// r, p, q vectors are FP32; products are accumulated in FP64 (an FP32 x FP32
// product is exact in FP64), every dot is accumulated in FP64, and the
// solution x is kept and updated in FP64.  The FP32 recursive residual is
// replaced by r = b - A x, computed in FP64 from the FP64 matrix and rounded
// to FP32, at "reliable updates" (Sleijpen & van der Vorst 1996): whenever
// ||r|| has fallen by reliable_delta since the last one (at least
// ru_min_gap iterations after it), after ru_gap
// iterations, or when the recurrence says the tolerance is met; the search
// direction is kept.  Convergence is decided only on that FP64 residual.
// Measured on the 2048^2 cavity pressure matrix: FP32 storage of x or FP32
// accumulation of the products each inflates the iteration count by 1.4-1.8x,
// and so do sparse reliable updates (reliable_delta 0.1: +33%); with x and
// the accumulations in FP64 and reliable_delta 0.5 the count equals FP64 CG's.
//
// Grid reduction of a solve over at most kFlatMax blocks: each block posts
// its partial sums and an epoch flag, then every block waits for all flags
// and adds all partials itself, in block order -- one write and one round of
// polling on the critical path, and every block holds the same bits.
// Larger grids use a tree: arrivals count on one counter per group of
// kGroup blocks, the last arrival of a group adds the group's partials in
// block order and arrives at the top counter, the last top arrival adds the
// group sums in group order and releases every group.  Either way the
// summation order is fixed, so results are deterministic.  Measured on the
// 512x128 channel's pressure matrix (65,656 rows), per reduction: flat
// 2.8-5.5 us against the tree's 5.1-10 us at 16-64 blocks, about even at
// 128, and the tree ahead at 256 (the flags of all blocks sit on a few
// cache lines that every block polls).

namespace {

constexpr int kGroup = 16;
constexpr int kMaxGroups = kMaxGridBlocks / kGroup;
constexpr int kFlatMax = 128;

struct alignas(256) SyncState {
    unsigned int arrive_top;
    unsigned int pad0[63];
    unsigned int abort;
    unsigned int pad1[63];
    unsigned int arrive_group[kMaxGroups][64];
    unsigned int release_group[kMaxGroups][64];
    double result[kMaxRed];
    double pad2[24];
    double gpartial[kMaxRed][kMaxGroups];
    double partial[kMaxRed][kMaxGridBlocks];
    // flat reduction: per block the last epoch whose partials are posted,
    // and the partials by epoch parity (a block overwrites a parity's slots
    // only after every block has posted the next epoch, i.e. has read them)
    unsigned int flag[kFlatMax];
    double fpart[2][kMaxRed][kFlatMax];
};

#ifdef NSFEM_CG_TRACE
// Phase timestamps of the first kTraceIters iterations (100 MHz wall clock):
// [0] block 0 iteration start, [1] latest phase A end over all blocks,
// [2] block 0 after the delta reduction, [3] latest phase B end,
// [4] block 0 after the gamma reduction.
constexpr int kTraceIters = 4096;
__device__ unsigned long long g_cg_trace[kTraceIters][5];
#define CG_TRACE_STAMP(it, slot)                                                    \
    do {                                                                            \
        if (threadIdx.x == 0 && (it) < kTraceIters && L == 0)                       \
            g_cg_trace[(it)][(slot)] = wall_clock64();                              \
    } while (0)
#define CG_TRACE_MAX(it, slot)                                                      \
    do {                                                                            \
        __syncthreads();                                                            \
        if (threadIdx.x == 0 && (it) < kTraceIters)                                 \
            atomicMax(&g_cg_trace[(it)][(slot)], (unsigned long long)wall_clock64()); \
    } while (0)
}  // namespace
// trace builds only: read / clear the phase timestamps
extern "C" int nsfem_trace_read(unsigned long long *out, int iters) {
    const int n = iters < kTraceIters ? iters : kTraceIters;
    if (hipMemcpyFromSymbol(out, HIP_SYMBOL(g_cg_trace), sizeof(unsigned long long) * 5 * n) !=
        hipSuccess)
        return -1;
    return n;
}
extern "C" int nsfem_trace_clear() {
    static unsigned long long zero[kTraceIters][5];
    return hipMemcpyToSymbol(HIP_SYMBOL(g_cg_trace), zero, sizeof(zero)) == hipSuccess ? 0 : -1;
}
namespace {
#else
#define CG_TRACE_STAMP(it, slot) \
    do {                         \
    } while (0)
#define CG_TRACE_MAX(it, slot) \
    do {                       \
    } while (0)
#endif

struct CgArgs {
    int n;
    const int4 *meta;  // wave tile descriptors
    const int32_t *tiles;
    const int32_t *block_tiles;
    const int32_t *rs;
    const int32_t *col;
    const int16_t *dcol;  // col - wave tile's first row (int16 kernels)
    const double *val64;
    const float *val32;
    const double *b;  // [NC][n]
    double *x;        // [NC][n]
    void *r, *p0, *p1, *q;  // interleaved [n][NC] (two-column solves)
    void *st[2];            // packed {r, p, q} rows of mixed solves
    int packed2;  // two-column mixed solves: 1 packed, 2 narrow-gather iteration (0 two-phase)
    double rtol, atol, rdelta;
    int max_iter, max_refine;
    int ru_gap;      // mixed: force a reliable update after this many iterations
    int ru_min_gap;  // mixed: no drop-triggered update sooner than this
    unsigned long long timeout_ticks;
    SyncState *sync;
    CgColumnResult *res;
    int nblocks;  // blocks of this solve (batched launches)
    int flat;     // flat grid reduction for solves of <= kFlatMax blocks
    // Jacobi scaling (mixed solves): 1/sqrt(diag A) per row, or null.  The
    // FP32 iteration runs on D^-1/2 A D^-1/2 (val32 holds it) with r, p, q
    // in that scaled space; x, b and the FP64 residual passes stay unscaled
    // (r_scaled = scale * (b - A x)), so convergence is decided on the
    // unscaled residual, as without scaling.
    const double *scale;
};

// The FP64 step of x_row along the scaled search direction: x = D^-1/2 y.
__device__ __forceinline__ double x_step(const CgArgs &A, double pend, int row) {
    return A.scale ? pend * A.scale[row] : pend;
}

// Launched with at least kFlatMax threads.
__global__ void k_sync_reset(SyncState *s, int ngroups) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        s->arrive_top = 0;
        s->abort = 0;
    }
    if (i < kFlatMax) s->flag[i] = 0;
    if (i < ngroups) {
        s->arrive_group[i][0] = 0;
        s->release_group[i][0] = 0;
    }
}

__device__ __forceinline__ unsigned ld_acquire(const unsigned *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_relaxed(const double *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide sum of K values with barrier semantics.  Returns false if the
// barrier timed out or another block aborted (the caller must return).
// Memory ordering (gfx950: one L2 per XCD, not coherent across XCDs; an
// agent-scope release writes the L2 back, an agent-scope acquire invalidates
// it).  Every wave first waits for its own global stores to complete
// (s_waitcnt 0: the stores have reached its XCD's L2), the block
// synchronises, and then ONE thread per block performs the agent-scope
// release as part of its arrival (a single L2 write-back per block).  Waiting
// threads poll with relaxed agent-scope loads, which read the coherence
// point without touching the caches, and perform one acquire fence after the
// wait (a single invalidate per block), before the block reads data other
// blocks wrote.  The last arrival at a counter (the group's, then the top
// one) reads the counter's release sequence and performs an acquire fence
// before it adds the other blocks' partials.
//
// Flat form (nb <= kFlatMax): thread 0 posts the block's partials at the
// coherence point and then its flag with an agent-scope release (the
// block's stores, completed to L2 by every wave's s_waitcnt before the
// block barrier, are written back with it).  Wave 0 polls every block's
// flag with relaxed agent-scope loads, performs one acquire fence once all
// are posted, and adds the partials in block order (lane i takes blocks i,
// i + 64, ...; then a fixed butterfly), so every block computes the same
// sum.  A block posts epoch e + 1 only after its wave 0 has read epoch e's
// partials, which is what makes two parity slots enough.
template <int K, int BS>
__device__ bool grid_sum_flat(SyncState *S, double (&v)[K], const unsigned epoch,
                              unsigned long long timeout, const int b, const int nb) {
    static_assert(K <= kMaxRed, "reduction width");
    __shared__ double s_sum[K];
    __shared__ int s_ok;
    const int par = epoch & 1u;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k)
            __hip_atomic_store(&S->fpart[par][k][b], v[k], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(&S->flag[b], epoch, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int ok = 1;
        const unsigned long long t0 = wall_clock64();
        for (;;) {
            bool mine = true;
            for (int i = lane; i < nb; i += 64)
                mine = mine && __hip_atomic_load(&S->flag[i], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) >= epoch;
            if (__all(mine)) break;
            if (__any(__hip_atomic_load(&S->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) !=
                      0u)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > timeout) {
                if (lane == 0)
                    __hip_atomic_store(&S->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
            for (int i = lane; i < nb; i += 64) t += ld_relaxed(&S->fpart[par][k][i]);
            t = wave_sum(t);
            if (lane == 0) s_sum[k] = t;
        }
        if (lane == 0) s_ok = ok;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = s_sum[k];
    return s_ok != 0;
}

template <int K, int BS>
__device__ bool grid_sum(SyncState *S, double (&v)[K], unsigned &epoch, double *lds,
                         unsigned long long timeout, const int b, const int nb, const bool flat) {
    // Each flag is written by thread 0 before a __syncthreads and read by
    // the block after it; none is rewritten until the next call, which
    // every thread reaches only after passing several more barriers.
    __shared__ int s_group_last, s_top_last, s_ok;
    block_sum<K, BS>(v, lds);
    if (flat && nb <= kFlatMax) return grid_sum_flat<K, BS>(S, v, ++epoch, timeout, b, nb);
    const int g = b / kGroup;
    const int ng = (nb + kGroup - 1) / kGroup;
    const int gsize = min(kGroup, nb - g * kGroup);
    ++epoch;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores are complete
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) S->partial[k][b] = v[k];
        const unsigned prev = __hip_atomic_fetch_add(&S->arrive_group[g][0], 1u, __ATOMIC_RELEASE,
                                                     __HIP_MEMORY_SCOPE_AGENT);
        const bool last = prev == epoch * (unsigned)gsize - 1u;
        // the last arrival read the counter's release sequence (every
        // member's release RMW): an acquire fence after it synchronises with
        // all of them before the block reads their partials (the block
        // barrier below carries the order to the reading threads)
        if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_group_last = last ? 1 : 0;
    }
    __syncthreads();
    if (s_group_last) {
        // group sum, members in block order (lane i of wave 0 holds member i);
        // the partials are read at the coherence point
        double gs[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double t = 0.0;
            if (threadIdx.x < gsize) t = ld_relaxed(&S->partial[k][g * kGroup + threadIdx.x]);
            gs[k] = threadIdx.x < 64 ? wave_sum(t) : 0.0;
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                __hip_atomic_store(&S->gpartial[k][g], gs[k], __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            const unsigned prev = __hip_atomic_fetch_add(&S->arrive_top, 1u, __ATOMIC_RELEASE,
                                                         __HIP_MEMORY_SCOPE_AGENT);
            const bool last = prev == epoch * (unsigned)ng - 1u;
            if (last) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // as above
            s_top_last = last ? 1 : 0;
        }
        __syncthreads();
        if (s_top_last) {
            double ts[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                double acc = 0.0;
                for (int i = threadIdx.x; i < ng; i += BS) acc += ld_relaxed(&S->gpartial[k][i]);
                ts[k] = acc;
            }
            block_sum<K, BS>(ts, lds);
            if (threadIdx.x == 0) {
#pragma unroll
                for (int k = 0; k < K; ++k)
                    __hip_atomic_store(&S->result[k], ts[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            }
            __syncthreads();
            for (int i = threadIdx.x; i < ng; i += BS)
                __hip_atomic_store(&S->release_group[i][0], epoch, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // wait for the release of this block's group (bounded)
    if (threadIdx.x == 0) {
        int ok = 1;
        const unsigned long long t0 = wall_clock64();
        while (__hip_atomic_load(&S->release_group[g][0], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT) < epoch) {
            if (__hip_atomic_load(&S->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                ok = 0;
                break;
            }
            if (wall_clock64() - t0 > timeout) {
                __hip_atomic_store(&S->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        s_ok = ok;
    }
    __syncthreads();
    const bool ok = s_ok != 0;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = ld_relaxed(&S->result[k]);
    return ok;
}

template <int NC, typename T>
struct VecN {
    T v[NC];
};

template <int NC, typename T>
__device__ __forceinline__ VecN<NC, T> ldv(const T *p, int i) {
    VecN<NC, T> o;
    if constexpr (NC == 1) {
        o.v[0] = p[i];
    } else if constexpr (std::is_same<T, double>::value) {
        const double2 t = reinterpret_cast<const double2 *>(p)[i];
        o.v[0] = t.x;
        o.v[1] = t.y;
    } else {
        const float2 t = reinterpret_cast<const float2 *>(p)[i];
        o.v[0] = t.x;
        o.v[1] = t.y;
    }
    return o;
}

template <int NC, typename T>
__device__ __forceinline__ void stv(T *p, int i, const VecN<NC, T> &o) {
    if constexpr (NC == 1) {
        p[i] = o.v[0];
    } else if constexpr (std::is_same<T, double>::value) {
        reinterpret_cast<double2 *>(p)[i] = make_double2(o.v[0], o.v[1]);
    } else {
        reinterpret_cast<float2 *>(p)[i] = make_float2(o.v[0], o.v[1]);
    }
}

// ---------------------------------------------------------------------
// Wave tiles: the persistent CG partitions rows into tiles of at most 64
// rows and kWaveNnz nonzeros; each wavefront stages its own tile through
// its own LDS region (no block-wide synchronisation inside a phase), so the
// 16 waves of a block overlap their memory latencies independently.
constexpr int kWItems = kWaveNnz / 64;  // nonzeros per lane per tile

struct WTile {
    int r0, r1, k0, nz;
};

// Sum of a row's products p[ra..rb) in column order, four LDS reads in
// flight at a time (the adds stay sequential: bit-identical to one by one).
__device__ __forceinline__ double row_sum(const double *p, int ra, int rb) {
    double s = 0.0;
    int k = ra;
    for (; k + 4 <= rb; k += 4) {
        const double a0 = p[k], a1 = p[k + 1], a2 = p[k + 2], a3 = p[k + 3];
        s += a0;
        s += a1;
        s += a2;
        s += a3;
    }
    for (; k < rb; ++k) s += p[k];
    return s;
}

// The same for NC interleaved columns (products [k][NC]), added to s[c];
// two nonzeros' products in flight at a time, each column in column order.
template <int NC>
__device__ __forceinline__ void row_sums(const double *p, int ra, int rb, double (&s)[NC]) {
    if constexpr (NC == 1) {
        s[0] += row_sum(p, ra, rb);
    } else {
        int k = ra;
        for (; k + 2 <= rb; k += 2) {
            double a[2][NC];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int c = 0; c < NC; ++c) a[j][c] = p[(k + j) * NC + c];
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int c = 0; c < NC; ++c) s[c] += a[j][c];
        }
        for (; k < rb; ++k)
#pragma unroll
            for (int c = 0; c < NC; ++c) s[c] += p[k * NC + c];
    }
}

// Tile metadata: read from global memory, or from the block's copy in LDS
// (MetaLds: indexed from the block's first tile), which keeps the per-tile
// descriptor read out of the vector-memory counter the gathers wait on.
// The LDS copy holds {first row, first nonzero} of the block's tiles and of
// one past its last (a tile ends where the next begins).
struct MetaLds {
    const int2 *m;
    int t0;
};
__device__ __forceinline__ WTile wave_tile(const int4 *meta, int t);
__device__ __forceinline__ WTile wave_tile(const MetaLds &meta, int t) {
    const int2 a = meta.m[t - meta.t0], b = meta.m[t - meta.t0 + 1];
    WTile w;
    w.r0 = __builtin_amdgcn_readfirstlane(a.x);
    w.r1 = __builtin_amdgcn_readfirstlane(b.x);
    w.k0 = __builtin_amdgcn_readfirstlane(a.y);
    w.nz = __builtin_amdgcn_readfirstlane(b.y) - w.k0;
    return w;
}

// Copy tiles T0 .. T1 - 1 of the block's descriptors (T1 > T0) into `lds`
// as MetaLds entries, T1 - T0 + 1 of them.
template <int BS>
__device__ __forceinline__ void load_meta_lds(const int4 *meta, int T0, int T1, int2 *lds) {
    for (int i = threadIdx.x; i <= T1 - T0; i += BS) {
        const int4 m = meta[i < T1 - T0 ? T0 + i : T1 - 1];
        lds[i] = i < T1 - T0 ? make_int2(m.x, m.z) : make_int2(m.y, m.z + m.w);
    }
    __syncthreads();
}

__device__ __forceinline__ WTile wave_tile(const int4 *meta, int t) {
    const int4 m = meta[t];
    WTile w;
    w.r0 = __builtin_amdgcn_readfirstlane(m.x);
    w.r1 = __builtin_amdgcn_readfirstlane(m.y);
    w.k0 = __builtin_amdgcn_readfirstlane(m.z);
    w.nz = __builtin_amdgcn_readfirstlane(m.w);
    return w;
}

// One-column mixed solves keep each row's FP32 CG state packed as {r, p, q}
// in 12 bytes, so that one dwordx3 gather brings everything a neighbour
// contributes.  A wave64 gather costs the same return-path time whatever its
// width per lane, so one wide gather replaces the two narrow ones (r, p) of
// the two-phase iteration.  (The solution stays FP64 and is updated every
// iteration: an FP32 accumulator between reliable updates, multiplied by an
// ill-conditioned A, pollutes the true residual at ~eps32 * cond(A) of its
// last level and cost the 2048^2 cavity's pressure solve 50% more
// iterations.)
struct alignas(4) RowState {
    float r, p, q;
};

__device__ __forceinline__ RowState ld_state(const void *st, int i) {
    return reinterpret_cast<const RowState *>(st)[i];
}
__device__ __forceinline__ void st_state(void *st, int i, const RowState &v) {
    reinterpret_cast<RowState *>(st)[i] = v;
}

// Two-column mixed solves keep both columns' FP32 {r, p} of a row in one
// 16-byte word {r0, r1, p0, p1} and their q in an 8-byte word {q0, q1}: a
// state copy holds the n words of the first kind, then the n of the second.
// A neighbour's contribution to both columns takes two gathers.
__device__ __forceinline__ float4 *st2_rp(void *st) { return reinterpret_cast<float4 *>(st); }
__device__ __forceinline__ const float4 *st2_rp(const void *st) {
    return reinterpret_cast<const float4 *>(st);
}
__device__ __forceinline__ float2 *st2_q(void *st, int n) {
    return reinterpret_cast<float2 *>(reinterpret_cast<float4 *>(st) + n);
}
__device__ __forceinline__ const float2 *st2_q(const void *st, int n) {
    return reinterpret_cast<const float2 *>(reinterpret_cast<const float4 *>(st) + n);
}

// Column word of nonzero k: the column (IT = int32_t) or its offset from the
// first row of the nonzero's wave tile (IT = int16_t).
template <typename IT>
__device__ __forceinline__ int col_word(const CgArgs &A, int k) {
    if constexpr (std::is_same<IT, int16_t>::value)
        return (int)ld_stream(A.dcol + k);
    else
        return ld_stream(A.col + k);
}
// Column from a column word of a wave tile whose first row is r0.
template <typename IT>
__device__ __forceinline__ int col_of(int word, int r0) {
    if constexpr (std::is_same<IT, int16_t>::value)
        return r0 + word;
    else
        return word;
}
template <typename IT>
__device__ __forceinline__ int col_at(const CgArgs &A, int r0, int k) {
    return col_of<IT>(col_word<IT>(A, k), r0);
}

// FP64 residual over the block's wave tiles: r = b - A x (x column-major).
// SETUP: also p_old = 0 and b.b.
// Accumulates per column: acc[c] = r.r (FP64), acc[NC+c] = b.b,
// acc[2NC+c] = (FP32 r).(FP32 r).
// PACKED (one-column mixed solves): r goes to the packed state `st`; SETUP
// also clears its p and q.
template <int NC, int BS, bool MIXED, bool SETUP, typename IT, bool PACKED = false,
          typename MP = const int4 *>
__device__ void residual_phase(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                               double (&acc)[3 * NC], void *st = nullptr) {
    static_assert(!PACKED || ((NC == 1 || NC == 2) && MIXED), "packed state: FP32 columns");
    using VT = typename std::conditional<MIXED, float, double>::type;
    constexpr int NW = BS / 64;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    VT *r = (VT *)A.r;
    VT *pold = (VT *)A.p0;
    for (int t = T0 + wave; t < T1; t += NW) {
        const WTile w = wave_tile(meta, t);
        const int row = w.r0 + lane;
        const bool has_row = row < w.r1;
        const int rowc = has_row ? row : w.r0;
        double s[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) s[c] = 0.0;
        double bb[NC];
        int ra = A.rs[rowc] - w.k0;
        int rb = A.rs[rowc + 1] - w.k0;
#pragma unroll
        for (int c = 0; c < NC; ++c) bb[c] = A.b[(size_t)c * A.n + rowc];
        if (!has_row) rb = ra;
        if (w.nz <= kWaveNnz) {
            // Branch-free staging: out-of-tile lanes read index 0 of the
            // column/value arrays (always allocated) and gather x[0] (the
            // matrix has >= 1 column); their products are zeroed and never
            // summed.  All loads of a batch issue back to back.
            constexpr int IB = (NC == 2 || MIXED) ? 4 : kWItems;  // register budget
#pragma unroll
            for (int i0 = 0; i0 < kWItems; i0 += IB) {
                int cj[IB];
                double vj[IB];
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    const bool ok = k < w.nz;
                    const int idx = ok ? w.k0 + k : 0;
                    const int cc = col_at<IT>(A, w.r0, idx);
                    const double vv = A.val64[idx];
                    cj[i] = ok ? cc : 0;
                    vj[i] = ok ? vv : 0.0;
                }
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        lds_wave[k * NC + c] = vj[i] * A.x[(size_t)c * A.n + cj[i]];
                }
            }
            __syncwarp();
            if (has_row) row_sums<NC>(lds_wave, ra, rb, s);
            __syncwarp();
        } else {  // a single row longer than a tile: wave-wide sum
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double part = 0.0;
                for (int k = lane; k < w.nz; k += 64)
                    part += A.val64[w.k0 + k] *
                            A.x[(size_t)c * A.n + col_at<IT>(A, w.r0, w.k0 + k)];
                s[c] = wave_sum(part);
            }
        }
        if (has_row) {
            VecN<NC, VT> rv, zero;
            const double sc = A.scale ? A.scale[row] : 1.0;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const double rr = bb[c] - s[c];
                acc[c] += rr * rr;
                if (SETUP) acc[NC + c] += bb[c] * bb[c];
                rv.v[c] = (VT)(A.scale ? rr * sc : rr);
                if (MIXED) acc[2 * NC + c] += (double)rv.v[c] * (double)rv.v[c];
                zero.v[c] = (VT)0;
            }
            if constexpr (PACKED && NC == 1) {
                if constexpr (SETUP)
                    st_state(st, row, RowState{(float)rv.v[0], 0.f, 0.f});
                else
                    reinterpret_cast<RowState *>(st)[row].r = (float)rv.v[0];
            } else if constexpr (PACKED) {
                float4 *rp = st2_rp(st);
                if constexpr (SETUP) {
                    rp[row] = make_float4((float)rv.v[0], (float)rv.v[1], 0.f, 0.f);
                    st2_q(st, A.n)[row] = make_float2(0.f, 0.f);
                } else {
                    reinterpret_cast<float2 *>(rp + row)[0] =
                        make_float2((float)rv.v[0], (float)rv.v[1]);
                }
            } else {
                stv<NC, VT>(r, row, rv);
                if (SETUP) {
                    stv<NC, VT>(pold, row, zero);
                }
            }
        }
    }
}

// Phase A over the block's wave tiles:
//   p_new = r + beta p_old (owner rows), q = A p_new (gathering r, p_old),
//   deferred FP64 solution update x += pend * p_old, delta += (p_new, q).
// The products and row sums accumulate in FP64 in both precisions (the
// product of two FP32 numbers is exact in FP64); only q is stored as VT.
template <int NC, int BS, bool MIXED, typename IT, typename MP>
__device__ void phase_a(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                        const typename std::conditional<MIXED, float, double>::type *pold,
                        typename std::conditional<MIXED, float, double>::type *pnew,
                        const typename std::conditional<MIXED, float, double>::type (&beta)[NC],
                        const double (&pend)[NC], bool has_pend, double (&delta)[NC]) {
    using VT = typename std::conditional<MIXED, float, double>::type;
    using MT = VT;
    constexpr int NW = BS / 64;
    const MT *val = MIXED ? (const MT *)A.val32 : (const MT *)A.val64;
    const VT *r = (const VT *)A.r;
    VT *q = (VT *)A.q;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int t = T0 + wave;
    if (t >= T1) return;
    // Software pipeline: the column indices of the next tile load while the
    // current tile's values and gathers are in flight.  Lanes beyond a
    // tile's nonzeros read entry 0 of the column/value arrays (a valid
    // column: upload makes col[0] = 0 for an empty matrix); their products
    // land in LDS slots that no row sum reads, so no loaded value is masked
    // and no load has to be waited for before the products are staged.
    // The prefetched words stay raw until the tile is consumed (no ALU use
    // of a load in flight).  Lanes beyond the tile read its first nonzero
    // (a valid column of the tile); a tile without nonzeros gathers nothing.
    // Item slices wholly beyond a tile's nonzeros issue no loads (a 64-row
    // tile of a 7-point matrix fills 7 of its 8 slices): 5% faster here.
    WTile w = wave_tile(meta, t);
    int cn[kWItems];
    {
        const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
        for (int i = 0; i < kWItems; ++i) {
            const int k = lane + 64 * i;
            cn[i] = 64 * i < w.nz ? col_word<IT>(A, k < w.nz ? w.k0 + k : kf) : 0;
        }
    }
    for (; t < T1; t += NW) {
        const WTile cur = w;
        int cj[kWItems];
#pragma unroll
        for (int i = 0; i < kWItems; ++i) cj[i] = col_of<IT>(cn[i], cur.r0);
        // 1. next tile: descriptor and column words
        const int tn = t + NW;
        if (tn < T1) {
            w = wave_tile(meta, tn);
            const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
            for (int i = 0; i < kWItems; ++i) {
                const int k = lane + 64 * i;
                cn[i] = 64 * i < w.nz ? col_word<IT>(A, k < w.nz ? w.k0 + k : kf) : 0;
            }
        }
        // 2. owner rows (clamped so the loads stay unconditional)
        const int row = cur.r0 + lane;
        const bool has_row = row < cur.r1;
        const int rowc = has_row ? row : cur.r0;
        const int rsa = A.rs[rowc];
        const int rsb = A.rs[rowc + 1];
        const VecN<NC, VT> rv = ldv<NC, VT>(r, rowc);
        const VecN<NC, VT> pv = ldv<NC, VT>(pold, rowc);
        double xo[NC];
        if (has_pend) {
#pragma unroll
            for (int c = 0; c < NC; ++c) xo[c] = A.x[(size_t)c * A.n + rowc];
        }
        double s[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) s[c] = 0.0;
        if (cur.nz == 0) {
            // rows without nonzeros: q = 0
        } else if (cur.nz <= kWaveNnz) {
            // 3. values and gathers, then one wait for the whole tile
            constexpr int IB = (NC == 2) ? 4 : kWItems;  // register budget
#pragma unroll
            for (int i0 = 0; i0 < kWItems; i0 += IB) {
                if (64 * i0 >= cur.nz) break;
                MT vj[IB];
                VecN<NC, VT> rj[IB], pj[IB];
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    if (64 * (i0 + i) < cur.nz) {
                        vj[i] = ld_stream(val + (k < cur.nz ? cur.k0 + k : 0));
                        rj[i] = ldv<NC, VT>(r, cj[i0 + i]);
                        pj[i] = ldv<NC, VT>(pold, cj[i0 + i]);
                    } else {
                        vj[i] = (MT)0;
#pragma unroll
                        for (int c = 0; c < NC; ++c) rj[i].v[c] = pj[i].v[c] = (VT)0;
                    }
                }
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    if (64 * (i0 + i) < cur.nz) {
#pragma unroll
                        for (int c = 0; c < NC; ++c)
                            lds_wave[k * NC + c] =
                                (double)vj[i] * (double)fma_t<VT>(beta[c], pj[i].v[c], rj[i].v[c]);
                    }
                }
            }
            __syncwarp();
            const int ra = rsa - cur.k0;
            const int rb = has_row ? rsb - cur.k0 : ra;
            row_sums<NC>(lds_wave, ra, rb, s);
            __syncwarp();
        } else {
            // a single row longer than a tile: wave-wide sum
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double part = 0.0;
                for (int k = lane; k < cur.nz; k += 64) {
                    const int j = col_at<IT>(A, cur.r0, cur.k0 + k);
                    const VecN<NC, VT> rjj = ldv<NC, VT>(r, j);
                    const VecN<NC, VT> pjj = ldv<NC, VT>(pold, j);
                    part += (double)val[cur.k0 + k] *
                            (double)fma_t<VT>(beta[c], pjj.v[c], rjj.v[c]);
                }
                s[c] = wave_sum(part);
            }
        }
        if (has_row) {
            VecN<NC, VT> pn, qq;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                pn.v[c] = fma_t<VT>(beta[c], pv.v[c], rv.v[c]);
                qq.v[c] = (VT)s[c];
                delta[c] += (double)pn.v[c] * s[c];
            }
            stv<NC, VT>(pnew, row, pn);
            stv<NC, VT>(q, row, qq);
            if (has_pend) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
                    A.x[(size_t)c * A.n + row] =
                        __builtin_fma(x_step(A, pend[c], row), (double)pv.v[c], xo[c]);
            }
        }
    }
}

// Phase B over the block's rows: r -= alpha q, g += (r, r); rows are
// processed 4 per thread with all loads issued first.
template <int NC, int BS, typename VT>
__device__ void phase_b(const CgArgs &A, int R0, int R1, const VT (&alpha)[NC], double (&g)[NC]) {
    VT *r = (VT *)A.r;
    const VT *q = (const VT *)A.q;
    constexpr int U = 4;
    for (int base = R0 + (int)threadIdx.x; base < R1; base += U * BS) {
        VecN<NC, VT> rv[U], qv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {  // clamped rows keep the loads unconditional
            const int row = min(base + u * BS, R1 - 1);
            rv[u] = ldv<NC, VT>(r, row);
            qv[u] = ldv<NC, VT>(q, row);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int row = base + u * BS;
            if (row < R1) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    rv[u].v[c] = fma_t<VT>(-alpha[c], qv[u].v[c], rv[u].v[c]);
                    g[c] += (double)rv[u].v[c] * (double)rv[u].v[c];
                }
                stv<NC, VT>(r, row, rv[u]);
            }
        }
    }
}

// x += pend p over the block's rows (FP64 solution, p in VT)
template <int NC, int BS, typename VT>
__device__ void flush_x(const CgArgs &A, int R0, int R1, const VT *p, const double (&pend)[NC]) {
    for (int row = R0 + (int)threadIdx.x; row < R1; row += BS) {
        const VecN<NC, VT> pv = ldv<NC, VT>(p, row);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double *xp = A.x + (size_t)c * A.n + row;
            *xp = __builtin_fma(x_step(A, pend[c], row), (double)pv.v[c], *xp);
        }
    }
}

// The persistent CG run by one block of a solve that owns `nb` blocks: `bid`
// is the block's index among them (reduction slots), `L` its logical index
// (tile ownership).  prod_raw / lds are the block's LDS.
template <int NC, int BS, bool MIXED, typename IT, typename MP>
__device__ __forceinline__ void cg_body_m(const CgArgs &A, MP meta, const int bid, const int nb,
                                          const int L, double *prod_raw, double *lds) {
    using VT = typename std::conditional<MIXED, float, double>::type;
    const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
    const int R0 = A.tiles[T0], R1 = A.tiles[T1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *lds64 = prod_raw + wave * kWaveNnz * NC;
    VT *pbuf[2] = {(VT *)A.p0, (VT *)A.p1};
    unsigned epoch = 0;
    const bool writer = (L == 0 && threadIdx.x == 0);

    // ------------------------------------------------------------ setup
    double acc[3 * NC];
#pragma unroll
    for (int k = 0; k < 3 * NC; ++k) acc[k] = 0.0;
    residual_phase<NC, BS, MIXED, true, IT, false, MP>(A, meta, T0, T1, lds64, acc);
    if (!grid_sum<3 * NC, BS>(A.sync, acc, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
    {
        double tol2[NC], gamma[NC], rmax2[NC], rho[NC];
        bool active[NC];
        int iters[NC], status[NC], refines = 0;
        VT beta[NC];
        double pend[NC];
        double last_rr[NC];
        bool any = false;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double bn = sqrt(acc[NC + c]), r0n = sqrt(acc[c]);
            const double ref = bn > 0.0 ? bn : r0n;
            const double tol = fmax(A.rtol * ref, A.atol);
            tol2[c] = tol * tol;
            active[c] = acc[c] > tol2[c];
            any = any || active[c];
            gamma[c] = MIXED ? acc[2 * NC + c] : acc[c];
            // the drop triggers compare the recursion with the true residual
            // in the iteration's (scaled) space; rho converts a scaled norm
            // into the unscaled one the tolerance is on
            rmax2[c] = A.scale ? acc[2 * NC + c] : acc[c];
            rho[c] = A.scale && acc[2 * NC + c] > 0.0 ? acc[c] / acc[2 * NC + c] : 1.0;
            last_rr[c] = acc[c];
            iters[c] = 0;
            status[c] = NSFEM_OK;
            beta[c] = (VT)0;
            pend[c] = 0.0;
            if (writer) {
                A.res[c].rhs_norm = bn;
                A.res[c].initial_residual = r0n;
                A.res[c].tolerance = tol;
            }
        }
        int k = 0;
        int cur = 0;  // pbuf[cur] = p_old
        bool has_pend = false;
        int last_ru = 0;
        int stall[NC];
        double ru_rr[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            stall[c] = 0;
            ru_rr[c] = acc[c];
        }
        while (any) {
            VT *pold = pbuf[cur];
            VT *pnew = pbuf[cur ^ 1];
            double delta[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) delta[c] = 0.0;
            CG_TRACE_STAMP(k, 0);
            phase_a<NC, BS, MIXED, IT, MP>(A, meta, T0, T1, lds64, pold, pnew, beta, pend, has_pend, delta);
            CG_TRACE_MAX(k, 1);
            if (!grid_sum<NC, BS>(A.sync, delta, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            CG_TRACE_STAMP(k, 2);
            double alpha[NC];
            VT alpha_v[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                alpha[c] = 0.0;
                if (active[c]) {
                    if (!(delta[c] > 0.0)) {
                        status[c] = NSFEM_BREAKDOWN;
                        active[c] = false;
                    } else {
                        alpha[c] = gamma[c] / delta[c];
                    }
                }
                alpha_v[c] = (VT)alpha[c];
            }
            double g[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) g[c] = 0.0;
            phase_b<NC, BS, VT>(A, R0, R1, alpha_v, g);
            CG_TRACE_MAX(k, 3);
            if (!grid_sum<NC, BS>(A.sync, g, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            CG_TRACE_STAMP(k, 4);
            ++k;
            bool want_ru = false;
            bool dropped[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                dropped[c] = false;
                if (active[c]) {
                    iters[c] = k;
                    if (MIXED) {
                        // a drop by rdelta triggers the update once
                        // ru_min_gap iterations have passed, a drop by
                        // rdelta^2 at once (FP32 rounding must not pile up
                        // over a steep fall of the residual)
                        const double d2 = A.rdelta * A.rdelta;
                        dropped[c] = g[c] * rho[c] <= tol2[c] ||
                                     (g[c] < d2 * rmax2[c] &&
                                      (k - last_ru >= A.ru_min_gap || g[c] < d2 * d2 * rmax2[c]));
                        want_ru = want_ru || dropped[c];
                    } else if (g[c] <= tol2[c]) {
                        active[c] = false;  // its last update is still pending
                        last_rr[c] = g[c];
                    }
                }
                pend[c] = alpha[c];
            }
            has_pend = true;
            const bool out_of_iters = k >= A.max_iter;
            if (MIXED && k - last_ru >= A.ru_gap) want_ru = true;
            if (MIXED && (want_ru || out_of_iters)) {
                last_ru = k;
                // apply the pending x += alpha p_new (FP64), recompute r
                flush_x<NC, BS, VT>(A, R0, R1, pnew, pend);
                has_pend = false;
#pragma unroll
                for (int c = 0; c < NC; ++c) pend[c] = 0.0;
                double dummy[1] = {0.0};
                if (!grid_sum<1, BS>(A.sync, dummy, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
                double ra[3 * NC];
#pragma unroll
                for (int kk = 0; kk < 3 * NC; ++kk) ra[kk] = 0.0;
                residual_phase<NC, BS, MIXED, false, IT, false, MP>(A, meta, T0, T1, lds64, ra);
                if (!grid_sum<3 * NC, BS>(A.sync, ra, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
                ++refines;
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    last_rr[c] = ra[c];
                    if (active[c]) {
                        if (ra[c] <= tol2[c]) {
                            active[c] = false;
                        } else {
                            g[c] = ra[2 * NC + c];
                            rmax2[c] = A.scale ? ra[2 * NC + c] : ra[c];
                            rho[c] = A.scale && ra[2 * NC + c] > 0.0 ? ra[c] / ra[2 * NC + c]
                                                                     : 1.0;
                            // when the recursion reports a drop by rdelta,
                            // the FP64 residual must fall below
                            // (1 + rdelta)/2 of its best value so far; three
                            // misses in a row end the FP32 iteration for
                            // this column (an update forced by ru_gap alone
                            // is not judged)
                            if (dropped[c]) {
                                const double s = 0.5 * (1.0 + A.rdelta);
                                stall[c] = ra[c] > s * s * ru_rr[c] ? stall[c] + 1 : 0;
                            }
                            ru_rr[c] = fmin(ru_rr[c], ra[c]);
                            if (stall[c] >= 3) {
                                status[c] = NSFEM_STAGNATION;
                                active[c] = false;
                            }
                        }
                    }
                }
                if (refines >= A.max_refine) {
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        if (active[c]) {
                            status[c] = NSFEM_MAX_ITERATIONS_REACHED;
                            active[c] = false;
                        }
                }
            }
            any = false;
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                if (active[c] && out_of_iters) {
                    status[c] = NSFEM_MAX_ITERATIONS_REACHED;
                    active[c] = false;
                    last_rr[c] = g[c];
                }
                any = any || active[c];
                beta[c] = active[c] ? (VT)(g[c] / gamma[c]) : (VT)0;
                if (active[c]) gamma[c] = g[c];
            }
            if (!any) {
                // flush the pending x update of the last iteration
                if (has_pend) flush_x<NC, BS, VT>(A, R0, R1, pnew, pend);
                break;
            }
            cur ^= 1;
        }
        if (writer) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                A.res[c].iterations = iters[c];
                A.res[c].refinements = refines;
                A.res[c].status = status[c];
                A.res[c].last_residual = sqrt(last_rr[c]);
            }
        }
        return;
    }
timeout:
    if (writer) {
        for (int c = 0; c < NC; ++c) A.res[c].status = NSFEM_BARRIER_TIMEOUT;
    }
}

// Tile descriptors of a block: copied into LDS once per launch when they fit
// (one-column solves; the two-column solves' product slices fill the LDS).
// 4096 tiles cover the 2048^2 cavity's P2 matrices on 256 blocks (~2,500
// tiles each).
constexpr int kMetaLds = 4096;

template <int NC, int BS, bool MIXED, typename IT>
__device__ __forceinline__ void cg_body(const CgArgs &A, const int bid, const int nb, const int L,
                                        double *prod_raw, double *lds) {
    if constexpr (NC == 1) {
        __shared__ int2 tmeta[kMetaLds + 1];
        const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
        if (T1 > T0 && T1 - T0 <= kMetaLds) {
            load_meta_lds<BS>(A.meta, T0, T1, tmeta);
            cg_body_m<NC, BS, MIXED, IT>(A, MetaLds{tmeta, T0}, bid, nb, L, prod_raw, lds);
            return;
        }
    }
    cg_body_m<NC, BS, MIXED, IT>(A, A.meta, bid, nb, L, prod_raw, lds);
}

// ---------------------------------------------------------------------
// One-column mixed solves: single-reduction iteration over the packed FP32
// row state.
//
// With r_{k-1}, p_{k-1}, q_{k-1} = A p_{k-1} of the last phase, one fused
// phase forms, over the block's wave tiles,
//   r_k = r_{k-1} - alpha q_{k-1},  p_k = r_k + beta p_{k-1}  (owner rows),
//   q_k = A p_k, each neighbour's p_k recomputed from its gathered state by
//   the same two roundings as its owner's (products summed in FP64),
//   the deferred FP64 solution update x += pend p_{k-1},
// and the FP64 sums (p_k, q_k), (r_k, r_k), (r_k, q_k), (q_k, q_k), all in one
// grid reduction: gamma_{k+1} = (r,r) - 2 alpha (r,q) + alpha^2 (q,q) gives
// beta_{k+1} without a second reduction (Meurant's single-synchronisation
// form of Hestenes-Stiefel CG, equal to it in exact arithmetic).  The state
// of step k goes to the other buffer, so no block overwrites rows that
// another still gathers.  (The FP64 solve keeps the two-phase iteration: its
// packed state would need two gathers per nonzero, and measured slower.)
__device__ __forceinline__ float next_r(float alpha, float r, float q) {
    return __builtin_fmaf(-alpha, q, r);
}

// WI: nonzeros per lane per tile (the host's bound on every multi-row tile's
// nonzeros over 64); a single longer row is summed wave-wide.
template <int BS, typename IT, int WI, typename MP>
__device__ void phase_f1(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                         const void *st_old,
                         void *st_new, const float alpha, const float beta, const double pend,
                         const bool has_pend, double (&sums)[4]) {
    constexpr int NW = BS / 64;
    const float *val = A.val32;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int t = T0 + wave;
    if (t >= T1) return;
    // column words of the next tile stay raw until the tile is consumed;
    // lanes beyond a tile read its first nonzero (a valid column)
    WTile w = wave_tile(meta, t);
    int cn[WI];
    {
        const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
        for (int i = 0; i < WI; ++i) {
            const int k = lane + 64 * i;
            cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
        }
    }
    for (; t < T1; t += NW) {
        const WTile cur = w;
        int cj[WI];
#pragma unroll
        for (int i = 0; i < WI; ++i) cj[i] = col_of<IT>(cn[i], cur.r0);
        const int tn = t + NW;
        if (tn < T1) {
            w = wave_tile(meta, tn);
            const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
            for (int i = 0; i < WI; ++i) {
                const int k = lane + 64 * i;
                cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
            }
        }
        const int row = cur.r0 + lane;
        const bool has_row = row < cur.r1;
        const int rowc = has_row ? row : cur.r0;
        const int rsa = A.rs[rowc];
        const int rsb = A.rs[rowc + 1];
        const RowState so = ld_state(st_old, rowc);
        double xo = 0.0;
        if (has_pend) xo = A.x[rowc];
        double s = 0.0;
        if (cur.nz == 0) {
            // rows without nonzeros: q = 0
        } else if (cur.nz <= 64 * WI) {
            constexpr int IB = 4;  // registers in flight
#pragma unroll
            // (loads stay unconditional: guarding the item slices beyond a
            // tile's nonzeros serialised this loop's gathers, 2x slower)
            for (int i0 = 0; i0 < WI; i0 += IB) {
                float vj[IB];
                RowState gj[IB];
#pragma unroll
                for (int i = 0; i < IB && i0 + i < WI; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    vj[i] = ld_stream(val + (k < cur.nz ? cur.k0 + k : 0));
                    gj[i] = ld_state(st_old, cj[i0 + i]);
                }
#pragma unroll
                for (int i = 0; i < IB && i0 + i < WI; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    lds_wave[k] = (double)vj[i] *
                                  (double)__builtin_fmaf(beta, gj[i].p, next_r(alpha, gj[i].r, gj[i].q));
                }
            }
            __syncwarp();
            const int ra = rsa - cur.k0;
            const int rb = has_row ? rsb - cur.k0 : ra;
            s = row_sum(lds_wave, ra, rb);
            __syncwarp();
        } else {
            // a single row longer than a tile: wave-wide sum
            double part = 0.0;
            for (int k = lane; k < cur.nz; k += 64) {
                const RowState g = ld_state(st_old, col_at<IT>(A, cur.r0, cur.k0 + k));
                part += (double)val[cur.k0 + k] *
                        (double)__builtin_fmaf(beta, g.p, next_r(alpha, g.r, g.q));
            }
            s = wave_sum(part);
        }
        if (has_row) {
            RowState sn;
            sn.r = next_r(alpha, so.r, so.q);
            sn.p = __builtin_fmaf(beta, so.p, sn.r);
            sn.q = (float)s;
            const double rd = (double)sn.r, qd = (double)sn.q;
            sums[0] += (double)sn.p * s;
            sums[1] += rd * rd;
            sums[2] += rd * qd;
            sums[3] += qd * qd;
            st_state(st_new, row, sn);
            if (has_pend) A.x[row] = __builtin_fma(x_step(A, pend, row), (double)so.p, xo);
        }
    }
}

// The same fused phase with group loads (one-column mixed solves on
// matrices with 16-bit offsets).  Each lane owns 8 consecutive nonzeros of a
// wave tile, counted from the tile's first nonzero rounded down to a
// multiple of 8, so a tile's values take two 16-byte loads per lane and its
// column words one, instead of one load per nonzero each.  Entries of the
// neighbouring tiles in the first and last group are computed too (their
// offsets, decoded against this tile's first row, are clamped into the
// matrix) into the slice's 8 entries of padding or its unused tail, and are
// never summed.  The next tile's column words and values load while the
// current tile's gathers are in flight.  Products and sums are those of
// phase_f1 (p_k with the owner's two roundings, FP64 products of FP32
// factors, row sums in column order), so the iterates are identical.
template <int BS, typename MP>
__device__ void phase_f1g(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                          const void *st_old, void *st_new, const float alpha, const float beta,
                          const double pend, const bool has_pend, double (&sums)[4]) {
    constexpr int NW = BS / 64;
    typedef short i16x8 __attribute__((ext_vector_type(8)));
    typedef float f32x4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int t = T0 + wave;
    if (t >= T1) return;
    const int nmax = A.n - 1;
    struct Pre {  // a tile's group: column words and values
        i16x8 w;
        f32x4 v0, v1;
    };
    auto fetch = [&](const WTile &w, Pre &p) {
        const int kk = (w.k0 & ~7) + 8 * lane;
        const int kc = kk < w.k0 + w.nz ? kk : (w.k0 & ~7);  // idle lanes: the first group
        p.w = ld_stream(reinterpret_cast<const i16x8 *>(A.dcol + kc));
        p.v0 = ld_stream(reinterpret_cast<const f32x4 *>(A.val32 + kc));
        p.v1 = ld_stream(reinterpret_cast<const f32x4 *>(A.val32 + kc + 4));
    };
    WTile w = wave_tile(meta, t);
    Pre pn;
    if (w.nz > 0 && w.nz <= kStageNnz) fetch(w, pn);
    for (; t < T1; t += NW) {
        const WTile cur = w;
        const Pre pc = pn;
        const int tn = t + NW;
        if (tn < T1) {
            w = wave_tile(meta, tn);
            if (w.nz > 0 && w.nz <= kStageNnz) fetch(w, pn);
        }
        const int row = cur.r0 + lane;
        const bool has_row = row < cur.r1;
        const int rowc = has_row ? row : cur.r0;
        const int rsa = A.rs[rowc];
        const int rsb = A.rs[rowc + 1];
        const RowState so = ld_state(st_old, rowc);
        double xo = 0.0;
        if (has_pend) xo = A.x[rowc];
        double s = 0.0;
        if (cur.nz == 0) {
            // rows without nonzeros: q = 0
        } else if (cur.nz <= kStageNnz) {
            const int kk = (cur.k0 & ~7) + 8 * lane;
            RowState g[8];
#pragma unroll
            for (int e = 0; e < 8; ++e)
                g[e] = ld_state(st_old, min(max(cur.r0 + (int)pc.w[e], 0), nmax));
            if (kk < cur.k0 + cur.nz) {
                double *dst = lds_wave + (kk - cur.k0);
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float v = e < 4 ? pc.v0[e] : pc.v1[e - 4];
                    dst[e] = (double)v *
                             (double)__builtin_fmaf(beta, g[e].p, next_r(alpha, g[e].r, g[e].q));
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int ra = rsa - cur.k0;
            const int rb = has_row ? rsb - cur.k0 : ra;
            s = row_sum(lds_wave, ra, rb);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            __builtin_amdgcn_wave_barrier();
        } else {
            // a single row longer than a tile: wave-wide sum
            double part = 0.0;
            for (int k = lane; k < cur.nz; k += 64) {
                const RowState g = ld_state(st_old, col_at<int16_t>(A, cur.r0, cur.k0 + k));
                part += (double)A.val32[cur.k0 + k] *
                        (double)__builtin_fmaf(beta, g.p, next_r(alpha, g.r, g.q));
            }
            s = wave_sum(part);
        }
        if (has_row) {
            RowState sn;
            sn.r = next_r(alpha, so.r, so.q);
            sn.p = __builtin_fmaf(beta, so.p, sn.r);
            sn.q = (float)s;
            const double rd = (double)sn.r, qd = (double)sn.q;
            sums[0] += (double)sn.p * s;
            sums[1] += rd * rd;
            sums[2] += rd * qd;
            sums[3] += qd * qd;
            st_state(st_new, row, sn);
            if (has_pend) A.x[row] = __builtin_fma(x_step(A, pend, row), (double)so.p, xo);
        }
    }
}

// x += pend p over the block's rows (p from the packed state)
template <int BS>
__device__ void flush_x1(const CgArgs &A, int R0, int R1, const void *st, double pend) {
    for (int row = R0 + (int)threadIdx.x; row < R1; row += BS)
        A.x[row] = __builtin_fma(x_step(A, pend, row), (double)ld_state(st, row).p, A.x[row]);
}

// Control state of a one-column solve, in LDS: thread 0 updates it after
// every grid reduction and the block reads the step coefficients from it,
// so the bookkeeping holds no registers during the phases.
struct CgCtl1 {
    double tol2, gamma, rmax2, last_rr, ru_rr, rho;
    double pend, alpha, beta, g;
    int iters, status, stall, active, dropped;
    int k, last_ru, refines, ru, has_pend, out_of_iters;
};

// After the phase reduction: alpha_k, gamma_{k+1} and the reliable-update
// trigger.  A drop of the recursive residual by rdelta triggers the update
// once ru_min_gap iterations have passed, a drop by rdelta^2 at once (FP32
// rounding must not pile up over a steep fall); ru_gap iterations and the
// iteration limit force one.
__device__ void ctl_after_phase(CgCtl1 &C, const CgArgs &A, const double (&sums)[4]) {
    const int k = ++C.k;
    double alpha = 0.0, g = 0.0;
    bool dropped = false;
    if (C.active) {
        const double delta = sums[0], gk = sums[1];
        if (!(delta > 0.0)) {
            C.status = NSFEM_BREAKDOWN;
            C.active = 0;
            C.last_rr = gk;
        } else {
            alpha = gk / delta;
            g = fmax(gk - 2.0 * alpha * sums[2] + alpha * alpha * sums[3], 0.0);
            C.gamma = gk;
            C.iters = k;
            const double d2 = A.rdelta * A.rdelta;
            dropped = g * C.rho <= C.tol2 ||
                      (g < d2 * C.rmax2 && (k - C.last_ru >= A.ru_min_gap || g < d2 * d2 * C.rmax2));
        }
    }
    C.pend = alpha;
    C.alpha = alpha;
    C.g = g;
    C.dropped = dropped;
    C.has_pend = 1;
    C.out_of_iters = k >= A.max_iter;
    C.ru = C.active && (dropped || k - C.last_ru >= A.ru_gap || C.out_of_iters);
    if (C.ru) C.last_ru = k;
}

// After a reliable update (x flushed, r := b - A x in FP64): the FP64
// residual decides convergence; after a drop reported by the recursion the
// FP64 residual must fall below (1 + rdelta)/2 of its best value so far, and
// three misses in a row end the FP32 iteration (stagnation; an update
// forced by ru_gap alone is not judged).
__device__ void ctl_after_update(CgCtl1 &C, const CgArgs &A, const double (&ra)[3]) {
    C.has_pend = 0;
    C.pend = 0.0;
    C.alpha = 0.0;
    ++C.refines;
    C.last_rr = ra[0];
    if (ra[0] <= C.tol2) {
        C.active = 0;
    } else {
        C.g = ra[2];
        C.rmax2 = A.scale ? ra[2] : ra[0];
        C.rho = A.scale && ra[2] > 0.0 ? ra[0] / ra[2] : 1.0;
        if (C.dropped) {
            const double sr = 0.5 * (1.0 + A.rdelta);
            C.stall = ra[0] > sr * sr * C.ru_rr ? C.stall + 1 : 0;
        }
        C.ru_rr = fmin(C.ru_rr, ra[0]);
        if (C.stall >= 3) {
            C.status = NSFEM_STAGNATION;
            C.active = 0;
        } else if (C.refines >= A.max_refine) {
            C.status = NSFEM_MAX_ITERATIONS_REACHED;
            C.active = 0;
        }
    }
}

// End of an iteration: iteration limit, beta_{k+1} = gamma_{k+1} / gamma_k.
__device__ void ctl_end(CgCtl1 &C) {
    if (C.active && C.out_of_iters) {
        C.status = NSFEM_MAX_ITERATIONS_REACHED;
        C.active = 0;
        C.last_rr = C.g;
    }
    C.beta = C.active ? C.g / C.gamma : 0.0;
    if (!C.active) C.alpha = 0.0;
}

// The one-column mixed persistent CG run by one block (arguments as
// cg_body).  The recursive residual is replaced by the FP64 b - A x whenever
// it has fallen by rdelta (reliable updates).
template <int BS, typename IT, int WI, typename MP>
__device__ __forceinline__ void cg_body1_m(const CgArgs &A, MP meta, const int bid, const int nb,
                                           const int L, double *prod_raw, double *lds) {
    constexpr bool MIXED = true;  // residual_phase precision
    // WI 1: group loads (16-bit offsets only); else WI items per lane
    constexpr bool GROUPS = WI == 1;
    static_assert(!GROUPS || std::is_same<IT, int16_t>::value, "group loads read 16-bit words");
    __shared__ CgCtl1 C;
    const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
    const int R0 = A.tiles[T0], R1 = A.tiles[T1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // the group-load product writes up to 7 entries before a tile's first
    // nonzero: each wave's slice starts 8 entries in
    double *lds64 = prod_raw + wave * (kWaveNnz + (GROUPS ? 8 : 0)) + (GROUPS ? 8 : 0);
    unsigned epoch = 0;
    const bool t0 = threadIdx.x == 0;
    const bool writer = (L == 0 && t0);
    int cur = 0;  // st[cur]: state of step k - 1

    {
        double acc[3] = {0.0, 0.0, 0.0};
        residual_phase<1, BS, MIXED, true, IT, true, MP>(A, meta, T0, T1, lds64, acc, A.st[0]);
        if (!grid_sum<3, BS>(A.sync, acc, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        if (t0) {
            const double bn = sqrt(acc[1]), r0n = sqrt(acc[0]);
            const double ref = bn > 0.0 ? bn : r0n;
            const double tol = fmax(A.rtol * ref, A.atol);
            C.tol2 = tol * tol;
            C.active = acc[0] > C.tol2;
            C.gamma = acc[2];
            C.rmax2 = C.last_rr = C.ru_rr = acc[0];
            // drop triggers in the iteration's (scaled) space; rho converts
            // a scaled norm into the unscaled one the tolerance is on
            if (A.scale) C.rmax2 = acc[2];
            C.rho = A.scale && acc[2] > 0.0 ? acc[0] / acc[2] : 1.0;
            C.g = 0.0;
            C.iters = 0;
            C.status = NSFEM_OK;
            C.stall = C.dropped = 0;
            C.pend = C.alpha = C.beta = 0.0;
            C.k = C.last_ru = C.refines = C.ru = C.has_pend = C.out_of_iters = 0;
            if (writer) {
                A.res[0].rhs_norm = bn;
                A.res[0].initial_residual = r0n;
                A.res[0].tolerance = tol;
            }
        }
        __syncthreads();
    }
    while (C.active) {
        const int nx = cur ^ 1;
        const float alpha = (float)C.alpha, beta = (float)C.beta;
        const double pend = C.pend;
        const bool has_pend = C.has_pend != 0;
        double sums[4] = {0.0, 0.0, 0.0, 0.0};
        CG_TRACE_STAMP(C.k, 0);
        if constexpr (GROUPS)
            phase_f1g<BS, MP>(A, meta, T0, T1, lds64, A.st[cur], A.st[nx], alpha, beta, pend,
                              has_pend, sums);
        else
            phase_f1<BS, IT, WI, MP>(A, meta, T0, T1, lds64, A.st[cur], A.st[nx], alpha, beta,
                                     pend, has_pend, sums);
        CG_TRACE_MAX(C.k, 1);
        if (!grid_sum<4, BS>(A.sync, sums, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        CG_TRACE_STAMP(C.k, 2);
        if (t0) {
            ctl_after_phase(C, A, sums);
            if (!C.ru) ctl_end(C);
        }
        __syncthreads();
        if (C.ru) {
            // apply the pending x += alpha p_k (FP64), then replace the
            // recursive residual: r_k := b - A x, alpha := 0
            flush_x1<BS>(A, R0, R1, A.st[nx], C.pend);
            double dummy[1] = {0.0};
            if (!grid_sum<1, BS>(A.sync, dummy, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            double ra[3] = {0.0, 0.0, 0.0};
            residual_phase<1, BS, MIXED, false, IT, true, MP>(A, meta, T0, T1, lds64, ra, A.st[nx]);
            if (!grid_sum<3, BS>(A.sync, ra, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            if (t0) {
                ctl_after_update(C, A, ra);
                ctl_end(C);
            }
            __syncthreads();
        }
        if (!C.active) {
            if (C.has_pend) flush_x1<BS>(A, R0, R1, A.st[nx], C.pend);
            break;
        }
        cur = nx;
    }
    if (writer) {
        A.res[0].iterations = C.iters;
        A.res[0].refinements = C.refines;
        A.res[0].status = C.status;
        A.res[0].last_residual = sqrt(C.last_rr);
    }
    return;
timeout:
    if (writer) A.res[0].status = NSFEM_BARRIER_TIMEOUT;
}

template <int BS, typename IT, int WI>
__device__ __forceinline__ void cg_body1(const CgArgs &A, const int bid, const int nb, const int L,
                                         double *prod_raw, double *lds) {
    __shared__ int2 tmeta[kMetaLds + 1];
    const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
    if (T1 > T0 && T1 - T0 <= kMetaLds) {
        load_meta_lds<BS>(A.meta, T0, T1, tmeta);
        cg_body1_m<BS, IT, WI>(A, MetaLds{tmeta, T0}, bid, nb, L, prod_raw, lds);
    } else {
        cg_body1_m<BS, IT, WI>(A, A.meta, bid, nb, L, prod_raw, lds);
    }
}

// ---------------------------------------------------------------------
// Two-column mixed solves: the single-reduction iteration of the one-column
// solves (phase_f1) over both columns at once.  Each nonzero's column and
// value load once for both columns; a neighbour's {r0, r1, p0, p1} and
// {q0, q1} come in two gathers; each column's p_k is recomputed from its
// gathered state by the same two roundings as its owner's, and the FP64
// sums (p_k, q_k), (r_k, r_k), (r_k, q_k), (q_k, q_k) of both columns take
// one grid reduction of 8 values.  sums[4c + i] holds column c's i-th sum.
// IB: nonzeros per lane whose loads are in flight together (4 or 8).
template <int BS, typename IT, int IB, typename MP>
__device__ void phase_f2(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                         const void *st_old, void *st_new, const float (&alpha)[2],
                         const float (&beta)[2], const double (&pend)[2], const bool has_pend,
                         double (&sums)[8]) {
    constexpr int NW = BS / 64;
    constexpr int WI = kWItems;
    const int n = A.n;
    const float *val = A.val32;
    const float4 *rp_old = st2_rp(st_old);
    const float2 *q_old = st2_q(st_old, n);
    float4 *rp_new = st2_rp(st_new);
    float2 *q_new = st2_q(st_new, n);
    const bool pend0 = has_pend && pend[0] != 0.0, pend1 = has_pend && pend[1] != 0.0;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int t = T0 + wave;
    if (t >= T1) return;
    WTile w = wave_tile(meta, t);
    int cn[WI];
    {
        const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
        for (int i = 0; i < WI; ++i) {
            const int k = lane + 64 * i;
            cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
        }
    }
    for (; t < T1; t += NW) {
        const WTile cur = w;
        int cj[WI];
#pragma unroll
        for (int i = 0; i < WI; ++i) cj[i] = col_of<IT>(cn[i], cur.r0);
        const int tn = t + NW;
        if (tn < T1) {
            w = wave_tile(meta, tn);
            const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
            for (int i = 0; i < WI; ++i) {
                const int k = lane + 64 * i;
                cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
            }
        }
        const int row = cur.r0 + lane;
        const bool has_row = row < cur.r1;
        const int rowc = has_row ? row : cur.r0;
        const int rsa = A.rs[rowc];
        const int rsb = A.rs[rowc + 1];
        const float4 so = rp_old[rowc];
        const float2 sq = q_old[rowc];
        // the row's solution entries and scale load with its state, ahead of
        // the gathers (a late load would wait behind the next tile's
        // prefetch)
        double xo0 = 0.0, xo1 = 0.0, scl = 1.0;
        if (pend0) xo0 = A.x[rowc];
        if (pend1) xo1 = A.x[(size_t)n + rowc];
        if (has_pend && A.scale) scl = A.scale[rowc];
        double s[2] = {0.0, 0.0};
        if (cur.nz == 0) {
            // rows without nonzeros: q = 0
        } else if (cur.nz <= 64 * WI) {
#pragma unroll
            for (int i0 = 0; i0 < WI; i0 += IB) {
                float vj[IB];
                float4 gj[IB];
                float2 gq[IB];
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    vj[i] = ld_stream(val + (k < cur.nz ? cur.k0 + k : 0));
                    gj[i] = rp_old[cj[i0 + i]];
                    gq[i] = q_old[cj[i0 + i]];
                }
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    const double v = (double)vj[i];
                    lds_wave[2 * k] =
                        v * (double)__builtin_fmaf(beta[0], gj[i].z, next_r(alpha[0], gj[i].x, gq[i].x));
                    lds_wave[2 * k + 1] =
                        v * (double)__builtin_fmaf(beta[1], gj[i].w, next_r(alpha[1], gj[i].y, gq[i].y));
                }
            }
            __syncwarp();
            const int ra = rsa - cur.k0;
            const int rb = has_row ? rsb - cur.k0 : ra;
            row_sums<2>(lds_wave, ra, rb, s);
            __syncwarp();
        } else {
            // a single row longer than a tile: wave-wide sums
            double part0 = 0.0, part1 = 0.0;
            for (int k = lane; k < cur.nz; k += 64) {
                const int j = col_at<IT>(A, cur.r0, cur.k0 + k);
                const float4 g = rp_old[j];
                const float2 q = q_old[j];
                const double v = (double)val[cur.k0 + k];
                part0 += v * (double)__builtin_fmaf(beta[0], g.z, next_r(alpha[0], g.x, q.x));
                part1 += v * (double)__builtin_fmaf(beta[1], g.w, next_r(alpha[1], g.y, q.y));
            }
            s[0] = wave_sum(part0);
            s[1] = wave_sum(part1);
        }
        if (has_row) {
            const float r0 = next_r(alpha[0], so.x, sq.x), r1 = next_r(alpha[1], so.y, sq.y);
            const float p0 = __builtin_fmaf(beta[0], so.z, r0);
            const float p1 = __builtin_fmaf(beta[1], so.w, r1);
            const float q0 = (float)s[0], q1 = (float)s[1];
            const double r0d = (double)r0, q0d = (double)q0, r1d = (double)r1, q1d = (double)q1;
            sums[0] += (double)p0 * s[0];
            sums[1] += r0d * r0d;
            sums[2] += r0d * q0d;
            sums[3] += q0d * q0d;
            sums[4] += (double)p1 * s[1];
            sums[5] += r1d * r1d;
            sums[6] += r1d * q1d;
            sums[7] += q1d * q1d;
            rp_new[row] = make_float4(r0, r1, p0, p1);
            q_new[row] = make_float2(q0, q1);
            // x_c += x_step(pend_c) p_c, x_step = pend_c * scale[row] (or pend_c)
            if (pend0) A.x[row] = __builtin_fma(pend[0] * scl, (double)so.z, xo0);
            if (pend1) A.x[(size_t)n + row] = __builtin_fma(pend[1] * scl, (double)so.w, xo1);
        }
    }
}

// x_c += pend_c p_c over the block's rows (p from the packed state)
template <int BS>
__device__ void flush_x2(const CgArgs &A, int R0, int R1, const void *st,
                         const double (&pend)[2]) {
    const float4 *rp = st2_rp(st);
    for (int row = R0 + (int)threadIdx.x; row < R1; row += BS) {
        const float4 v = rp[row];
        if (pend[0] != 0.0) A.x[row] = __builtin_fma(x_step(A, pend[0], row), (double)v.z, A.x[row]);
        if (pend[1] != 0.0) {
            double *x = A.x + (size_t)A.n + row;
            *x = __builtin_fma(x_step(A, pend[1], row), (double)v.w, *x);
        }
    }
}

// Control state of a two-column solve (LDS, updated by thread 0): the
// one-column rules per column, with one iteration count and one reliable
// update for both (the FP64 residual pass reads the matrix once for both
// columns; a column that did not ask for the update restarts its recursion
// from the true residual, like one that did).
struct CgCtl2 {
    double tol2[2], gamma[2], rmax2[2], last_rr[2], ru_rr[2], rho[2];
    double pend[2], alpha[2], beta[2], g[2];
    int iters[2], status[2], stall[2], active[2], dropped[2];
    int k, last_ru, refines, ru, has_pend, out_of_iters, any;
};

__device__ void ctl2_after_phase(CgCtl2 &C, const CgArgs &A, const double (&sums)[8]) {
    const int k = ++C.k;
    const double d2 = A.rdelta * A.rdelta;
    bool want = false;
    for (int c = 0; c < 2; ++c) {
        double alpha = 0.0, g = 0.0;
        bool dropped = false;
        if (C.active[c]) {
            const double delta = sums[4 * c], gk = sums[4 * c + 1];
            if (!(delta > 0.0)) {
                C.status[c] = NSFEM_BREAKDOWN;
                C.active[c] = 0;
                C.last_rr[c] = gk;
            } else {
                alpha = gk / delta;
                g = fmax(gk - 2.0 * alpha * sums[4 * c + 2] + alpha * alpha * sums[4 * c + 3], 0.0);
                C.gamma[c] = gk;
                C.iters[c] = k;
                dropped = g * C.rho[c] <= C.tol2[c] ||
                          (g < d2 * C.rmax2[c] &&
                           (k - C.last_ru >= A.ru_min_gap || g < d2 * d2 * C.rmax2[c]));
            }
        }
        want = want || dropped;
        C.pend[c] = alpha;
        C.alpha[c] = alpha;
        C.g[c] = g;
        C.dropped[c] = dropped;
    }
    C.has_pend = 1;
    C.out_of_iters = k >= A.max_iter;
    const bool any = C.active[0] || C.active[1];
    C.ru = any && (want || k - C.last_ru >= A.ru_gap || C.out_of_iters);
    if (C.ru) C.last_ru = k;
}

// After a reliable update; ra as residual_phase<2>: ra[c] = FP64 (r, r),
// ra[4 + c] = (FP32 r, FP32 r).
__device__ void ctl2_after_update(CgCtl2 &C, const CgArgs &A, const double (&ra)[6]) {
    C.has_pend = 0;
    ++C.refines;
    for (int c = 0; c < 2; ++c) {
        C.pend[c] = 0.0;
        C.alpha[c] = 0.0;
        C.last_rr[c] = ra[c];
        if (!C.active[c]) continue;
        if (ra[c] <= C.tol2[c]) {
            C.active[c] = 0;
            continue;
        }
        C.g[c] = ra[4 + c];
        C.rmax2[c] = A.scale ? ra[4 + c] : ra[c];
        C.rho[c] = A.scale && ra[4 + c] > 0.0 ? ra[c] / ra[4 + c] : 1.0;
        if (C.dropped[c]) {
            const double sr = 0.5 * (1.0 + A.rdelta);
            C.stall[c] = ra[c] > sr * sr * C.ru_rr[c] ? C.stall[c] + 1 : 0;
        }
        C.ru_rr[c] = fmin(C.ru_rr[c], ra[c]);
        if (C.stall[c] >= 3) {
            C.status[c] = NSFEM_STAGNATION;
            C.active[c] = 0;
        } else if (C.refines >= A.max_refine) {
            C.status[c] = NSFEM_MAX_ITERATIONS_REACHED;
            C.active[c] = 0;
        }
    }
}

__device__ void ctl2_end(CgCtl2 &C) {
    for (int c = 0; c < 2; ++c) {
        if (C.active[c] && C.out_of_iters) {
            C.status[c] = NSFEM_MAX_ITERATIONS_REACHED;
            C.active[c] = 0;
            C.last_rr[c] = C.g[c];
        }
        C.beta[c] = C.active[c] ? C.g[c] / C.gamma[c] : 0.0;
        if (!C.active[c]) C.alpha[c] = 0.0;
    }
    C.any = C.active[0] || C.active[1];
}

// The two-column mixed persistent CG run by one block (arguments as
// cg_body), single-reduction iteration over the packed state.
template <int BS, typename IT, int IB, typename MP>
__device__ __forceinline__ void cg_body2_m(const CgArgs &A, MP meta, const int bid, const int nb,
                                           const int L, double *prod_raw, double *lds) {
    __shared__ CgCtl2 C;
    const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
    const int R0 = A.tiles[T0], R1 = A.tiles[T1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *lds64 = prod_raw + wave * kWaveNnz * 2;
    unsigned epoch = 0;
    const bool t0 = threadIdx.x == 0;
    const bool writer = (L == 0 && t0);
    int cur = 0;  // st[cur]: state of step k - 1

    {
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        residual_phase<2, BS, true, true, IT, true, MP>(A, meta, T0, T1, lds64, acc, A.st[0]);
        if (!grid_sum<6, BS>(A.sync, acc, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        if (t0) {
            for (int c = 0; c < 2; ++c) {
                const double bn = sqrt(acc[2 + c]), r0n = sqrt(acc[c]);
                const double ref = bn > 0.0 ? bn : r0n;
                const double tol = fmax(A.rtol * ref, A.atol);
                C.tol2[c] = tol * tol;
                C.active[c] = acc[c] > C.tol2[c];
                C.gamma[c] = acc[4 + c];
                C.rmax2[c] = A.scale ? acc[4 + c] : acc[c];
                C.last_rr[c] = C.ru_rr[c] = acc[c];
                C.rho[c] = A.scale && acc[4 + c] > 0.0 ? acc[c] / acc[4 + c] : 1.0;
                C.g[c] = 0.0;
                C.iters[c] = 0;
                C.status[c] = NSFEM_OK;
                C.stall[c] = C.dropped[c] = 0;
                C.pend[c] = C.alpha[c] = C.beta[c] = 0.0;
                if (writer) {
                    A.res[c].rhs_norm = bn;
                    A.res[c].initial_residual = r0n;
                    A.res[c].tolerance = tol;
                }
            }
            C.k = C.last_ru = C.refines = C.ru = C.has_pend = C.out_of_iters = 0;
            C.any = C.active[0] || C.active[1];
        }
        __syncthreads();
    }
    while (C.any) {
        const int nx = cur ^ 1;
        const float alpha[2] = {(float)C.alpha[0], (float)C.alpha[1]};
        const float beta[2] = {(float)C.beta[0], (float)C.beta[1]};
        const double pend[2] = {C.pend[0], C.pend[1]};
        const bool has_pend = C.has_pend != 0;
        double sums[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        CG_TRACE_STAMP(C.k, 0);
        phase_f2<BS, IT, IB, MP>(A, meta, T0, T1, lds64, A.st[cur], A.st[nx], alpha, beta, pend,
                             has_pend, sums);
        CG_TRACE_MAX(C.k, 1);
        if (!grid_sum<8, BS>(A.sync, sums, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        CG_TRACE_STAMP(C.k, 2);
        if (t0) {
            ctl2_after_phase(C, A, sums);
            if (!C.ru) ctl2_end(C);
        }
        __syncthreads();
        if (C.ru) {
            // apply the pending x += alpha p_k (FP64), then replace both
            // recursive residuals: r_k := b - A x, alpha := 0
            const double pf[2] = {C.pend[0], C.pend[1]};
            flush_x2<BS>(A, R0, R1, A.st[nx], pf);
            double dummy[1] = {0.0};
            if (!grid_sum<1, BS>(A.sync, dummy, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            double ra[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            residual_phase<2, BS, true, false, IT, true, MP>(A, meta, T0, T1, lds64, ra, A.st[nx]);
            if (!grid_sum<6, BS>(A.sync, ra, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            if (t0) {
                ctl2_after_update(C, A, ra);
                ctl2_end(C);
            }
            __syncthreads();
        }
        if (!C.any) {
            if (C.has_pend) {
                const double pf[2] = {C.pend[0], C.pend[1]};
                flush_x2<BS>(A, R0, R1, A.st[nx], pf);
            }
            break;
        }
        cur = nx;
    }
    if (writer) {
        for (int c = 0; c < 2; ++c) {
            A.res[c].iterations = C.iters[c];
            A.res[c].refinements = C.refines;
            A.res[c].status = C.status[c];
            A.res[c].last_residual = sqrt(C.last_rr[c]);
        }
    }
    return;
timeout:
    if (writer) {
        for (int c = 0; c < 2; ++c) A.res[c].status = NSFEM_BARRIER_TIMEOUT;
    }
}

// ---------------------------------------------------------------------
// Two-column mixed solves, narrow-gather form: the packed iteration's
// arithmetic with p_k stored once instead of recomputed by every reader.
// After the reduction, a stream pass over the block's own rows forms
//   r_k = r_{k-1} - alpha q_{k-1},  p_k = r_k + beta p_{k-1},
//   x += pend p_{k-1} (FP64),
// a grid barrier follows, and the product q_k = A p_k gathers only the
// neighbours' {p0, p1} (8 bytes, one load per nonzero instead of two) with
// the sums (p_k, q_k), (r_k, r_k), (r_k, q_k), (q_k, q_k) of the packed
// form.  r, p and q are interleaved [n][2] FP32 arrays (one copy each: the
// stream pass touches only its own rows, the product writes only q).  The
// FP32 roundings, products and sums are those of the packed form, so the
// iterates are the same bits.
template <int BS>
__device__ void phase_u2(const CgArgs &A, int R0, int R1, float2 *r, float2 *p, const float2 *q,
                         const float (&alpha)[2], const float (&beta)[2], const double (&pend)[2],
                         const bool has_pend) {
    const bool pend0 = has_pend && pend[0] != 0.0, pend1 = has_pend && pend[1] != 0.0;
    const size_t n = (size_t)A.n;
    for (int row = R0 + (int)threadIdx.x; row < R1; row += BS) {
        const float2 ro = r[row], po = p[row], qo = q[row];
        double x0 = 0.0, x1 = 0.0, scl = 1.0;
        if (pend0) x0 = A.x[row];
        if (pend1) x1 = A.x[n + row];
        if (has_pend && A.scale) scl = A.scale[row];
        const float r0 = next_r(alpha[0], ro.x, qo.x), r1 = next_r(alpha[1], ro.y, qo.y);
        r[row] = make_float2(r0, r1);
        p[row] = make_float2(__builtin_fmaf(beta[0], po.x, r0), __builtin_fmaf(beta[1], po.y, r1));
        if (pend0) A.x[row] = __builtin_fma(pend[0] * scl, (double)po.x, x0);
        if (pend1) A.x[n + row] = __builtin_fma(pend[1] * scl, (double)po.y, x1);
    }
}

template <int BS, typename IT, int IB, typename MP>
__device__ void phase_q2(const CgArgs &A, MP meta, int T0, int T1, double *lds_wave,
                         const float2 *p, const float2 *r, float2 *q, double (&sums)[8]) {
    constexpr int NW = BS / 64;
    constexpr int WI = kWItems;
    const float *val = A.val32;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int t = T0 + wave;
    if (t >= T1) return;
    WTile w = wave_tile(meta, t);
    int cn[WI];
    {
        const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
        for (int i = 0; i < WI; ++i) {
            const int k = lane + 64 * i;
            cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
        }
    }
    for (; t < T1; t += NW) {
        const WTile cur = w;
        int cj[WI];
#pragma unroll
        for (int i = 0; i < WI; ++i) cj[i] = col_of<IT>(cn[i], cur.r0);
        const int tn = t + NW;
        if (tn < T1) {
            w = wave_tile(meta, tn);
            const int kf = w.nz > 0 ? w.k0 : 0;
#pragma unroll
            for (int i = 0; i < WI; ++i) {
                const int k = lane + 64 * i;
                cn[i] = col_word<IT>(A, k < w.nz ? w.k0 + k : kf);
            }
        }
        const int row = cur.r0 + lane;
        const bool has_row = row < cur.r1;
        const int rowc = has_row ? row : cur.r0;
        const int rsa = A.rs[rowc];
        const int rsb = A.rs[rowc + 1];
        const float2 po = p[rowc], ro = r[rowc];
        double s[2] = {0.0, 0.0};
        if (cur.nz == 0) {
            // rows without nonzeros: q = 0
        } else if (cur.nz <= 64 * WI) {
#pragma unroll
            for (int i0 = 0; i0 < WI; i0 += IB) {
                float vj[IB];
                float2 gj[IB];
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    vj[i] = ld_stream(val + (k < cur.nz ? cur.k0 + k : 0));
                    gj[i] = p[cj[i0 + i]];
                }
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const int k = lane + 64 * (i0 + i);
                    const double v = (double)vj[i];
                    lds_wave[2 * k] = v * (double)gj[i].x;
                    lds_wave[2 * k + 1] = v * (double)gj[i].y;
                }
            }
            __syncwarp();
            const int ra = rsa - cur.k0;
            const int rb = has_row ? rsb - cur.k0 : ra;
            row_sums<2>(lds_wave, ra, rb, s);
            __syncwarp();
        } else {
            // a single row longer than a tile: wave-wide sums
            double part0 = 0.0, part1 = 0.0;
            for (int k = lane; k < cur.nz; k += 64) {
                const float2 g = p[col_at<IT>(A, cur.r0, cur.k0 + k)];
                const double v = (double)val[cur.k0 + k];
                part0 += v * (double)g.x;
                part1 += v * (double)g.y;
            }
            s[0] = wave_sum(part0);
            s[1] = wave_sum(part1);
        }
        if (has_row) {
            const float q0 = (float)s[0], q1 = (float)s[1];
            const double r0d = (double)ro.x, q0d = (double)q0, r1d = (double)ro.y, q1d = (double)q1;
            sums[0] += (double)po.x * s[0];
            sums[1] += r0d * r0d;
            sums[2] += r0d * q0d;
            sums[3] += q0d * q0d;
            sums[4] += (double)po.y * s[1];
            sums[5] += r1d * r1d;
            sums[6] += r1d * q1d;
            sums[7] += q1d * q1d;
            q[row] = make_float2(q0, q1);
        }
    }
}

// The narrow-gather two-column mixed CG run by one block (control as
// cg_body2_m; one grid barrier and one reduction an iteration).
template <int BS, typename IT, int IB, typename MP>
__device__ __forceinline__ void cg_body2n_m(const CgArgs &A, MP meta, const int bid, const int nb,
                                            const int L, double *prod_raw, double *lds) {
    __shared__ CgCtl2 C;
    const int T0 = A.block_tiles[L], T1 = A.block_tiles[L + 1];
    const int R0 = A.tiles[T0], R1 = A.tiles[T1];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double *lds64 = prod_raw + wave * kWaveNnz * 2;
    float2 *r = (float2 *)A.r, *p = (float2 *)A.p0, *q = (float2 *)A.q;
    unsigned epoch = 0;
    const bool t0 = threadIdx.x == 0;
    const bool writer = (L == 0 && t0);

    {
        // r = b - A x (FP64, stored FP32), p = 0, q = 0
        double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        residual_phase<2, BS, true, true, IT, false, MP>(A, meta, T0, T1, lds64, acc);
        for (int row = R0 + (int)threadIdx.x; row < R1; row += BS) q[row] = make_float2(0.f, 0.f);
        if (!grid_sum<6, BS>(A.sync, acc, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        if (t0) {
            for (int c = 0; c < 2; ++c) {
                const double bn = sqrt(acc[2 + c]), r0n = sqrt(acc[c]);
                const double ref = bn > 0.0 ? bn : r0n;
                const double tol = fmax(A.rtol * ref, A.atol);
                C.tol2[c] = tol * tol;
                C.active[c] = acc[c] > C.tol2[c];
                C.gamma[c] = acc[4 + c];
                C.rmax2[c] = A.scale ? acc[4 + c] : acc[c];
                C.last_rr[c] = C.ru_rr[c] = acc[c];
                C.rho[c] = A.scale && acc[4 + c] > 0.0 ? acc[c] / acc[4 + c] : 1.0;
                C.g[c] = 0.0;
                C.iters[c] = 0;
                C.status[c] = NSFEM_OK;
                C.stall[c] = C.dropped[c] = 0;
                C.pend[c] = C.alpha[c] = C.beta[c] = 0.0;
                if (writer) {
                    A.res[c].rhs_norm = bn;
                    A.res[c].initial_residual = r0n;
                    A.res[c].tolerance = tol;
                }
            }
            C.k = C.last_ru = C.refines = C.ru = C.has_pend = C.out_of_iters = 0;
            C.any = C.active[0] || C.active[1];
        }
        __syncthreads();
    }
    while (C.any) {
        {
            const float alpha[2] = {(float)C.alpha[0], (float)C.alpha[1]};
            const float beta[2] = {(float)C.beta[0], (float)C.beta[1]};
            const double pend[2] = {C.pend[0], C.pend[1]};
            phase_u2<BS>(A, R0, R1, r, p, q, alpha, beta, pend, C.has_pend != 0);
        }
        double dummy[1] = {0.0};
        if (!grid_sum<1, BS>(A.sync, dummy, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        double sums[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        phase_q2<BS, IT, IB, MP>(A, meta, T0, T1, lds64, p, r, q, sums);
        if (!grid_sum<8, BS>(A.sync, sums, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
        if (t0) {
            ctl2_after_phase(C, A, sums);
            if (!C.ru) ctl2_end(C);
        }
        __syncthreads();
        if (C.ru) {
            // apply the pending x += alpha p_k (FP64), then replace both
            // recursive residuals: r_k := b - A x, alpha := 0
            const double pf[2] = {C.pend[0], C.pend[1]};
            flush_x<2, BS, float>(A, R0, R1, (const float *)p, pf);
            if (!grid_sum<1, BS>(A.sync, dummy, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            double ra[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            residual_phase<2, BS, true, false, IT, false, MP>(A, meta, T0, T1, lds64, ra);
            if (!grid_sum<6, BS>(A.sync, ra, epoch, lds, A.timeout_ticks, bid, nb, A.flat)) goto timeout;
            if (t0) {
                ctl2_after_update(C, A, ra);
                ctl2_end(C);
            }
            __syncthreads();
        }
        if (!C.any) {
            if (C.has_pend) {
                const double pf[2] = {C.pend[0], C.pend[1]};
                flush_x<2, BS, float>(A, R0, R1, (const float *)p, pf);
            }
            break;
        }
    }
    if (writer) {
        for (int c = 0; c < 2; ++c) {
            A.res[c].iterations = C.iters[c];
            A.res[c].refinements = C.refines;
            A.res[c].status = C.status[c];
            A.res[c].last_residual = sqrt(C.last_rr[c]);
        }
    }
    return;
timeout:
    if (writer) {
        for (int c = 0; c < 2; ++c) A.res[c].status = NSFEM_BARRIER_TIMEOUT;
    }
}

// One block per CU: BS / 256 waves per SIMD (4 for 1024 threads, 128 VGPRs
// per wave; 2 for 512 threads, up to 256 VGPRs).
// WI (items per lane of a tile) < 8 is used by the one-column mixed solve only.
template <int NC, int BS, bool MIXED, typename IT, int WI>
__global__ __launch_bounds__(BS, BS / 256) void k_cg(const CgArgs A) {
    constexpr int NW = BS / 64;
    __shared__ double prod_raw[NW * (kWaveNnz * NC + (WI == 1 ? 8 : 0))];
    __shared__ double lds[kMaxRed * NW];
    const int L = xcd_swizzle(blockIdx.x, gridDim.x);
    if constexpr (NC == 1 && MIXED)
        cg_body1<BS, IT, WI>(A, blockIdx.x, gridDim.x, L, prod_raw, lds);
    else
        cg_body<NC, BS, MIXED, IT>(A, blockIdx.x, gridDim.x, L, prod_raw, lds);
}

// The two-column mixed solve's packed iteration: a kernel of its own, so
// that its registers are allocated for it alone.  (The tile descriptors
// stay in global memory: a copy in LDS beside the product slices measured
// no faster, 2,495 against 2,525 GB/s in the 2048^2 cavity's mass solve.)
template <int BS, typename IT, int IB>
__global__ __launch_bounds__(BS, BS / 256) void k_cg_p2(const CgArgs A) {
    constexpr int NW = BS / 64;
    __shared__ double prod_raw[NW * kWaveNnz * 2];
    __shared__ double lds[kMaxRed * NW];
    const int L = xcd_swizzle(blockIdx.x, gridDim.x);
    if (A.packed2 == 2)
        cg_body2n_m<BS, IT, IB>(A, A.meta, blockIdx.x, gridDim.x, L, prod_raw, lds);
    else
        cg_body2_m<BS, IT, IB>(A, A.meta, blockIdx.x, gridDim.x, L, prod_raw, lds);
}

// Many independent systems in one persistent launch: block i works for
// instance map[i].x as its block map[i].y of args[instance].nblocks; each
// instance reduces over its own blocks with its own SyncState, and its blocks
// leave the kernel when it converges.
template <int NC, int BS, bool MIXED, typename IT, int WI>
__global__ __launch_bounds__(BS, BS / 256) void k_cg_batch(const CgArgs *__restrict__ args,
                                                           const int2 *__restrict__ map) {
    constexpr int NW = BS / 64;
    __shared__ double prod_raw[NW * (kWaveNnz * NC + (WI == 1 ? 8 : 0))];
    __shared__ double lds[kMaxRed * NW];
    const int2 m = map[blockIdx.x];
    const int inst = __builtin_amdgcn_readfirstlane(m.x);
    const int lb = __builtin_amdgcn_readfirstlane(m.y);
    const CgArgs A = args[inst];
    if constexpr (NC == 1 && MIXED)
        cg_body1<BS, IT, WI>(A, lb, A.nblocks, lb, prod_raw, lds);
    else
        cg_body<NC, BS, MIXED, IT>(A, lb, A.nblocks, lb, prod_raw, lds);
}

constexpr int kCgBlock = 1024;
// One-column mixed solves may run 768-thread blocks: 12 waves per CU with
// 168 VGPRs each instead of 16 with 128, which the tile loop needs to hold
// a prefetched tile without spilling (a scratch reload inside the loop waits
// for every load in flight, the prefetch included).
constexpr int kCgBlockWide = 768;

// Block size of a solve: one-column mixed solves run the wide-register
// blocks unless NSFEM_CG_BLOCK=1024 (read at every solve; measured at the
// 2048^2 cavity, 768 threads with group loads: pressure 95.9 us and v
// diffusion 550 us per iteration against 98.8 / 564 us for 1024 threads
// with one load per nonzero per lane).  The two-column mixed solve was
// slower in 768-thread blocks (mass 1,195 against 1,044 us per iteration).
int cg_block(int nc, bool mixed) {
    if (!(nc == 1 && mixed)) return kCgBlock;
    const char *e = std::getenv("NSFEM_CG_BLOCK");
    return (e && std::atoi(e) == kCgBlock) ? kCgBlock : kCgBlockWide;
}

template <int NC, bool MIXED, typename IT, int WI, int BS>
void cg_launch(Device &d, const CgArgs &args, int grid) {
    void *kargs[] = {(void *)&args};
    NSFEM_HIP(hipLaunchCooperativeKernel((const void *)k_cg<NC, BS, MIXED, IT, WI>, dim3(grid),
                                         dim3(BS), kargs, 0, d.stream));
}
template <int NC, bool MIXED, typename IT, int WI, int BS>
int cg_max_grid(Device &d) {
    int per_cu = 0;
    NSFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg<NC, BS, MIXED, IT, WI>,
                                                           BS, 0));
    return std::max(1, std::min(per_cu * d.cu_count, kMaxGridBlocks));
}

// Calls f.template run<NC, MIXED, IT, WI>() for the solve's column count,
// precision, column-index width and (one-column mixed solves) items per lane.
template <int WI, typename F>
void cg_dispatch_wi(int nc, bool mixed, bool idx16, F &f) {
    if (nc == 1) {
        if (mixed) {
            if (idx16) f.template run<1, true, int16_t, WI>();
            else f.template run<1, true, int32_t, WI>();
        } else {
            if (idx16) f.template run<1, false, int16_t, kWItems>();
            else f.template run<1, false, int32_t, kWItems>();
        }
    } else {
        if (mixed) {
            if (idx16) f.template run<2, true, int16_t, kWItems>();
            else f.template run<2, true, int32_t, kWItems>();
        } else {
            if (idx16) f.template run<2, false, int16_t, kWItems>();
            else f.template run<2, false, int32_t, kWItems>();
        }
    }
}
// wi 1: the one-column mixed iteration with group loads (16-bit offsets)
template <typename F>
void cg_dispatch(int nc, bool mixed, bool idx16, int wi, F &&f) {
    if (wi == 1 && nc == 1 && mixed && idx16) f.template run<1, true, int16_t, 1>();
    else if (wi <= 4) cg_dispatch_wi<4>(nc, mixed, idx16, f);
    else if (wi <= 7) cg_dispatch_wi<7>(nc, mixed, idx16, f);
    else cg_dispatch_wi<kWItems>(nc, mixed, idx16, f);
}

// Items per lane the one-column mixed iteration needs for a matrix: every
// multi-row wave tile's nonzeros over 64, rounded up to 4, 7 or 8.
int items_per_lane(const Matrix &m) {
    const int need = (m.max_multirow_nz + 63) / 64;
    return need <= 4 ? 4 : need <= 7 ? 7 : kWItems;
}

// Form of the one-column mixed iteration on matrices with 16-bit offsets:
// NSFEM_CG_PHASE = items (one load per nonzero per lane) or groups (8
// nonzeros per lane); without it, groups where `groups_default` (the
// single solves' wide blocks) and the matrix allows them.  Read at every
// solve (comparison runs switch it in one process).
int phase_form(const Matrix &m, bool groups_default) {
    const char *e = std::getenv("NSFEM_CG_PHASE");
    const std::string v = e ? e : "";
    if (v == "items") return -1;
    if (v == "groups" || (v.empty() && groups_default)) return m.groups_ok ? 1 : -1;
    return -1;
}

template <typename IT, int BS, int IB>
void cg_launch_p2(Device &d, const CgArgs &args, int grid) {
    void *kargs[] = {(void *)&args};
    NSFEM_HIP(hipLaunchCooperativeKernel((const void *)k_cg_p2<BS, IT, IB>, dim3(grid), dim3(BS),
                                         kargs, 0, d.stream));
}
template <typename IT, int BS, int IB>
int cg_max_grid_p2(Device &d) {
    int per_cu = 0;
    NSFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_cg_p2<BS, IT, IB>, BS, 0));
    return std::max(1, std::min(per_cu * d.cu_count, kMaxGridBlocks));
}
// Calls f.template run<IT, BS, IB>() for the packed two-column iteration's
// block size (768 or 1024) and loads in flight (4 or 8 nonzeros per lane).
template <typename IT, typename F>
void p2_dispatch(int bs, int ib, F &f) {
    if (bs == kCgBlockWide) {
        if (ib == 8) f.template run2<IT, kCgBlockWide, 8>();
        else f.template run2<IT, kCgBlockWide, 4>();
    } else {
        if (ib == 8) f.template run2<IT, kCgBlock, 8>();
        else f.template run2<IT, kCgBlock, 4>();
    }
}

// p2: the two-column mixed solve's packed iteration (k_cg_p2)
struct CgMaxGrid {
    Device &d;
    int out;
    int bs = kCgBlock;
    bool p2 = false;
    int ib = 4;
    template <typename IT, int BS, int IB>
    void run2() {
        out = cg_max_grid_p2<IT, BS, IB>(d);
    }
    template <int NC, bool MX, typename IT, int WI>
    void run() {
        if constexpr (NC == 2 && MX) {
            if (p2) {
                p2_dispatch<IT>(bs, ib, *this);
                return;
            }
        }
        if constexpr (NC == 1 && MX) {
            if (bs == kCgBlockWide) {
                out = cg_max_grid<NC, MX, IT, WI, kCgBlockWide>(d);
                return;
            }
        }
        out = cg_max_grid<NC, MX, IT, WI, kCgBlock>(d);
    }
};
struct CgLaunch {
    Device &d;
    const CgArgs &a;
    int grid;
    int bs = kCgBlock;
    bool p2 = false;
    int ib = 4;
    template <typename IT, int BS, int IB>
    void run2() {
        cg_launch_p2<IT, BS, IB>(d, a, grid);
    }
    template <int NC, bool MX, typename IT, int WI>
    void run() {
        if constexpr (NC == 2 && MX) {
            if (p2) {
                p2_dispatch<IT>(bs, ib, *this);
                return;
            }
        }
        if constexpr (NC == 1 && MX) {
            if (bs == kCgBlockWide) {
                cg_launch<NC, MX, IT, WI, kCgBlockWide>(d, a, grid);
                return;
            }
        }
        cg_launch<NC, MX, IT, WI, kCgBlock>(d, a, grid);
    }
};


}  // namespace


namespace {
// Device memory one persistent solve works in: six vectors of `vec_bytes`
// (r, p_old, p_new, q or the two packed {r, p, q} states), the grid
// reduction's state, the per-column results and the dot product's scratch.
struct CgWorkspace {
    unsigned char *vectors = nullptr;
    size_t vec_bytes = 0;
    SyncState *sync = nullptr;
    CgColumnResult *res = nullptr;  // [2]
    double *dot_partial = nullptr;  // [kDotBlocks]
    double *scalars = nullptr;      // [2]
};

// true residual ||b_c - A x_c|| for each column (FP64), into host res[c]
void true_residuals(Device &d, Matrix &m, const double *b, const double *x, int nc,
                    double *tmp, const CgWorkspace &ws, double *out) {
    launch_fill(d, nc, 0.0, ws.scalars);
    for (int c = 0; c < nc; ++c) {
        launch_spmv(d, m, x + (size_t)c * m.rows, b + (size_t)c * m.rows, tmp, 1);
        dot_accumulate(d, m.rows, tmp, tmp, ws.scalars + c, ws.dot_partial);
    }
    double h[2] = {0, 0};
    NSFEM_HIP(hipMemcpyAsync(h, ws.scalars, sizeof(double) * nc, hipMemcpyDeviceToHost,
                             d.stream));
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    for (int c = 0; c < nc; ++c) out[c] = std::sqrt(h[c]);
}

// Bytes of one CG vector of a solve (rounded to 256).
size_t cg_vector_bytes(int n, int nc, bool mixed) {
    const size_t vt = mixed ? sizeof(float) : sizeof(double);
    return ((size_t)n * nc * vt + 255) / 256 * 256;
}

// The workspace in the device's own arrays, grown as needed.
CgWorkspace device_workspace(Device &d, int n, int nc, bool mixed) {
    CgWorkspace ws;
    ws.vec_bytes = cg_vector_bytes(n, nc, mixed);
    d.work.ensure(ws.vec_bytes * 6);
    d.sync_mem.ensure(sizeof(SyncState) + 256);
    d.cg_res.ensure(2);
    d.dot_partial.ensure(kDotBlocks);
    d.scalars.ensure(4);
    ws.vectors = d.work.ptr;
    unsigned char *sm = d.sync_mem.ptr;
    ws.sync = reinterpret_cast<SyncState *>(sm + (256 - ((uintptr_t)sm % 256)) % 256);
    ws.res = d.cg_res.ptr;
    ws.dot_partial = d.dot_partial.ptr;
    ws.scalars = d.scalars.ptr;
    return ws;
}
}  // namespace

namespace {

// Tolerances, iteration limits and reliable-update knobs of one solve.
void set_cg_params(const Device &d, const CgRequest &req, int n, CgArgs &a) {
    a.rtol = req.rtol;
    a.atol = req.atol;
    // reliable updates of the mixed solver: when the recursive residual has
    // fallen by rdelta since the last one, or after ru_gap iterations.
    // Measured on the 2048^2 cavity pressure matrix from x0 = 0: forced
    // updates every 1000 iterations cost 26% more iterations, and so does
    // holding updates back for 8 iterations after the last one (the default
    // minimum gap is therefore 0; it saves time only on systems that
    // converge in a few dozen iterations).  Tuning overrides:
    // NSFEM_CG_RDELTA, NSFEM_CG_RU_GAP, NSFEM_CG_RU_MIN_GAP.
    static const double rdelta_env = [] {
        const char *e = std::getenv("NSFEM_CG_RDELTA");
        return e ? std::atof(e) : 0.0;
    }();
    static const int ru_gap_env = [] {
        const char *e = std::getenv("NSFEM_CG_RU_GAP");
        return e ? std::atoi(e) : 0;
    }();
    a.rdelta = rdelta_env > 0.0 && rdelta_env < 1.0 ? rdelta_env : req.reliable_delta;
    a.max_iter = req.max_iter < 0 ? n : req.max_iter;
    a.max_refine = 400;
    static const int ru_min_gap_env = [] {
        const char *e = std::getenv("NSFEM_CG_RU_MIN_GAP");
        return e ? std::atoi(e) : -1;
    }();
    a.ru_gap = ru_gap_env > 0 ? ru_gap_env : (req.ru_gap > 0 ? req.ru_gap : 20000);
    a.ru_min_gap = ru_min_gap_env >= 0 ? ru_min_gap_env : 0;
    a.timeout_ticks = (unsigned long long)d.wall_khz * 1000ull * 20ull;  // 20 s per barrier
    // the flat grid reduction unless NSFEM_CG_FLAT=0 (read at every solve)
    const char *flat = std::getenv("NSFEM_CG_FLAT");
    a.flat = !(flat && std::atoi(flat) == 0);
}

// Packed FP32 {r, p, q} rows of a one-column mixed solve: two copies of 3
// vectors of `vb` bytes (a vector holds n values of the solve's precision).
void set_state_buffers(CgArgs &a, unsigned char *w, size_t vb) {
    a.st[0] = w;
    a.st[1] = w + 3 * vb;
}

// Iteration of two-column mixed solves: the single-reduction iteration
// (k_cg_p2) unless NSFEM_CG_PAIRED=twophase, in its narrow-gather form
// unless NSFEM_CG_PAIRED=packed (read at every solve; comparison runs switch
// it in one process).  Measured on the 2048^2 cavity's mass solve (textbook
// bytes / time): narrow 2,581 (4 loads in flight) and 2,603 GB/s (8),
// packed 2,521, two-phase 2,463.  The batched solves run the two-phase form.
bool paired_packed() {
    const char *e = std::getenv("NSFEM_CG_PAIRED");
    return !(e && std::string(e) == "twophase");
}
// 2: the narrow-gather form, 1: the packed form.
int paired_form() {
    const char *e = std::getenv("NSFEM_CG_PAIRED");
    return (e && std::string(e) == "packed") ? 1 : 2;
}
// Block size of the packed two-column iteration: NSFEM_CG_P2_BLOCK=768 or
// 1024 (read at every solve).
int paired_block() {
    const char *e = std::getenv("NSFEM_CG_P2_BLOCK");
    return (e && std::atoi(e) == kCgBlockWide) ? kCgBlockWide : kCgBlock;
}
// Nonzeros per lane whose loads the two-column iteration keeps in flight
// together: NSFEM_CG_P2_LOADS=4 or 8 (read at every solve); by default 8 for
// the narrow-gather form (one 8-byte gather a nonzero), 4 for the packed
// one (two gathers a nonzero; 8 measured 2,497 against 2,527 GB/s).
int paired_loads(int form) {
    const char *e = std::getenv("NSFEM_CG_P2_LOADS");
    if (e && (std::atoi(e) == 4 || std::atoi(e) == 8)) return std::atoi(e);
    return form == 2 ? 8 : 4;
}

struct CgPass {
    CgColumnResult res[2];
    int iters[2] = {0, 0}, refines[2] = {0, 0};
    double final_res[2] = {0, 0};
    float ms = 0.f;
};

// One persistent-CG solve in the given precision.  The FP64 path ends on the
// recursive residual; the true residual is recomputed and, should rounding
// drift have left it above the tolerance, the solve restarts from the
// current x (at most twice).  `fixed`: a workspace to work in (the
// no-allocation solve), else the device's own arrays.
void cg_pass(Device &d, const CgRequest &req, bool mixed, CgPass &P,
             const CgWorkspace *fixed = nullptr) {
    Matrix &m = *req.m;
    const int n = m.rows, nc = req.ncol;
    if (mixed) ensure_val32(d, m);
    const bool jacobi = mixed && req.jacobi && ensure_jacobi(d, m);
    // 1024 threads per block (16 waves per CU, 128 VGPRs each) for FP64
    // and two-column solves: the loop is bound by gather latency, so waves
    // in flight matter more than registers; 512 threads (8 waves, up to 256
    // VGPRs) was slower on every measured solve.  One-column mixed solves
    // run 768 (cg_block).
    const bool idx16 = m.dcol_ok && want_idx16();
    const bool p2 = nc == 2 && mixed && paired_packed();
    const int bs = p2 ? paired_block() : cg_block(nc, mixed);
    const int form = idx16 && mixed && nc == 1 ? phase_form(m, bs == kCgBlockWide) : -1;
    const int wi = form >= 0 ? form : items_per_lane(m);
    const int ib = p2 ? paired_loads(paired_form()) : 4;
    CgMaxGrid mg{d, 0, bs, p2, ib};
    cg_dispatch(nc, mixed, idx16, wi, mg);
    const int maxg = mg.out;
    int want = req.grid;
    if (want <= 0) {
        // tuning override: NSFEM_CG_GRID=<blocks> (capped by co-residency)
        const char *env = std::getenv("NSFEM_CG_GRID");
        if (env) want = std::atoi(env);
    }
    int grid = want > 0 ? std::min(want, maxg)
                        : std::min(maxg, std::max(1, m.n_wtiles / (2 * (bs / 64))));
    grid = std::max(1, std::min(grid, std::max(1, m.n_wtiles)));
    const bool verbose = std::getenv("NSFEM_CG_VERBOSE") != nullptr;
    ensure_block_tiles(d, m, grid);

    const CgWorkspace ws = fixed ? *fixed : device_workspace(d, n, nc, mixed);
    const size_t vec_bytes = cg_vector_bytes(n, nc, mixed);
    NSFEM_REQUIRE(vec_bytes <= ws.vec_bytes, NSFEM_ERROR, "CG workspace too small");

    CgArgs a{};
    a.n = n;
    a.meta = m.wmeta.ptr;
    a.tiles = m.wtiles.ptr;
    a.block_tiles = m.block_tiles.ptr;
    a.rs = m.row_start.ptr;
    a.col = m.col.ptr;
    a.dcol = m.dcol.ptr;
    a.val64 = m.val.ptr;
    a.val32 = jacobi ? m.jval32.ptr : m.val32.ptr;
    a.scale = jacobi ? m.jscale.ptr : nullptr;
    a.b = req.b;
    a.x = req.x;
    unsigned char *w = ws.vectors;
    a.r = w;
    a.p0 = w + vec_bytes;
    a.p1 = w + 2 * vec_bytes;
    a.q = w + 3 * vec_bytes;
    set_state_buffers(a, w, vec_bytes);
    a.packed2 = p2 ? paired_form() : 0;
    set_cg_params(d, req, n, a);
    a.sync = ws.sync;
    a.res = ws.res;
    const int ngroups = (grid + kGroup - 1) / kGroup;

    for (int attempt = 0; attempt < 3; ++attempt) {
        if (verbose)  // one line per launch, in launch order
            std::fprintf(stderr,
                         "[nsfem] cg launch matrix=%p rows=%d cols=%d wave tiles=%d grid=%d "
                         "maxgrid=%d mixed=%d idx16=%d form=%d block=%d attempt=%d\n",
                         (const void *)&m, n, nc, m.n_wtiles, grid, maxg, (int)mixed, (int)idx16,
                         wi, bs, attempt);
        NSFEM_HIP(hipEventRecord(d.ev_start, d.stream));
        k_sync_reset<<<(ngroups + 255) / 256, 256, 0, d.stream>>>(a.sync, ngroups);
        NSFEM_HIP(hipGetLastError());
        CgLaunch ln{d, a, grid, bs, p2, ib};
        cg_dispatch(nc, mixed, idx16, wi, ln);
        NSFEM_HIP(hipEventRecord(d.ev_stop, d.stream));
        NSFEM_HIP(hipMemcpyAsync(P.res, ws.res, sizeof(CgColumnResult) * nc,
                                 hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        float ms = 0.f;
        NSFEM_HIP(hipEventElapsedTime(&ms, d.ev_start, d.ev_stop));
        P.ms += ms;
        bool drift = false;
        if (!mixed) true_residuals(d, m, req.b, req.x, nc, (double *)a.q, ws, P.final_res);
        for (int c = 0; c < nc; ++c) {
            P.iters[c] += P.res[c].iterations;
            P.refines[c] += P.res[c].refinements;
            if (mixed) P.final_res[c] = P.res[c].last_residual;
            if (P.res[c].status == NSFEM_OK && P.final_res[c] > P.res[c].tolerance * (1.0 + 1e-12))
                drift = true;
            if (P.res[c].status == NSFEM_BARRIER_TIMEOUT)
                fail(NSFEM_BARRIER_TIMEOUT, "CG kernel barrier timed out");
        }
        if (!drift || mixed) break;
    }
}

}  // namespace

namespace {
int run_cg_in(Device &d, const CgRequest &req, nsfem_cg_stats *stats,
              const CgWorkspace *fixed) {
    Matrix &m = *req.m;
    NSFEM_REQUIRE(m.ready, NSFEM_NOT_READY, "matrix slot is empty");
    NSFEM_REQUIRE(m.rows == m.cols, NSFEM_INVALID_ARGUMENT, "CG needs a square matrix");
    NSFEM_REQUIRE(req.ncol == 1 || req.ncol == 2, NSFEM_INVALID_ARGUMENT,
                  "columns must be 1 or 2");
    const bool mixed = req.precision == NSFEM_PRECISION_FP32_REFINED;
    NSFEM_REQUIRE(mixed || req.precision == NSFEM_PRECISION_FP64, NSFEM_INVALID_ARGUMENT,
                  "unknown precision");
    const int nc = req.ncol;
    CgPass P;
    cg_pass(d, req, mixed, P, fixed);
    CgPass F;  // FP64 continuation after an FP32 stagnation
    bool fallback = false;
    if (mixed) {
        for (int c = 0; c < nc; ++c) fallback = fallback || P.res[c].status == NSFEM_STAGNATION;
        if (fallback) {
            CgRequest r2 = req;
            if (r2.max_iter > 0) r2.max_iter = std::max(1, r2.max_iter - std::max(P.iters[0], P.iters[nc - 1]));
            // the same stopping rule: ||b|| (or ||r0||) of the original solve
            r2.rtol = 0.0;
            r2.atol = std::max(P.res[0].tolerance, nc == 2 ? P.res[1].tolerance : 0.0);
            if (nc == 2) r2.atol = std::min(P.res[0].tolerance, P.res[1].tolerance);
            cg_pass(d, r2, false, F, fixed);
        }
    }
    int worst = NSFEM_OK;
    for (int c = 0; c < nc; ++c) {
        const int status = fallback ? F.res[c].status : P.res[c].status;
        if (status != NSFEM_OK && worst == NSFEM_OK) worst = status;
        if (stats) {
            nsfem_cg_stats &s = stats[c];
            s.iterations = P.iters[c] + (fallback ? F.iters[c] : 0);
            s.refinements = P.refines[c] + (fallback ? 1 : 0);
            s.status = status;
            s.columns = nc;
            s.rhs_norm = P.res[c].rhs_norm;
            s.initial_residual = P.res[c].initial_residual;
            s.final_residual = fallback ? F.final_res[c] : P.final_res[c];
            s.tolerance = P.res[c].tolerance;
            s.seconds = (P.ms + (fallback ? F.ms : 0.f)) * 1e-3;
        }
    }
    return worst;
}
}  // namespace

int run_cg(Device &d, const CgRequest &req, nsfem_cg_stats *stats) {
    return run_cg_in(d, req, stats, nullptr);
}

// ------------------------------------------------------ no-allocation CG
// The device memory of the no-allocation solve: one static array per GPU
// (each device the library runs on holds its own copy until the process
// ends).  Host code never uses this name as an address: scratch_init looks
// up the device address once per device handle, and every kernel argument
// and copy uses that address.  Layout: the reduction state, the per-column
// results, the dot product's partials and scalars, then six CG vectors of
// kNoAllocVecBytes (an FP64 system of rows x columns <= kNoAllocMaxValues;
// an FP32 + refinement solve needs the same room for its FP64 fallback).
namespace {
constexpr size_t kNoAllocVecBytes = (size_t)kNoAllocMaxValues * sizeof(double);
constexpr size_t kScratchSync = 0;
constexpr size_t kScratchRes = kScratchSync + (sizeof(SyncState) + 255) / 256 * 256;
constexpr size_t kScratchDot = kScratchRes + 256;
constexpr size_t kScratchScal = kScratchDot + kDotBlocks * sizeof(double);
constexpr size_t kScratchVec = kScratchScal + 256;
constexpr size_t kScratchBytes = kScratchVec + 6 * kNoAllocVecBytes;
static_assert(2 * sizeof(CgColumnResult) <= 256, "results slot");
static_assert(sizeof(SyncState) % 256 == 0, "reduction state alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];
}  // namespace

void scratch_init(Device &d) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        fail(NSFEM_HIP_ERROR, std::string("hipGetSymbolAddress(scratch): ") +
                                  hipGetErrorString(e));
    d.scratch = static_cast<unsigned char *>(ptr);
}

int run_cg_noalloc(Device &d, const CgRequest &req, nsfem_cg_stats *stats) {
    NSFEM_REQUIRE(d.scratch, NSFEM_NOT_READY, "device scratch not set up");
    NSFEM_REQUIRE(req.m && req.m->ready, NSFEM_NOT_READY, "matrix slot is empty");
    const int64_t values = (int64_t)req.m->rows * req.ncol;
    NSFEM_REQUIRE(values <= kNoAllocMaxValues, NSFEM_INVALID_ARGUMENT,
                  "no-allocation CG: rows x columns = " + std::to_string(values) +
                      " exceeds the scratch capacity of " + std::to_string(kNoAllocMaxValues) +
                      " values; use nsfem_cg_solve");
    NSFEM_REQUIRE(req.precision != NSFEM_PRECISION_FP32_REFINED || req.m->val32.ptr,
                  NSFEM_NOT_READY, "the matrix has no FP32 copy");
    CgWorkspace ws;
    ws.sync = reinterpret_cast<SyncState *>(d.scratch + kScratchSync);
    ws.res = reinterpret_cast<CgColumnResult *>(d.scratch + kScratchRes);
    ws.dot_partial = reinterpret_cast<double *>(d.scratch + kScratchDot);
    ws.scalars = reinterpret_cast<double *>(d.scratch + kScratchScal);
    ws.vectors = d.scratch + kScratchVec;
    ws.vec_bytes = kNoAllocVecBytes;
    return run_cg_in(d, req, stats, &ws);
}

// ------------------------------------------------------------ batched CG
namespace {

template <int NC, bool MIXED, typename IT, int WI>
void cg_batch_launch(Device &d, const CgArgs *args, const int2 *map, int grid) {
    void *kargs[] = {(void *)&args, (void *)&map};
    NSFEM_HIP(hipLaunchCooperativeKernel((const void *)k_cg_batch<NC, kCgBlock, MIXED, IT, WI>,
                                         dim3(grid), dim3(kCgBlock), kargs, 0, d.stream));
}

template <int NC, bool MIXED, typename IT, int WI>
int cg_batch_max_grid(Device &d) {
    int per_cu = 0;
    NSFEM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, k_cg_batch<NC, kCgBlock, MIXED, IT, WI>, kCgBlock, 0));
    return std::max(1, std::min(per_cu * d.cu_count, kMaxGridBlocks));
}

struct CgBatchMaxGrid {
    Device &d;
    int out;
    template <int NC, bool MX, typename IT, int WI>
    void run() { out = cg_batch_max_grid<NC, MX, IT, WI>(d); }
};
struct CgBatchLaunch {
    Device &d;
    const CgArgs *args;
    const int2 *map;
    int grid;
    template <int NC, bool MX, typename IT, int WI>
    void run() { cg_batch_launch<NC, MX, IT, WI>(d, args, map, grid); }
};

inline size_t round256(size_t b) { return (b + 255) / 256 * 256; }

// Blocks per instance, at most `maxg` in total.  Alone, an instance would get
// one block per 32 wave tiles (`scale` times that, its cap here).  When the
// batch wants more blocks than the device holds, every instance starts with
// one and the rest go, one at a time, to the instance with the largest
// expected time: wave tiles x iterations / blocks, with the iterations of a
// Poisson-type solve growing like the mesh width (tiles^0.5).  So the
// largest system, which needs the most iterations, does not become the
// batch's critical path on a fair share.
std::vector<int> split_blocks(const std::vector<const CgRequest *> &rq, int maxg, int scale) {
    const size_t I = rq.size();
    std::vector<int> want(I), nb(I, 1);
    std::vector<double> work(I);
    int64_t total = 0;
    for (size_t i = 0; i < I; ++i) {
        const int tiles = std::max(1, rq[i]->m->n_wtiles);
        want[i] = std::min(tiles, std::max(1, tiles * scale / 32));
        work[i] = std::pow((double)tiles, 1.5);
        total += want[i];
    }
    if (total <= maxg) return want;
    int64_t left = maxg - (int64_t)I;
    while (left > 0) {
        size_t big = I;
        double worst = -1.0;
        for (size_t i = 0; i < I; ++i) {
            if (nb[i] >= want[i]) continue;
            const double t = work[i] / nb[i];
            if (t > worst) {
                worst = t;
                big = i;
            }
        }
        if (big == I) break;
        ++nb[big];
        --left;
    }
    return nb;
}

}  // namespace

int run_cg_batch(Device &d, const std::vector<CgRequest> &reqs, nsfem_cg_stats *stats,
                 int block_scale, double *batch_seconds, bool finish) {
    if (batch_seconds) *batch_seconds = 0.0;
    if (reqs.empty()) return NSFEM_OK;
    const int nc = reqs[0].ncol;
    const bool mixed = reqs[0].precision == NSFEM_PRECISION_FP32_REFINED;
    for (const CgRequest &r : reqs) {
        NSFEM_REQUIRE(r.m && r.m->ready && r.m->rows == r.m->cols, NSFEM_INVALID_ARGUMENT,
                      "batched CG needs square uploaded matrices");
        NSFEM_REQUIRE(r.ncol == nc && (r.precision == NSFEM_PRECISION_FP32_REFINED) == mixed &&
                          (mixed || r.precision == NSFEM_PRECISION_FP64),
                      NSFEM_INVALID_ARGUMENT,
                      "a batch needs one column count and one precision");
        NSFEM_REQUIRE(nc == 1 || nc == 2, NSFEM_INVALID_ARGUMENT, "columns must be 1 or 2");
    }
    bool idx16 = want_idx16();
    // batched launches run 1024-thread blocks: one load per nonzero per
    // lane unless NSFEM_CG_PHASE says otherwise
    int form = mixed && nc == 1 ? phase_form(*reqs[0].m, false) : -1;
    int wi = 4;
    for (const CgRequest &r : reqs) {
        idx16 = idx16 && r.m->dcol_ok;
        if (phase_form(*r.m, false) != form) form = -1;
        wi = std::max(wi, items_per_lane(*r.m));
    }
    if (form >= 0 && idx16) wi = form;
    CgBatchMaxGrid mg{d, 0};
    cg_dispatch(nc, mixed, idx16, wi, mg);
    const int maxg = mg.out;
    int worst = NSFEM_OK;
    const size_t vt = mixed ? sizeof(float) : sizeof(double);
    // instances beyond the co-resident block count run in further launches
    for (size_t first = 0; first < reqs.size(); first += (size_t)maxg) {
        const size_t I = std::min(reqs.size() - first, (size_t)maxg);
        std::vector<const CgRequest *> rq(I);
        for (size_t i = 0; i < I; ++i) rq[i] = &reqs[first + i];
        const std::vector<int> nb = split_blocks(rq, maxg, std::max(1, block_scale));
        // per-instance memory
        std::vector<size_t> voff(I), vbytes(I);
        size_t work = 0;
        for (size_t i = 0; i < I; ++i) {
            if (mixed) ensure_val32(d, *rq[i]->m);
            ensure_block_tiles(d, *rq[i]->m, nb[i]);
            vbytes[i] = round256((size_t)rq[i]->m->rows * nc * vt);
            voff[i] = work;
            work += 6 * vbytes[i];
        }
        const size_t sync_bytes = round256(sizeof(SyncState));
        d.batch_work.ensure(std::max<size_t>(work, 256));
        d.batch_sync.ensure(sync_bytes * I + 256);
        d.batch_res.ensure(2 * I);
        d.batch_args.ensure(sizeof(CgArgs) * I);
        d.batch_scal.ensure(2 * I);
        unsigned char *sm = d.batch_sync.ptr + (256 - ((uintptr_t)d.batch_sync.ptr % 256)) % 256;
        std::vector<CgArgs> args(I);
        std::vector<int2> map;
        for (size_t i = 0; i < I; ++i) {
            const CgRequest &r = *rq[i];
            Matrix &m = *r.m;
            CgArgs &a = args[i];
            a = CgArgs{};
            a.n = m.rows;
            a.meta = m.wmeta.ptr;
            a.tiles = m.wtiles.ptr;
            a.block_tiles = m.block_tiles.ptr;
            a.rs = m.row_start.ptr;
            a.col = m.col.ptr;
            a.dcol = m.dcol.ptr;
            a.val64 = m.val.ptr;
            a.val32 = m.val32.ptr;
            a.b = r.b;
            a.x = r.x;
            unsigned char *w = d.batch_work.ptr + voff[i];
            a.r = w;
            a.p0 = w + vbytes[i];
            a.p1 = w + 2 * vbytes[i];
            a.q = w + 3 * vbytes[i];
            set_state_buffers(a, w, vbytes[i]);
            set_cg_params(d, r, m.rows, a);
            a.sync = reinterpret_cast<SyncState *>(sm + i * sync_bytes);
            a.res = d.batch_res.ptr + 2 * i;
            a.nblocks = nb[i];
            for (int k = 0; k < nb[i]; ++k) map.push_back(make_int2((int)i, k));
        }
        const int grid = (int)map.size();
        NSFEM_REQUIRE(grid <= maxg, NSFEM_ERROR, "batched CG: block split exceeds co-residency");
        d.batch_map.ensure(map.size());
        NSFEM_HIP(hipMemcpyAsync(d.batch_args.ptr, args.data(), sizeof(CgArgs) * I,
                                 hipMemcpyHostToDevice, d.stream));
        NSFEM_HIP(hipMemcpyAsync(d.batch_map.ptr, map.data(), sizeof(int2) * map.size(),
                                 hipMemcpyHostToDevice, d.stream));
        for (size_t i = 0; i < I; ++i) {
            const int ng = (nb[i] + kGroup - 1) / kGroup;
            k_sync_reset<<<(ng + 255) / 256, 256, 0, d.stream>>>(args[i].sync, ng);
        }
        NSFEM_HIP(hipGetLastError());
        NSFEM_HIP(hipEventRecord(d.ev_start, d.stream));
        const CgArgs *dargs = reinterpret_cast<const CgArgs *>(d.batch_args.ptr);
        CgBatchLaunch ln{d, dargs, d.batch_map.ptr, grid};
        cg_dispatch(nc, mixed, idx16, wi, ln);
        NSFEM_HIP(hipEventRecord(d.ev_stop, d.stream));
        // FP64 true residuals of every instance, one host synchronisation
        if (!mixed) {
            launch_fill(d, (int)(nc * I), 0.0, d.batch_scal.ptr);
            for (size_t i = 0; i < I; ++i) {
                const CgRequest &r = *rq[i];
                double *tmp = (double *)args[i].q;  // n * nc doubles of scratch
                for (int c = 0; c < nc; ++c) {
                    const size_t off = (size_t)c * r.m->rows;
                    launch_spmv(d, *r.m, r.x + off, r.b + off, tmp, 1);
                    launch_dot_accumulate(d, r.m->rows, tmp, tmp, d.batch_scal.ptr + nc * i + c);
                }
            }
        }
        std::vector<CgColumnResult> res(2 * I);
        std::vector<double> tres(nc * I, 0.0);
        NSFEM_HIP(hipMemcpyAsync(res.data(), d.batch_res.ptr, sizeof(CgColumnResult) * 2 * I,
                                 hipMemcpyDeviceToHost, d.stream));
        if (!mixed)
            NSFEM_HIP(hipMemcpyAsync(tres.data(), d.batch_scal.ptr, sizeof(double) * nc * I,
                                     hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        float ms = 0.f;
        NSFEM_HIP(hipEventElapsedTime(&ms, d.ev_start, d.ev_stop));
        if (batch_seconds) *batch_seconds += ms * 1e-3;
        for (size_t i = 0; i < I; ++i) {
            const CgRequest &r = *rq[i];
            bool redo = false;
            for (int c = 0; c < nc; ++c) {
                const CgColumnResult &cr = res[2 * i + c];
                if (cr.status == NSFEM_BARRIER_TIMEOUT)
                    fail(NSFEM_BARRIER_TIMEOUT, "batched CG kernel barrier timed out");
                const double fin = mixed ? cr.last_residual : std::sqrt(tres[nc * i + c]);
                if (cr.status != NSFEM_OK || fin > cr.tolerance * (1.0 + 1e-12)) redo = true;
                nsfem_cg_stats &st = stats[2 * (first + i) + c];
                st.iterations = cr.iterations;
                st.refinements = cr.refinements;
                st.status = cr.status;
                st.columns = nc;
                st.rhs_norm = cr.rhs_norm;
                st.initial_residual = cr.initial_residual;
                st.final_residual = fin;
                st.tolerance = cr.tolerance;
                st.seconds = ms * 1e-3;
            }
            if (redo && finish) {
                // stagnation, the iteration limit or FP64 drift: finish this
                // instance alone from its current x, to its original
                // absolute tolerance
                CgRequest r2 = r;
                r2.rtol = 0.0;
                r2.atol = res[2 * i].tolerance;
                if (nc == 2) r2.atol = std::min(r2.atol, res[2 * i + 1].tolerance);
                if (r2.max_iter > 0)
                    r2.max_iter = std::max(1, r2.max_iter - std::max(res[2 * i].iterations,
                                                                     res[2 * i + nc - 1].iterations));
                nsfem_cg_stats s2[2]{};
                const int st2 = run_cg(d, r2, s2);
                for (int c = 0; c < nc; ++c) {
                    nsfem_cg_stats &st = stats[2 * (first + i) + c];
                    st.iterations += s2[c].iterations;
                    st.refinements += s2[c].refinements + 1;
                    st.status = s2[c].status;
                    st.final_residual = s2[c].final_residual;
                    st.seconds += s2[c].seconds;
                }
                if (batch_seconds) *batch_seconds += s2[0].seconds;
                if (st2 != NSFEM_OK && worst == NSFEM_OK) worst = st2;
            }
        }
    }
    return worst;
}

}  // namespace nsfem
