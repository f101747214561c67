// extern "C" entry points of libnsfem_hip.so (declared in include/nsfem_gpu.h).
// Each call validates its arguments on the host before any kernel is
// launched, so that every index a kernel reads is in range.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "nsfem_internal.hpp"

namespace nsfem {

static thread_local std::string g_last_error;

void set_last_error(const std::string &msg) { g_last_error = msg; }

void fail(int code, const std::string &msg) { throw Error(code, msg); }

void upload_csr(Device &d, Matrix &m, int rows, int cols, int64_t nnz, const int32_t *rs,
                const int32_t *col, const double *val) {
    NSFEM_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, NSFEM_INVALID_ARGUMENT,
                  "matrix sizes must be non-negative");
    NSFEM_REQUIRE(nnz < (int64_t)INT32_MAX, NSFEM_INVALID_ARGUMENT, "nnz must fit int32");
    NSFEM_REQUIRE(rs && (nnz == 0 || (col && val)), NSFEM_INVALID_ARGUMENT, "null CSR array");
    NSFEM_REQUIRE(rs[0] == 0 && rs[rows] == nnz, NSFEM_INVALID_ARGUMENT,
                  "row_start must start at 0 and end at nnz");
    for (int r = 0; r < rows; ++r) {
        NSFEM_REQUIRE(rs[r] <= rs[r + 1], NSFEM_INVALID_ARGUMENT, "row_start must not decrease");
        for (int32_t k = rs[r]; k < rs[r + 1]; ++k) {
            NSFEM_REQUIRE(col[k] >= 0 && col[k] < cols, NSFEM_INVALID_ARGUMENT,
                          "column index out of range");
            NSFEM_REQUIRE(k == rs[r] || col[k - 1] < col[k], NSFEM_INVALID_ARGUMENT,
                          "columns must ascend strictly within a row");
        }
    }
    std::lock_guard<std::mutex> lk(m.mtx);
    d.set_current();
    m.ready = false;
    m.rows = rows;
    m.cols = cols;
    m.nnz = nnz;
    m.row_start.resize((size_t)rows + 1);
    m.col.resize((size_t)std::max<int64_t>(nnz, 1));
    m.val.resize((size_t)std::max<int64_t>(nnz, 1));
    m.val32.release();
    m.jscale.release();
    m.jval32.release();
    m.jacobi_built = m.jacobi_ok = false;
    NSFEM_HIP(hipMemcpy(m.row_start.ptr, rs, sizeof(int32_t) * ((size_t)rows + 1),
                        hipMemcpyHostToDevice));
    if (nnz) {
        NSFEM_HIP(hipMemcpy(m.col.ptr, col, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
        NSFEM_HIP(hipMemcpy(m.val.ptr, val, sizeof(double) * nnz, hipMemcpyHostToDevice));
    } else {
        // kernels may read entry 0 of an empty matrix speculatively: keep it
        // a valid column index with a zero value
        const int32_t zc = 0;
        const double zv = 0.0;
        NSFEM_HIP(hipMemcpy(m.col.ptr, &zc, sizeof(int32_t), hipMemcpyHostToDevice));
        NSFEM_HIP(hipMemcpy(m.val.ptr, &zv, sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<int32_t> hrs(rs, rs + rows + 1);
    build_tiles(m, hrs);
    // the FP32 copy the refined solver streams, made here so that no solve
    // has to allocate it (the no-allocation entry point relies on it)
    if (rows == cols && nnz > 0) ensure_val32(d, m);
    m.ready = true;
}

void adopt_csr(Device &d, Matrix &m, DeviceCsr &c) {
    std::lock_guard<std::mutex> lk(m.mtx);
    d.set_current();
    m.ready = false;
    m.rows = c.rows;
    m.cols = c.cols;
    m.nnz = c.nnz;
    m.row_start = std::move(c.rs);
    m.col = std::move(c.col);
    m.val = std::move(c.val);
    m.val32.release();
    m.jscale.release();
    m.jval32.release();
    m.jacobi_built = m.jacobi_ok = false;
    std::vector<int32_t> hrs((size_t)m.rows + 1);
    NSFEM_HIP(hipMemcpy(hrs.data(), m.row_start.ptr, sizeof(int32_t) * hrs.size(),
                        hipMemcpyDeviceToHost));
    build_tiles(m, hrs);
    if (m.rows == m.cols && m.nnz > 0) ensure_val32(d, m);
    m.ready = true;
}

void upload_tree(Device &d, KdTreeGpu &t, int32_t tn, const uint32_t *w, int64_t lc,
                 const int32_t *leaf, int V, const float *xy, int E, const int32_t *el) {
    NSFEM_REQUIRE(tn > 0 && V > 0 && E > 0 && lc >= 0, NSFEM_INVALID_ARGUMENT,
                  "KD tree upload: empty tree or mesh");
    NSFEM_REQUIRE(tn < (1 << 30), NSFEM_INVALID_ARGUMENT, "KD tree: too many nodes");
    for (int64_t k = 0; k < 6ll * E; ++k)
        NSFEM_REQUIRE(el[k] >= 0 && el[k] < V, NSFEM_INVALID_ARGUMENT,
                      "KD tree upload: element node index out of range");
    for (int64_t k = 0; k < lc; ++k)
        NSFEM_REQUIRE(leaf[k] >= 0 && leaf[k] < E, NSFEM_INVALID_ARGUMENT,
                      "KD tree upload: leaf triangle index out of range");
    // Node checks: children strictly after their parent (so every descent
    // terminates), leaf ranges inside the leaf array, depth <= 30 (so the
    // 32-entry nearest-neighbour stack cannot overflow).
    std::vector<uint8_t> depth((size_t)tn, 0);
    std::vector<uint8_t> seen((size_t)tn, 0);
    seen[0] = 1;
    for (int32_t i = 0; i < tn; ++i) {
        NSFEM_REQUIRE(seen[i], NSFEM_INVALID_ARGUMENT, "KD tree upload: unreachable node");
        const uint32_t f = w[2 * i];
        const uint32_t kind = f & 3u;
        NSFEM_REQUIRE(kind != 3u, NSFEM_INVALID_ARGUMENT, "KD tree upload: bad node flags");
        if (kind == 2u) {
            const int64_t off = w[2 * i + 1], cnt = f >> 2;
            NSFEM_REQUIRE(off + cnt <= lc, NSFEM_INVALID_ARGUMENT,
                          "KD tree upload: leaf range outside the leaf array");
        } else {
            const int64_t right = f >> 2;
            NSFEM_REQUIRE(i + 1 < tn && right > i + 1 && right < tn, NSFEM_INVALID_ARGUMENT,
                          "KD tree upload: child index out of order");
            NSFEM_REQUIRE(depth[i] < 30, NSFEM_INVALID_ARGUMENT, "KD tree upload: depth > 30");
            depth[i + 1] = depth[right] = (uint8_t)(depth[i] + 1);
            seen[i + 1] = seen[right] = 1;
        }
    }
    d.set_current();
    t.ready = false;
    t.nodes.resize(2 * (size_t)tn);
    t.leaf.resize((size_t)std::max<int64_t>(lc, 1));
    t.xy.resize(2 * (size_t)V);
    t.elements.resize(6 * (size_t)E);
    NSFEM_HIP(hipMemcpy(t.nodes.ptr, w, sizeof(uint32_t) * 2 * tn, hipMemcpyHostToDevice));
    if (lc) NSFEM_HIP(hipMemcpy(t.leaf.ptr, leaf, sizeof(int32_t) * lc, hipMemcpyHostToDevice));
    NSFEM_HIP(hipMemcpy(t.xy.ptr, xy, sizeof(float) * 2 * V, hipMemcpyHostToDevice));
    NSFEM_HIP(hipMemcpy(t.elements.ptr, el, sizeof(int32_t) * 6 * E, hipMemcpyHostToDevice));
    t.node_count = tn;
    t.leaf_count = lc;
    t.mesh_nodes = V;
    t.elements_count = E;
    build_tri_xy(d, t);
    NSFEM_HIP(hipStreamSynchronize(d.stream));
    t.ready = true;
    if (&t == &d.tree) {
        // initVelocityBuffers (cpp/gpu_simulation_device.cpp:76)
        for (auto &b : d.vel32) b.resize((size_t)V);
        d.vel_n = V;
    }
}

static Buffer &buf(const nsfem_buffer *b) {
    NSFEM_REQUIRE(b != nullptr, NSFEM_INVALID_ARGUMENT, "null buffer");
    NSFEM_REQUIRE(b->impl.dev != nullptr, NSFEM_INVALID_ARGUMENT,
                  "buffer belongs to a destroyed device");
    return const_cast<nsfem_buffer *>(b)->impl;
}

static Device &dev(nsfem_device *d) {
    NSFEM_REQUIRE(d != nullptr, NSFEM_INVALID_ARGUMENT, "null device");
    return d->impl;
}

static Matrix &slot(Device &d, int s) {
    NSFEM_REQUIRE(s >= 0 && s < NSFEM_MAX_SLOTS, NSFEM_INVALID_ARGUMENT, "matrix slot out of range");
    NSFEM_REQUIRE(d.slots[s].ready, NSFEM_NOT_READY, "matrix slot is empty");
    return d.slots[s];
}

static double *dptr(const nsfem_buffer *b, int64_t count, const char *what) {
    Buffer &B = buf(b);
    NSFEM_REQUIRE((int64_t)B.mem.count >= count * (int64_t)sizeof(double), NSFEM_INVALID_ARGUMENT,
                  std::string(what) + ": buffer smaller than the vector");
    return reinterpret_cast<double *>(B.mem.ptr);
}

}  // namespace nsfem

using namespace nsfem;

extern "C" {

const char *nsfem_last_error(void) { return g_last_error.c_str(); }

int nsfem_get_device_count(int *count) {
    return guard([&] {
        NSFEM_REQUIRE(count, NSFEM_INVALID_ARGUMENT, "null count");
        int n = 0;
        NSFEM_HIP(hipGetDeviceCount(&n));
        *count = n;
    });
}

int nsfem_device_create(int index, nsfem_device **out) {
    return guard([&] {
        NSFEM_REQUIRE(out, NSFEM_INVALID_ARGUMENT, "null output");
        *out = nullptr;
        int n = 0;
        NSFEM_HIP(hipGetDeviceCount(&n));
        NSFEM_REQUIRE(index >= 0 && index < n, NSFEM_INVALID_ARGUMENT,
                      "device index " + std::to_string(index) + " out of range (" +
                          std::to_string(n) + " devices)");
        auto *h = new nsfem_device();
        Device &d = h->impl;
        try {
            d.index = index;
            NSFEM_HIP(hipSetDevice(index));
            NSFEM_HIP(hipGetDeviceProperties(&d.props, index));
            d.cu_count = d.props.multiProcessorCount;
            int coop = 0;
            NSFEM_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, index));
            NSFEM_REQUIRE(coop, NSFEM_HIP_ERROR, "device does not support cooperative launches");
            int wall = 0;
            NSFEM_HIP(hipDeviceGetAttribute(&wall, hipDeviceAttributeWallClockRate, index));
            d.wall_khz = wall > 0 ? wall : 100000;
            NSFEM_HIP(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
            NSFEM_HIP(hipEventCreate(&d.ev_start));
            NSFEM_HIP(hipEventCreate(&d.ev_stop));
            scratch_init(d);  // this device's address of the static scratch array
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

int nsfem_device_destroy(nsfem_device *h) {
    return guard([&] {
        if (!h) return;
        Device &d = h->impl;
        (void)hipSetDevice(d.index);
        if (d.stream) (void)hipStreamSynchronize(d.stream);
        if (d.ev_start) (void)hipEventDestroy(d.ev_start);
        if (d.ev_stop) (void)hipEventDestroy(d.ev_stop);
        {
            std::lock_guard<std::mutex> lk(d.buf_mtx);
            for (Buffer *b : d.buffers) {
                b->mem.release();
                b->dev = nullptr;
            }
            d.buffers.clear();
            for (Simulation *sm : d.sims) simulation_detach(sm);
            d.sims.clear();
            for (GnnNet *g : d.gnns) {
                gnn_release(*g);
                g->dev = nullptr;
            }
            d.gnns.clear();
            for (GnnGraph *g : d.graphs) {
                gnn_graph_release(*g);
                g->dev = nullptr;
            }
            d.graphs.clear();
        }
        hipStream_t s = d.stream;
        delete h;  // frees every DeviceArray
        if (s) (void)hipStreamDestroy(s);
    });
}

int nsfem_device_info(nsfem_device *h, char *name, int name_len, int *cu, int64_t *hbm,
                      int *clock_khz) {
    return guard([&] {
        Device &d = dev(h);
        if (name && name_len > 0) {
            std::strncpy(name, d.props.gcnArchName, name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cu) *cu = d.cu_count;
        if (hbm) *hbm = (int64_t)d.props.totalGlobalMem;
        if (clock_khz) *clock_khz = d.props.clockRate;
    });
}

int nsfem_device_synchronize(nsfem_device *h) {
    return guard([&] {
        Device &d = dev(h);
        d.set_current();
        NSFEM_HIP(hipStreamSynchronize(d.stream));
    });
}

int nsfem_upload_matrix_csr(nsfem_device *h, int s, int rows, int cols, int64_t nnz,
                            const int32_t *rs, const int32_t *col, const double *val) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(s >= 0 && s < NSFEM_MAX_SLOTS, NSFEM_INVALID_ARGUMENT,
                      "matrix slot out of range");
        upload_csr(d, d.slots[s], rows, cols, nnz, rs, col, val);
    });
}

int nsfem_upload_matrix_triplet(nsfem_device *h, int s, int rows, int cols, int64_t count,
                                const int32_t *ti, const int32_t *tj, const double *tv) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(s >= 0 && s < NSFEM_MAX_SLOTS, NSFEM_INVALID_ARGUMENT,
                      "matrix slot out of range");
        NSFEM_REQUIRE(rows >= 0 && cols >= 0 && count >= 0, NSFEM_INVALID_ARGUMENT,
                      "negative size");
        NSFEM_REQUIRE(count == 0 || (ti && tj && tv), NSFEM_INVALID_ARGUMENT, "null triplet array");
        for (int64_t k = 0; k < count; ++k)
            NSFEM_REQUIRE(ti[k] >= 0 && ti[k] < rows && tj[k] >= 0 && tj[k] < cols,
                          NSFEM_INVALID_ARGUMENT, "triplet index out of range");
        // counting sort by row (stable), stable sort by column within a row,
        // duplicates summed in input order (cpp/gpu_sparse_matrix.cpp:7-49)
        std::vector<int64_t> cnt((size_t)rows + 1, 0);
        for (int64_t k = 0; k < count; ++k) cnt[ti[k] + 1]++;
        for (int r = 0; r < rows; ++r) cnt[r + 1] += cnt[r];
        std::vector<int64_t> perm((size_t)count);
        {
            std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
            for (int64_t k = 0; k < count; ++k) perm[pos[ti[k]]++] = k;
        }
        std::vector<int32_t> rs((size_t)rows + 1, 0), col;
        std::vector<double> val;
        col.reserve((size_t)count);
        val.reserve((size_t)count);
        for (int r = 0; r < rows; ++r) {
            auto b = perm.begin() + cnt[r], e = perm.begin() + cnt[r + 1];
            std::stable_sort(b, e, [&](int64_t a, int64_t c) { return tj[a] < tj[c]; });
            for (auto it = b; it != e;) {
                const int32_t c = tj[*it];
                double sum = tv[*it];
                for (++it; it != e && tj[*it] == c; ++it) sum += tv[*it];
                col.push_back(c);
                val.push_back(sum);
            }
            rs[r + 1] = (int32_t)col.size();
        }
        upload_csr(d, d.slots[s], rows, cols, (int64_t)col.size(), rs.data(), col.data(),
                   val.data());
    });
}

int nsfem_matrix_info(nsfem_device *h, int s, int *rows, int *cols, int64_t *nnz) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        if (rows) *rows = m.rows;
        if (cols) *cols = m.cols;
        if (nnz) *nnz = m.nnz;
    });
}

int nsfem_free_matrix(nsfem_device *h, int s) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(s >= 0 && s < NSFEM_MAX_SLOTS, NSFEM_INVALID_ARGUMENT,
                      "matrix slot out of range");
        Matrix &m = d.slots[s];
        std::lock_guard<std::mutex> lk(m.mtx);
        d.set_current();
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        matrix_release(m);
        m.rows = m.cols = 0;
        m.nnz = 0;
    });
}

int nsfem_buffer_create(nsfem_device *h, int64_t bytes, nsfem_buffer **out) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(out && bytes >= 0, NSFEM_INVALID_ARGUMENT, "bad buffer request");
        d.set_current();
        auto *b = new nsfem_buffer();
        b->impl.dev = &d;
        try {
            b->impl.mem.resize((size_t)bytes);
        } catch (...) {
            delete b;
            throw;
        }
        {
            std::lock_guard<std::mutex> lk(d.buf_mtx);
            d.buffers.insert(&b->impl);
        }
        *out = b;
    });
}

int nsfem_buffer_destroy(nsfem_buffer *b) {
    return guard([&] {
        if (!b) return;
        if (b->impl.dev) {  // still attached: its device is alive
            Device &d = *b->impl.dev;
            (void)hipSetDevice(d.index);
            (void)hipStreamSynchronize(d.stream);
            std::lock_guard<std::mutex> lk(d.buf_mtx);
            d.buffers.erase(&b->impl);
        }
        delete b;
    });
}

int nsfem_buffer_upload(nsfem_buffer *b, const void *src, int64_t bytes) {
    return guard([&] {
        Buffer &B = buf(b);
        NSFEM_REQUIRE(bytes >= 0 && (size_t)bytes <= B.mem.count && (src || bytes == 0),
                      NSFEM_INVALID_ARGUMENT, "upload larger than the buffer");
        B.dev->set_current();
        if (bytes)
            NSFEM_HIP(hipMemcpyAsync(B.mem.ptr, src, bytes, hipMemcpyHostToDevice, B.dev->stream));
        NSFEM_HIP(hipStreamSynchronize(B.dev->stream));
    });
}

int nsfem_buffer_download(const nsfem_buffer *b, void *dst, int64_t bytes) {
    return guard([&] {
        Buffer &B = buf(b);
        NSFEM_REQUIRE(bytes >= 0 && (size_t)bytes <= B.mem.count && (dst || bytes == 0),
                      NSFEM_INVALID_ARGUMENT, "download larger than the buffer");
        B.dev->set_current();
        if (bytes)
            NSFEM_HIP(hipMemcpyAsync(dst, B.mem.ptr, bytes, hipMemcpyDeviceToHost, B.dev->stream));
        NSFEM_HIP(hipStreamSynchronize(B.dev->stream));
    });
}

int64_t nsfem_buffer_byte_size(const nsfem_buffer *b) {
    return b ? (int64_t)b->impl.mem.count : -1;
}

void *nsfem_buffer_device_pointer(const nsfem_buffer *b) { return b ? b->impl.mem.ptr : nullptr; }

int nsfem_sp_r_mult(nsfem_device *h, int s, const nsfem_buffer *x, nsfem_buffer *y) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        const double *xp = dptr(x, m.cols, "spRMult mult");
        double *yp = dptr(y, m.rows, "spRMult res");
        NSFEM_REQUIRE((const void *)xp != (const void *)yp, NSFEM_INVALID_ARGUMENT,
                      "spRMult: mult must not alias res");
        d.set_current();
        launch_spmv(d, m, xp, nullptr, yp, 0);
    });
}

int nsfem_time_sp_r_mult(nsfem_device *h, int s, const nsfem_buffer *x, nsfem_buffer *y,
                         int repeats, double *seconds) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        const double *xp = dptr(x, m.cols, "spRMult mult");
        double *yp = dptr(y, m.rows, "spRMult res");
        NSFEM_REQUIRE((const void *)xp != (const void *)yp, NSFEM_INVALID_ARGUMENT,
                      "spRMult: mult must not alias res");
        NSFEM_REQUIRE(repeats > 0 && seconds, NSFEM_INVALID_ARGUMENT,
                      "time_sp_r_mult: repeats must be positive and seconds non-null");
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        NSFEM_HIP(hipEventRecord(d.ev_start, d.stream));
        for (int i = 0; i < repeats; ++i) launch_spmv(d, m, xp, nullptr, yp, 0);
        NSFEM_HIP(hipEventRecord(d.ev_stop, d.stream));
        NSFEM_HIP(hipEventSynchronize(d.ev_stop));
        float ms = 0.f;
        NSFEM_HIP(hipEventElapsedTime(&ms, d.ev_start, d.ev_stop));
        *seconds = (double)ms * 1e-3 / repeats;
    });
}

int nsfem_sp_r_mult_sub(nsfem_device *h, int s, const nsfem_buffer *lhs, const nsfem_buffer *x,
                        nsfem_buffer *y) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        const double *lp = dptr(lhs, m.rows, "spRMultSub lhs");
        const double *xp = dptr(x, m.cols, "spRMultSub mult");
        double *yp = dptr(y, m.rows, "spRMultSub res");
        NSFEM_REQUIRE((const void *)xp != (const void *)yp, NSFEM_INVALID_ARGUMENT,
                      "spRMultSub: mult must not alias res");
        d.set_current();
        launch_spmv(d, m, xp, lp, yp, 1);
    });
}

int nsfem_saxpy(nsfem_device *h, int n, double a, const nsfem_buffer *x, nsfem_buffer *y) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0, NSFEM_INVALID_ARGUMENT, "negative length");
        const double *xp = dptr(x, n, "saxpy x");
        double *yp = dptr(y, n, "saxpy y");
        d.set_current();
        launch_axpy(d, n, a, xp, yp);
    });
}

int nsfem_saxpby(nsfem_device *h, int n, double a, double b, const nsfem_buffer *x,
                 const nsfem_buffer *y, nsfem_buffer *out) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0, NSFEM_INVALID_ARGUMENT, "negative length");
        const double *xp = dptr(x, n, "saxpby x");
        const double *yp = dptr(y, n, "saxpby y");
        double *op = dptr(out, n, "saxpby result");
        d.set_current();
        launch_axpby(d, n, a, b, xp, yp, op);
    });
}

int nsfem_dot_product(nsfem_device *h, int n, const nsfem_buffer *a, const nsfem_buffer *b,
                      nsfem_buffer *result) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0, NSFEM_INVALID_ARGUMENT, "negative length");
        const double *ap = dptr(a, n, "dotProduct a");
        const double *bp = dptr(b, n, "dotProduct b");
        double *rp = dptr(result, 1, "dotProduct result");
        d.set_current();
        launch_dot_accumulate(d, n, ap, bp, rp);
    });
}

int nsfem_cg_solve(nsfem_device *h, int s, const nsfem_buffer *b, nsfem_buffer *x, int columns,
                   int max_iterations, double rtol, double atol, int precision,
                   nsfem_cg_stats *stats) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        NSFEM_REQUIRE(columns == 1 || columns == 2, NSFEM_INVALID_ARGUMENT,
                      "columns must be 1 or 2");
        NSFEM_REQUIRE(rtol >= 0 && atol >= 0, NSFEM_INVALID_ARGUMENT, "negative tolerance");
        const double *bp = dptr(b, (int64_t)m.rows * columns, "cg b");
        double *xp = dptr(x, (int64_t)m.rows * columns, "cg x");
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        CgRequest rq{};
        rq.m = &m;
        rq.b = bp;
        rq.x = xp;
        rq.ncol = columns;
        rq.max_iter = max_iterations;
        rq.rtol = rtol;
        rq.atol = atol;
        rq.precision = precision;
        const int st = run_cg(d, rq, stats);
        if (st != NSFEM_OK)
            fail(st, st == NSFEM_MAX_ITERATIONS_REACHED ? "Max iterations reached!"
                                                        : "CG breakdown: (Ap, p) <= 0");
    });
}

int nsfem_cg_solve_noalloc(nsfem_device *h, int s, const nsfem_buffer *b, nsfem_buffer *x,
                           int columns, int max_iterations, double rtol, double atol,
                           int precision, nsfem_cg_stats *stats) {
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        NSFEM_REQUIRE(columns == 1 || columns == 2, NSFEM_INVALID_ARGUMENT,
                      "columns must be 1 or 2");
        NSFEM_REQUIRE(rtol >= 0 && atol >= 0, NSFEM_INVALID_ARGUMENT, "negative tolerance");
        const double *bp = dptr(b, (int64_t)m.rows * columns, "cg b");
        double *xp = dptr(x, (int64_t)m.rows * columns, "cg x");
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        CgRequest rq{};
        rq.m = &m;
        rq.b = bp;
        rq.x = xp;
        rq.ncol = columns;
        rq.max_iter = max_iterations;
        rq.rtol = rtol;
        rq.atol = atol;
        rq.precision = precision;
        const int st = run_cg_noalloc(d, rq, stats);
        if (st != NSFEM_OK)
            fail(st, st == NSFEM_MAX_ITERATIONS_REACHED ? "Max iterations reached!"
                                                        : "CG breakdown: (Ap, p) <= 0");
    });
}

int nsfem_conjugate_gradient(nsfem_device *h, int s, const double *b, const double *x0, double *x,
                             int max_iterations, double eps, int use_megakernel) {
    (void)use_megakernel;
    return guard([&] {
        Device &d = dev(h);
        Matrix &m = slot(d, s);
        NSFEM_REQUIRE(b && x0 && x, NSFEM_INVALID_ARGUMENT, "null vector");
        NSFEM_REQUIRE(eps >= 0, NSFEM_INVALID_ARGUMENT, "negative eps");
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        DeviceArray<double> db, dx;
        db.resize((size_t)m.rows);
        dx.resize((size_t)m.rows);
        NSFEM_HIP(hipMemcpyAsync(db.ptr, b, sizeof(double) * m.rows, hipMemcpyHostToDevice, d.stream));
        NSFEM_HIP(hipMemcpyAsync(dx.ptr, x0, sizeof(double) * m.rows, hipMemcpyHostToDevice, d.stream));
        CgRequest rq{};
        rq.m = &m;
        rq.b = db.ptr;
        rq.x = dx.ptr;
        rq.ncol = 1;
        rq.max_iter = max_iterations;
        rq.rtol = 0.0;
        rq.atol = eps;
        rq.precision = NSFEM_PRECISION_FP64;
        const int st = run_cg(d, rq, nullptr);
        NSFEM_HIP(hipMemcpyAsync(x, dx.ptr, sizeof(double) * m.rows, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
        if (st != NSFEM_OK)
            fail(st, st == NSFEM_MAX_ITERATIONS_REACHED ? "Max iterations reached!"
                                                        : "CG breakdown: (Ap, p) <= 0");
    });
}

int nsfem_kdtree_build(int V, const float *xy, int E, const int32_t *el, int max_depth,
                       int min_leaf, nsfem_kdtree **out) {
    return guard([&] {
        NSFEM_REQUIRE(out && xy && el, NSFEM_INVALID_ARGUMENT, "null argument");
        auto *t = new nsfem_kdtree();
        try {
            kd_build_host(V, xy, E, el, max_depth, min_leaf, t->impl);
        } catch (...) {
            delete t;
            throw;
        }
        *out = t;
    });
}

int nsfem_kdtree_destroy(nsfem_kdtree *t) {
    return guard([&] { delete t; });
}

int nsfem_kdtree_info(const nsfem_kdtree *t, int32_t *nc, int64_t *lc, int32_t *md, float bbox[4]) {
    return guard([&] {
        NSFEM_REQUIRE(t, NSFEM_INVALID_ARGUMENT, "null tree");
        if (nc) *nc = (int32_t)(t->impl.nodes.size() / 2);
        if (lc) *lc = (int64_t)t->impl.leaf.size();
        if (md) *md = t->impl.max_depth;
        if (bbox) std::memcpy(bbox, t->impl.bbox, sizeof(float) * 4);
    });
}

int nsfem_kdtree_copy(const nsfem_kdtree *t, uint32_t *words, int32_t *leaf) {
    return guard([&] {
        NSFEM_REQUIRE(t, NSFEM_INVALID_ARGUMENT, "null tree");
        if (words) std::memcpy(words, t->impl.nodes.data(), sizeof(uint32_t) * t->impl.nodes.size());
        if (leaf) std::memcpy(leaf, t->impl.leaf.data(), sizeof(int32_t) * t->impl.leaf.size());
    });
}

int nsfem_upload_kdtree(nsfem_device *h, const nsfem_kdtree *t, int V, const float *xy, int E,
                        const int32_t *el) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(t && xy && el, NSFEM_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(d.mtx);
        upload_tree(d, d.tree, (int32_t)(t->impl.nodes.size() / 2), t->impl.nodes.data(),
                    (int64_t)t->impl.leaf.size(), t->impl.leaf.data(), V, xy, E, el);
    });
}

int nsfem_upload_kdtree_arrays(nsfem_device *h, int32_t tn, const uint32_t *w, int64_t lc,
                               const int32_t *leaf, int V, const float *xy, int E,
                               const int32_t *el) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(w && xy && el && (lc == 0 || leaf), NSFEM_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(d.mtx);
        upload_tree(d, d.tree, tn, w, lc, leaf, V, xy, E, el);
    });
}

int nsfem_init_velocity_buffers(nsfem_device *h, int n) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0, NSFEM_INVALID_ARGUMENT, "negative size");
        d.set_current();
        for (auto &b : d.vel32) b.resize((size_t)n);
        d.vel_n = n;
    });
}

int nsfem_advect(nsfem_device *h, int n, const float *u, const float *v, float dt, float *uo,
                 float *vo) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0 && n <= d.vel_n, NSFEM_INVALID_ARGUMENT,
                      "advect: n exceeds the velocity buffers");
        NSFEM_REQUIRE(n == 0 || (u && v && uo && vo), NSFEM_INVALID_ARGUMENT, "null vector");
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        const size_t bytes = sizeof(float) * (size_t)n;
        NSFEM_HIP(hipMemcpyAsync(d.vel32[0].ptr, u, bytes, hipMemcpyHostToDevice, d.stream));
        NSFEM_HIP(hipMemcpyAsync(d.vel32[1].ptr, v, bytes, hipMemcpyHostToDevice, d.stream));
        launch_advect_f32(d, d.tree, n, d.vel32[0].ptr, d.vel32[1].ptr, dt, d.vel32[2].ptr, d.vel32[3].ptr);
        NSFEM_HIP(hipMemcpyAsync(uo, d.vel32[2].ptr, bytes, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipMemcpyAsync(vo, d.vel32[3].ptr, bytes, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
    });
}

int nsfem_advect_f64(nsfem_device *h, int n, const nsfem_buffer *u, const nsfem_buffer *v,
                     double dt, nsfem_buffer *uo, nsfem_buffer *vo) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(d.tree.ready, NSFEM_NOT_READY, "KD tree not uploaded");
        NSFEM_REQUIRE(n >= 0 && n <= d.tree.mesh_nodes, NSFEM_INVALID_ARGUMENT,
                      "advect: n exceeds the mesh nodes");
        const int64_t need = d.tree.mesh_nodes;  // interpolation reads any mesh node
        const double *up = dptr(u, need, "advect u");
        const double *vp = dptr(v, need, "advect v");
        double *uop = dptr(uo, n, "advect uOut");
        double *vop = dptr(vo, n, "advect vOut");
        d.set_current();
        launch_advect_f64(d, d.tree, n, up, vp, dt, uop, vop);
    });
}

int nsfem_locate(nsfem_device *h, int n, const float *pts, int32_t *elem, float *xi, float *eta,
                 int32_t *nearest) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(n >= 0, NSFEM_INVALID_ARGUMENT, "negative count");
        NSFEM_REQUIRE(n == 0 || (pts && elem && xi && eta && nearest), NSFEM_INVALID_ARGUMENT,
                      "null argument");
        NSFEM_REQUIRE(d.tree.ready, NSFEM_NOT_READY, "KD tree not uploaded");
        if (n == 0) return;
        std::lock_guard<std::mutex> lk(d.mtx);
        d.set_current();
        DeviceArray<float> dp, dxi, deta;
        DeviceArray<int32_t> de, dn;
        dp.resize(2 * (size_t)n);
        dxi.resize(n);
        deta.resize(n);
        de.resize(n);
        dn.resize(n);
        NSFEM_HIP(hipMemcpyAsync(dp.ptr, pts, sizeof(float) * 2 * n, hipMemcpyHostToDevice, d.stream));
        launch_locate(d, d.tree, n, dp.ptr, de.ptr, dxi.ptr, deta.ptr, dn.ptr);
        NSFEM_HIP(hipMemcpyAsync(elem, de.ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipMemcpyAsync(xi, dxi.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipMemcpyAsync(eta, deta.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, d.stream));
        NSFEM_HIP(hipMemcpyAsync(nearest, dn.ptr, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                                 d.stream));
        NSFEM_HIP(hipStreamSynchronize(d.stream));
    });
}

int nsfem_assemble(int V, int P, const float *xy, int E, const int32_t *el, int nvb,
                   const int32_t *vb, const double *gu, const double *gv, int npb,
                   const int32_t *pb, const double *gp, double nu, double dt, int threads,
                   nsfem_assembly **out) {
    return guard([&] {
        NSFEM_REQUIRE(out && xy && el, NSFEM_INVALID_ARGUMENT, "null argument");
        NSFEM_REQUIRE((nvb == 0 || (vb && gu && gv)) && (npb == 0 || (pb && gp)),
                      NSFEM_INVALID_ARGUMENT, "null boundary arrays");
        auto *a = new nsfem_assembly();
        try {
            assemble_host(V, P, xy, E, el, nvb, vb, gu, gv, npb, pb, gp, nu, dt, threads, a->impl);
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

int nsfem_assemble_device(nsfem_device *h, int V, int P, const float *xy, int E,
                          const int32_t *el, int nvb, const int32_t *vb, const double *gu,
                          const double *gv, int npb, const int32_t *pb, const double *gp,
                          double nu, double dt, nsfem_assembly **out) {
    return guard([&] {
        Device &d = dev(h);
        NSFEM_REQUIRE(out && xy && el, NSFEM_INVALID_ARGUMENT, "null argument");
        NSFEM_REQUIRE((nvb == 0 || (vb && gu && gv)) && (npb == 0 || (pb && gp)),
                      NSFEM_INVALID_ARGUMENT, "null boundary arrays");
        std::lock_guard<std::mutex> lk(d.mtx);
        DeviceAssembly da;
        assemble_device(d, V, P, xy, E, el, nvb, vb, gu, gv, npb, pb, gp, nu, dt, nullptr,
                        nullptr, da);
        auto *a = new nsfem_assembly();
        try {
            for (int k = 0; k < 5; ++k) {
                HostCsr &hc = a->impl.m[k];
                DeviceCsr &dc = da.m[k];
                hc.rows = dc.rows;
                hc.cols = dc.cols;
                hc.rs.resize((size_t)dc.rows + 1);
                hc.col.resize((size_t)dc.nnz);
                hc.val.resize((size_t)dc.nnz);
                NSFEM_HIP(hipMemcpy(hc.rs.data(), dc.rs.ptr, sizeof(int32_t) * hc.rs.size(),
                                    hipMemcpyDeviceToHost));
                if (dc.nnz) {
                    NSFEM_HIP(hipMemcpy(hc.col.data(), dc.col.ptr, sizeof(int32_t) * dc.nnz,
                                        hipMemcpyDeviceToHost));
                    NSFEM_HIP(hipMemcpy(hc.val.data(), dc.val.ptr, sizeof(double) * dc.nnz,
                                        hipMemcpyDeviceToHost));
                }
            }
            a->impl.c_u.resize(V);
            a->impl.c_v.resize(V);
            a->impl.c_p.resize(P);
            NSFEM_HIP(hipMemcpy(a->impl.c_u.data(), da.c_u.ptr, sizeof(double) * V,
                                hipMemcpyDeviceToHost));
            NSFEM_HIP(hipMemcpy(a->impl.c_v.data(), da.c_v.ptr, sizeof(double) * V,
                                hipMemcpyDeviceToHost));
            NSFEM_HIP(hipMemcpy(a->impl.c_p.data(), da.c_p.ptr, sizeof(double) * P,
                                hipMemcpyDeviceToHost));
        } catch (...) {
            delete a;
            throw;
        }
        *out = a;
    });
}

int nsfem_assembly_destroy(nsfem_assembly *a) {
    return guard([&] { delete a; });
}

int nsfem_assembly_matrix_info(const nsfem_assembly *a, int which, int *rows, int *cols,
                               int64_t *nnz) {
    return guard([&] {
        NSFEM_REQUIRE(a && which >= 0 && which < 5, NSFEM_INVALID_ARGUMENT, "bad matrix id");
        const HostCsr &c = a->impl.m[which];
        if (rows) *rows = c.rows;
        if (cols) *cols = c.cols;
        if (nnz) *nnz = (int64_t)c.col.size();
    });
}

int nsfem_assembly_matrix_copy(const nsfem_assembly *a, int which, int32_t *rs, int32_t *col,
                               double *val) {
    return guard([&] {
        NSFEM_REQUIRE(a && which >= 0 && which < 5, NSFEM_INVALID_ARGUMENT, "bad matrix id");
        const HostCsr &c = a->impl.m[which];
        if (rs) std::memcpy(rs, c.rs.data(), sizeof(int32_t) * c.rs.size());
        if (col) std::memcpy(col, c.col.data(), sizeof(int32_t) * c.col.size());
        if (val) std::memcpy(val, c.val.data(), sizeof(double) * c.val.size());
    });
}

int nsfem_assembly_bc_copy(const nsfem_assembly *a, double *cu, double *cv, double *cp) {
    return guard([&] {
        NSFEM_REQUIRE(a, NSFEM_INVALID_ARGUMENT, "null assembly");
        if (cu) std::memcpy(cu, a->impl.c_u.data(), sizeof(double) * a->impl.c_u.size());
        if (cv) std::memcpy(cv, a->impl.c_v.data(), sizeof(double) * a->impl.c_v.size());
        if (cp) std::memcpy(cp, a->impl.c_p.data(), sizeof(double) * a->impl.c_p.size());
    });
}

// ------------------------------------------------------ learned warm start
static GnnNet &net(nsfem_gnn *g) {
    NSFEM_REQUIRE(g && g->impl, NSFEM_INVALID_ARGUMENT, "null network");
    NSFEM_REQUIRE(g->impl->dev, NSFEM_INVALID_ARGUMENT, "network belongs to a destroyed device");
    return *g->impl;
}

static GnnGraph &graph(nsfem_gnn_graph *g, const GnnNet &n) {
    NSFEM_REQUIRE(g && g->impl && g->impl->dev, NSFEM_INVALID_ARGUMENT,
                  "null graph or graph of a destroyed device");
    NSFEM_REQUIRE(g->impl->dev == n.dev, NSFEM_INVALID_ARGUMENT,
                  "graph and network live on different devices");
    return *g->impl;
}

static const double *vec(const nsfem_buffer *b, const GnnGraph &g, const char *what) {
    Buffer &bb = buf(b);
    NSFEM_REQUIRE(bb.dev == g.dev, NSFEM_INVALID_ARGUMENT, std::string(what) + " is on another device");
    NSFEM_REQUIRE((int64_t)bb.mem.bytes() >= (int64_t)g.n * (int64_t)sizeof(double),
                  NSFEM_INVALID_ARGUMENT, std::string(what) + " is shorter than the graph");
    return (const double *)bb.mem.ptr;
}

int nsfem_gnn_create(nsfem_device *h, uint64_t seed, nsfem_gnn **out) {
    return guard([&] {
        NSFEM_REQUIRE(h && out, NSFEM_INVALID_ARGUMENT, "null argument");
        Device &d = h->impl;
        std::lock_guard<std::mutex> lk(d.mtx);
        auto *w = new nsfem_gnn();
        try {
            w->impl = gnn_create(d, seed);
        } catch (...) {
            delete w;
            throw;
        }
        std::lock_guard<std::mutex> lb(d.buf_mtx);
        d.gnns.insert(w->impl);
        *out = w;
    });
}

int nsfem_gnn_destroy(nsfem_gnn *g) {
    return guard([&] {
        if (!g) return;
        if (g->impl && g->impl->dev) {
            Device &d = *g->impl->dev;
            (void)hipSetDevice(d.index);
            (void)hipStreamSynchronize(d.stream);
            std::lock_guard<std::mutex> lk(d.buf_mtx);
            for (Simulation *s : d.sims) simulation_forget_gnn(s, g->impl);
            d.gnns.erase(g->impl);
        }
        delete g->impl;
        delete g;
    });
}

int nsfem_gnn_get_params(nsfem_gnn *g, float *params) {
    return guard([&] {
        GnnNet &n = net(g);
        NSFEM_REQUIRE(params, NSFEM_INVALID_ARGUMENT, "null params");
        gnn_get(n, n.params.ptr, params);
    });
}

int nsfem_gnn_set_params(nsfem_gnn *g, const float *params) {
    return guard([&] {
        GnnNet &n = net(g);
        NSFEM_REQUIRE(params, NSFEM_INVALID_ARGUMENT, "null params");
        std::lock_guard<std::mutex> lk(n.dev->mtx);
        gnn_set_params(n, params);
    });
}

int nsfem_gnn_graph_create(nsfem_device *h, int n, int64_t nnz, const int32_t *rs,
                           const int32_t *col, const double *val, const double *area,
                           const float *xy, nsfem_gnn_graph **out) {
    return guard([&] {
        NSFEM_REQUIRE(h && out, NSFEM_INVALID_ARGUMENT, "null argument");
        Device &d = h->impl;
        auto *w = new nsfem_gnn_graph();
        try {
            w->impl = gnn_graph_create(d, n, nnz, rs, col, val, area, xy);
        } catch (...) {
            delete w;
            throw;
        }
        std::lock_guard<std::mutex> lb(d.buf_mtx);
        d.graphs.insert(w->impl);
        *out = w;
    });
}

int nsfem_gnn_graph_destroy(nsfem_gnn_graph *g) {
    return guard([&] {
        if (!g) return;
        if (g->impl && g->impl->dev) {
            Device &d = *g->impl->dev;
            (void)hipSetDevice(d.index);
            (void)hipStreamSynchronize(d.stream);
            std::lock_guard<std::mutex> lk(d.buf_mtx);
            d.graphs.erase(g->impl);
        }
        delete g->impl;
        delete g;
    });
}

int nsfem_gnn_predict(nsfem_gnn *g, nsfem_gnn_graph *gr, const nsfem_buffer *b,
                      nsfem_buffer *p_hat) {
    return guard([&] {
        GnnNet &n = net(g);
        GnnGraph &G = graph(gr, n);
        const double *bb = vec(b, G, "b");
        double *out = const_cast<double *>(vec(p_hat, G, "p_hat"));
        std::lock_guard<std::mutex> lk(n.dev->mtx);
        gnn_predict(n, G, bb, out);
        NSFEM_HIP(hipStreamSynchronize(n.dev->stream));
    });
}

int nsfem_gnn_gradients(nsfem_gnn *g, nsfem_gnn_graph *gr, const nsfem_buffer *b,
                        const nsfem_buffer *p, float *grads, double *loss) {
    return guard([&] {
        GnnNet &n = net(g);
        GnnGraph &G = graph(gr, n);
        const double *bb = vec(b, G, "b");
        const double *pp = vec(p, G, "p");
        std::lock_guard<std::mutex> lk(n.dev->mtx);
        gnn_gradients(n, G, bb, pp, grads, loss);
    });
}

int nsfem_gnn_train_step(nsfem_gnn *g, nsfem_gnn_graph *gr, const nsfem_buffer *b,
                         const nsfem_buffer *p, float learning_rate, double *loss) {
    return guard([&] {
        GnnNet &n = net(g);
        GnnGraph &G = graph(gr, n);
        NSFEM_REQUIRE(learning_rate > 0.f, NSFEM_INVALID_ARGUMENT, "learning rate must be > 0");
        const double *bb = vec(b, G, "b");
        const double *pp = vec(p, G, "p");
        std::lock_guard<std::mutex> lk(n.dev->mtx);
        gnn_train_step(n, G, bb, pp, learning_rate, loss);
    });
}

}  // extern "C"
