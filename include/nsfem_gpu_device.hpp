// C++ face of the drop-in boundary: the member-function interface of the
// reference's GPUSimulation::GPUSimulationDevice
// (include/gpu_simulation_device.h:13-116), header-only over the C ABI of
// nsfem_gpu.h.  Method names, argument order and meaning follow the
// reference; every method returns an ErrorCode whose status is the nsfem
// status (0 = success) and whose message is nsfem_last_error().
//
// Differences a caller of the reference sees (all deliberate):
//   * device vectors (GPUBuffer) hold FP64 values; host arrays of the FP32
//     reference signatures are widened on the way in and rounded on the way
//     out, so the device arithmetic is the FP64 plain method;
//   * conjugateGradient always writes x, also when x0 already meets eps
//     (the reference returns early without writing it,
//     cpp/gpu_simulation_device.cpp:446-448), and reports
//     NSFEM_MAX_ITERATIONS_REACHED as a distinct status with x = last iterate;
//   * useMegakernel is accepted and ignored: the persistent kernel always
//     runs (the reference's multi-kernel variant, :293-383, is not rebuilt);
//   * loadModules has nothing to load (kernels are linked into the library)
//     and succeeds;
//   * uploadKDTree takes the mesh (node coordinates and P2 elements) and
//     builds the tree with the library's restatement of KDTreeBuilder::build
//     (cpp/kd_tree_builder.cpp:81-148), or takes prebuilt tree arrays.
#ifndef NSFEM_GPU_DEVICE_HPP
#define NSFEM_GPU_DEVICE_HPP

#include <algorithm>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "nsfem_gpu.h"

namespace nsfem {

// EC::ErrorCode (include/error_code.h:14-39): status + message.
class ErrorCode {
public:
    ErrorCode() : status_(NSFEM_OK) {}
    ErrorCode(int status, const char *message)
        : status_(status), message_(message ? message : "") {}
    int getStatus() const { return status_; }
    const char *getMessage() const { return message_.c_str(); }
    bool hasError() const { return status_ != NSFEM_OK; }

private:
    int status_;
    std::string message_;
};

inline ErrorCode make_error(int status) {
    return status == NSFEM_OK ? ErrorCode() : ErrorCode(status, nsfem_last_error());
}

// GPU::GPUBuffer (include/gpu_host_common.h:268-320): an owned device
// allocation; here it holds FP64 values.
class GPUBuffer {
public:
    GPUBuffer() = default;
    GPUBuffer(const GPUBuffer &) = delete;
    GPUBuffer &operator=(const GPUBuffer &) = delete;
    GPUBuffer(GPUBuffer &&o) noexcept : buf_(o.buf_) { o.buf_ = nullptr; }
    GPUBuffer &operator=(GPUBuffer &&o) noexcept {
        std::swap(buf_, o.buf_);
        return *this;
    }
    ~GPUBuffer() { freeMem(); }

    ErrorCode init(nsfem_device *dev, int64_t byteSize) {
        freeMem();
        return make_error(nsfem_buffer_create(dev, byteSize, &buf_));
    }
    ErrorCode uploadBuffer(const void *src, int64_t byteSize) {
        return make_error(nsfem_buffer_upload(buf_, src, byteSize));
    }
    ErrorCode downloadBuffer(void *dst, int64_t byteSize) const {
        return make_error(nsfem_buffer_download(buf_, dst, byteSize));
    }
    int64_t getByteSize() const { return buf_ ? nsfem_buffer_byte_size(buf_) : 0; }
    void freeMem() {
        if (buf_) nsfem_buffer_destroy(buf_);
        buf_ = nullptr;
    }
    nsfem_buffer *handle() const { return buf_; }

private:
    nsfem_buffer *buf_ = nullptr;
};

class GPUSimulationDevice {
public:
    // SimMatrix (include/gpu_simulation_device.h:16-23), same order and
    // spelling as the reference (pressureSitffness included).
    enum SimMatrix {
        pressureSitffness = NSFEM_PRESSURE_STIFFNESS,
        velocityMass = NSFEM_VELOCITY_MASS,
        diffusion = NSFEM_DIFFUSION,
        velocityDivergence = NSFEM_VELOCITY_DIVERGENCE,
        pressureDivergence = NSFEM_PRESSURE_DIVERGENCE,
        count
    };

    GPUSimulationDevice() = default;
    GPUSimulationDevice(const GPUSimulationDevice &) = delete;
    GPUSimulationDevice &operator=(const GPUSimulationDevice &) = delete;
    ~GPUSimulationDevice() {
        if (dev_) nsfem_device_destroy(dev_);
    }

    ErrorCode init(int index) {
        if (dev_) nsfem_device_destroy(dev_);
        dev_ = nullptr;
        return make_error(nsfem_device_create(index, &dev_));
    }
    ErrorCode loadModules(const char * /*advectionData*/, const char * /*sparseMatrixData*/) {
        return ErrorCode();
    }
    // uploadKDTree from the mesh: builds the tree on the host
    // (KDTreeBuilder::buildCPUOwner, cpp/kd_tree_builder.cpp:17-25) and
    // uploads it with the grid (KDTreeCPUOwner::upload, :36-79).
    ErrorCode uploadKDTree(int nodeCount, const float *xy, int elementCount,
                           const int32_t *elements, int maxDepth = -1, int minLeafSize = 16) {
        nsfem_kdtree *t = nullptr;
        int rc = nsfem_kdtree_build(nodeCount, xy, elementCount, elements, maxDepth,
                                    minLeafSize, &t);
        if (rc != NSFEM_OK) return make_error(rc);
        rc = nsfem_upload_kdtree(dev_, t, nodeCount, xy, elementCount, elements);
        ErrorCode ec = make_error(rc);
        nsfem_kdtree_destroy(t);
        return ec;
    }
    // uploadKDTree from prebuilt arrays in the reference's KDNode layout
    // (two 32-bit words per node, gpu_cpu_shared/kd_tree_common.cuh:63-154).
    ErrorCode uploadKDTree(int32_t treeNodeCount, const uint32_t *nodeWords,
                           int64_t leafIndexCount, const int32_t *leafIndices, int nodeCount,
                           const float *xy, int elementCount, const int32_t *elements) {
        return make_error(nsfem_upload_kdtree_arrays(dev_, treeNodeCount, nodeWords,
                                                     leafIndexCount, leafIndices, nodeCount,
                                                     xy, elementCount, elements));
    }
    ErrorCode initVelocityBuffers(const int numElements) {
        return make_error(nsfem_init_velocity_buffers(dev_, numElements));
    }
    // uploadMatrix(SimMatrix, const SMM::TripletMatrix<float>&)
    // (cpp/gpu_sparse_matrix.cpp:7-49): any triplet type with the SMM
    // interface (getDenseRowCount, getNonZeroCount, iteration yielding
    // getRow/getCol/getValue); duplicates are summed.
    template <typename Triplet>
    ErrorCode uploadMatrix(SimMatrix matrix, const Triplet &triplet, int cols = -1) {
        const int rows = triplet.getDenseRowCount();
        std::vector<int32_t> r, c;
        std::vector<double> v;
        r.reserve(triplet.getNonZeroCount());
        c.reserve(triplet.getNonZeroCount());
        v.reserve(triplet.getNonZeroCount());
        int maxc = -1;
        for (const auto &el : triplet) {
            r.push_back(el.getRow());
            c.push_back(el.getCol());
            v.push_back((double)el.getValue());
            maxc = std::max(maxc, (int)el.getCol());
        }
        if (cols < 0) cols = std::max(rows, maxc + 1);
        return uploadMatrix(matrix, rows, cols, (int64_t)r.size(), r.data(), c.data(), v.data());
    }
    ErrorCode uploadMatrix(SimMatrix matrix, int rows, int cols, int64_t count, const int32_t *row,
                           const int32_t *col, const double *value) {
        return make_error(nsfem_upload_matrix_triplet(dev_, (int)matrix, rows, cols, count, row,
                                                      col, value));
    }
    // advect (cpp/gpu_simulation_device.cpp:91-131): host FP32 in and out.
    ErrorCode advect(const int numElements, const float *uVelocity, const float *vVelocity,
                     const float dt, float *uVelocityOut, float *vVelocityOut) {
        return make_error(nsfem_advect(dev_, numElements, uVelocity, vVelocity, dt, uVelocityOut,
                                       vVelocityOut));
    }
    ErrorCode spRMult(SimMatrix matrix, const GPUBuffer &mult, GPUBuffer &res) {
        return make_error(nsfem_sp_r_mult(dev_, (int)matrix, mult.handle(), res.handle()));
    }
    ErrorCode spRMultSub(const SimMatrix matrix, const GPUBuffer &lhs, const GPUBuffer &mult,
                         GPUBuffer &res) {
        return make_error(
            nsfem_sp_r_mult_sub(dev_, (int)matrix, lhs.handle(), mult.handle(), res.handle()));
    }
    // conjugateGradient (include/gpu_simulation_device.h:62-76): host
    // arrays, absolute stopping rule ||r|| < eps, maxIterations -1 = rows.
    ErrorCode conjugateGradient(const SimMatrix matrix, const float *const b,
                                const float *const x0, float *const x, int maxIterations,
                                float eps, bool useMegakernel = true) {
        int rows = 0, cols = 0;
        int64_t nnz = 0;
        int rc = nsfem_matrix_info(dev_, (int)matrix, &rows, &cols, &nnz);
        if (rc != NSFEM_OK) return make_error(rc);
        std::vector<double> bd(b, b + rows), xd(x0, x0 + rows), out((size_t)rows);
        rc = nsfem_conjugate_gradient(dev_, (int)matrix, bd.data(), xd.data(), out.data(),
                                      maxIterations, (double)eps, useMegakernel ? 1 : 0);
        if (rc == NSFEM_OK || rc == NSFEM_MAX_ITERATIONS_REACHED)
            for (int i = 0; i < rows; ++i) x[i] = (float)out[i];
        return make_error(rc);
    }
    // FP64 overload: the plain method without rounding through float.
    ErrorCode conjugateGradient(const SimMatrix matrix, const double *const b,
                                const double *const x0, double *const x, int maxIterations,
                                double eps, bool useMegakernel = true) {
        return make_error(nsfem_conjugate_gradient(dev_, (int)matrix, b, x0, x, maxIterations,
                                                    eps, useMegakernel ? 1 : 0));
    }
    ErrorCode saxpby(const int vectorLength, const double a, const double b, const GPUBuffer &x,
                     const GPUBuffer &y, GPUBuffer &result) {
        return make_error(
            nsfem_saxpby(dev_, vectorLength, a, b, x.handle(), y.handle(), result.handle()));
    }
    ErrorCode saxpy(const int vectorLength, const double a, const GPUBuffer &x, GPUBuffer &y) {
        return make_error(nsfem_saxpy(dev_, vectorLength, a, x.handle(), y.handle()));
    }
    // dotProduct adds a.b to result[0] (the caller zeroes it first, as in
    // the reference, gpu/matrix_math.cu:162-194).
    ErrorCode dotProduct(const int vectorLength, const GPUBuffer &a, const GPUBuffer &b,
                         GPUBuffer &result) {
        return make_error(
            nsfem_dot_product(dev_, vectorLength, a.handle(), b.handle(), result.handle()));
    }

    // Not in the reference: a buffer bound to this device, and the handle
    // for the wider C ABI (nsfem_cg_solve, nsfem_simulation_*).
    ErrorCode makeBuffer(GPUBuffer &buf, int64_t byteSize) { return buf.init(dev_, byteSize); }
    nsfem_device *handle() const { return dev_; }

private:
    nsfem_device *dev_ = nullptr;
};

}  // namespace nsfem

#endif
