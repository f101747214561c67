// Drives libnsfem_hip.so through include/nsfem_gpu_device.hpp the way the
// reference's solver drives GPUSimulation::GPUSimulationDevice
// (include/assembly.h:1302-1557): triplet upload, products, BLAS-1, the
// CG call with an absolute eps, and advection on a small P2 mesh.  Checks
// are closed forms computed here on the host.  Built as a program (exit
// status 0 = all passed) and as a shared library whose
// nsfem_device_shim_selftest() the GPU test calls in its own process.
#include <cmath>
#include <cstdio>
#include <vector>

#include "nsfem_gpu_device.hpp"

using nsfem::ErrorCode;
using nsfem::GPUBuffer;
using Dev = nsfem::GPUSimulationDevice;

static int failures = 0;
#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                      \
            std::printf("\n");                             \
            ++failures;                                    \
        }                                                  \
    } while (0)
#define OK(ec) CHECK(!(ec).hasError(), "%s", (ec).getMessage())

// The SMM::TripletMatrix<float> surface the reference's upload uses
// (cpp/gpu_sparse_matrix.cpp:7-49).
struct Triplet {
    struct El {
        int r, c;
        float v;
        int getRow() const { return r; }
        int getCol() const { return c; }
        float getValue() const { return v; }
    };
    int rows = 0;
    std::vector<El> els;
    int getDenseRowCount() const { return rows; }
    int getNonZeroCount() const { return (int)els.size(); }
    std::vector<El>::const_iterator begin() const { return els.begin(); }
    std::vector<El>::const_iterator end() const { return els.end(); }
};

// 5-point Laplacian on a k x k interior grid (SPD), entries listed twice
// with half values to exercise duplicate summation.
static Triplet laplacian(int k) {
    Triplet t;
    t.rows = k * k;
    for (int j = 0; j < k; ++j)
        for (int i = 0; i < k; ++i) {
            const int r = j * k + i;
            t.els.push_back({r, r, 2.0f});
            t.els.push_back({r, r, 2.0f});
            if (i > 0) t.els.push_back({r, r - 1, -1.0f});
            if (i < k - 1) t.els.push_back({r, r + 1, -1.0f});
            if (j > 0) t.els.push_back({r, r - k, -1.0f});
            if (j < k - 1) t.els.push_back({r, r + k, -1.0f});
        }
    return t;
}

static std::vector<double> host_mult(const Triplet &t, const std::vector<double> &x) {
    std::vector<double> y(t.rows, 0.0);
    for (const auto &e : t.els) y[e.r] += (double)e.v * x[e.c];
    return y;
}

static int run_all() {
    failures = 0;
    Dev dev;
    ErrorCode ec = dev.init(0);
    if (ec.hasError()) {
        std::printf("FAIL init: %s\n", ec.getMessage());
        return 1;
    }
    OK(dev.loadModules(nullptr, nullptr));
    const int k = 40, n = k * k;
    const Triplet A = laplacian(k);
    OK(dev.uploadMatrix(Dev::pressureSitffness, A));

    std::vector<double> x(n), lhs(n);
    for (int i = 0; i < n; ++i) {
        x[i] = std::sin(0.37 * i) + 0.25;
        lhs[i] = std::cos(0.11 * i);
    }
    GPUBuffer bx, by, bl, bres;
    OK(dev.makeBuffer(bx, 8 * n));
    OK(dev.makeBuffer(by, 8 * n));
    OK(dev.makeBuffer(bl, 8 * n));
    OK(dev.makeBuffer(bres, 8));
    OK(bx.uploadBuffer(x.data(), 8 * n));
    OK(bl.uploadBuffer(lhs.data(), 8 * n));

    // spRMult / spRMultSub against the host product
    const std::vector<double> ax = host_mult(A, x);
    std::vector<double> y(n);
    OK(dev.spRMult(Dev::pressureSitffness, bx, by));
    OK(by.downloadBuffer(y.data(), 8 * n));
    double err = 0;
    for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(y[i] - ax[i]));
    CHECK(err < 1e-12, "spRMult max error %g", err);
    OK(dev.spRMultSub(Dev::pressureSitffness, bl, bx, by));
    OK(by.downloadBuffer(y.data(), 8 * n));
    err = 0;
    for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(y[i] - (lhs[i] - ax[i])));
    CHECK(err < 1e-12, "spRMultSub max error %g", err);

    // saxpy, saxpby, dotProduct (result accumulates into a zeroed buffer)
    OK(by.uploadBuffer(lhs.data(), 8 * n));
    OK(dev.saxpy(n, 2.5, bx, by));
    OK(by.downloadBuffer(y.data(), 8 * n));
    err = 0;
    for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(y[i] - (lhs[i] + 2.5 * x[i])));
    CHECK(err < 1e-14, "saxpy max error %g", err);
    OK(dev.saxpby(n, 2.0, -3.0, bx, bl, by));
    OK(by.downloadBuffer(y.data(), 8 * n));
    err = 0;
    for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(y[i] - (2.0 * x[i] - 3.0 * lhs[i])));
    CHECK(err < 1e-14, "saxpby max error %g", err);
    const double zero = 0.0;
    OK(bres.uploadBuffer(&zero, 8));
    OK(dev.dotProduct(n, bx, bl, bres));
    double dot = 0, dref = 0;
    OK(bres.downloadBuffer(&dot, 8));
    for (int i = 0; i < n; ++i) dref += x[i] * lhs[i];
    CHECK(std::fabs(dot - dref) < 1e-11 * std::fabs(dref) + 1e-12, "dot %g vs %g", dot, dref);

    // conjugateGradient with the reference's absolute eps
    std::vector<float> bf(n), x0f(n, 0.0f), xf(n);
    for (int i = 0; i < n; ++i) bf[i] = (float)std::sin(0.05 * i);
    ec = dev.conjugateGradient(Dev::pressureSitffness, bf.data(), x0f.data(), xf.data(), -1, 1e-5f);
    OK(ec);
    std::vector<double> xd(xf.begin(), xf.end());
    std::vector<double> r = host_mult(A, xd);
    double rn = 0;
    for (int i = 0; i < n; ++i) rn += (bf[i] - r[i]) * (bf[i] - r[i]);
    // x is rounded to float on the way out: allow the rounding's share
    CHECK(std::sqrt(rn) < 1e-5 + 8.0 * 6e-8 * n, "CG residual %g", std::sqrt(rn));
    std::vector<double> bd(bf.begin(), bf.end()), x0d(n, 0.0), xo(n);
    OK(dev.conjugateGradient(Dev::pressureSitffness, bd.data(), x0d.data(), xo.data(), -1, 1e-9));
    r = host_mult(A, xo);
    rn = 0;
    for (int i = 0; i < n; ++i) rn += (bd[i] - r[i]) * (bd[i] - r[i]);
    CHECK(std::sqrt(rn) < 1e-9, "FP64 CG residual %g", std::sqrt(rn));
    // iteration limit: distinct status, x written
    std::fill(xo.begin(), xo.end(), 0.0);
    ec = dev.conjugateGradient(Dev::pressureSitffness, bd.data(), x0d.data(), xo.data(), 3, 1e-12);
    CHECK(ec.getStatus() == NSFEM_MAX_ITERATIONS_REACHED, "status %d", ec.getStatus());
    double xn = 0;
    for (double v : xo) xn += v * v;
    CHECK(xn > 0, "x not written at the iteration limit");
    // x0 already converged: x = x0 is written
    ec = dev.conjugateGradient(Dev::pressureSitffness, bd.data(), xo.data(), x0d.data(), -1, 1e30);
    OK(ec);
    CHECK(x0d[7] == xo[7], "x not written when x0 already converged");

    // advect on a unit square of two P2 triangles, zero velocity: identity
    const float xy[] = {0, 0, 1, 0, 1, 1, 0, 1, 1, 0.5f, 0.5f, 0.5f, 0.5f, 0, 0.5f, 1, 0, 0.5f};
    // vertices 0..3, midpoints 4 (1-2), 5 (0-2), 6 (0-1), 7 (2-3), 8 (0-3)
    const int32_t el[] = {0, 1, 2, 4, 5, 6, 0, 2, 3, 7, 8, 5};
    OK(dev.uploadKDTree(9, xy, 2, el));
    OK(dev.initVelocityBuffers(9));
    float u[9], v[9], uo[9], vo[9];
    for (int i = 0; i < 9; ++i) {
        u[i] = 0.0f;
        v[i] = 0.0f;
    }
    OK(dev.advect(9, u, v, 0.1f, uo, vo));
    for (int i = 0; i < 9; ++i) CHECK(uo[i] == 0.0f && vo[i] == 0.0f, "advect node %d", i);
    // uniform velocity (c, 0): the constant field is reproduced exactly
    for (int i = 0; i < 9; ++i) u[i] = 0.25f;
    OK(dev.advect(9, u, v, 0.5f, uo, vo));
    for (int i = 0; i < 9; ++i) CHECK(std::fabs(uo[i] - 0.25f) < 1e-6f, "advect const node %d", i);

    // errors carry a status and a message
    ec = dev.spRMult(Dev::velocityDivergence, bx, by);  // slot never uploaded
    CHECK(ec.hasError() && ec.getMessage()[0] != 0, "empty slot accepted");

    std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
    std::fflush(stdout);
    return failures;
}

#ifdef NSFEM_SHIM_LIBRARY
// loaded into the test process (ctypes): returns the number of failures
extern "C" int nsfem_device_shim_selftest() { return run_all(); }
#else
int main() { return run_all() ? 1 : 0; }
#endif
