// Dense inverse of a small symmetric positive definite matrix given as CSR
// (the lower triangle is read), by Cholesky A = L L^T and the solves
// L y = e_c, L^T x = y for every column c.  Host code, O(n^3): meant for the
// deepest multigrid level (up to about a thousand nodes: about a second at 1089).
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

namespace nsfem {

// Fills inv (row-major n x n) and returns true, or returns false when a
// pivot is not positive (the matrix is not positive definite).
inline bool spd_inverse(int n, const int32_t *rs, const int32_t *col, const double *val,
                        std::vector<double> &inv) {
    std::vector<double> L((size_t)n * n, 0.0);  // row-major lower triangle
    for (int r = 0; r < n; ++r)
        for (int32_t k = rs[r]; k < rs[r + 1]; ++k)
            if (col[k] <= r) L[(size_t)r * n + col[k]] = val[k];
    for (int j = 0; j < n; ++j) {
        double dj = L[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) dj -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
        if (!(dj > 0.0)) return false;
        dj = std::sqrt(dj);
        L[(size_t)j * n + j] = dj;
        for (int i = j + 1; i < n; ++i) {
            double v = L[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
            L[(size_t)i * n + j] = v / dj;
        }
    }
    // L^T row-major, so that the backward solve reads rows too; the inverse
    // is symmetric, so column c of it is stored as row c
    std::vector<double> LT((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int k = 0; k <= i; ++k) LT[(size_t)k * n + i] = L[(size_t)i * n + k];
    inv.assign((size_t)n * n, 0.0);
    std::vector<double> y((size_t)n);
    for (int c = 0; c < n; ++c) {
        double *x = inv.data() + (size_t)c * n;
        // y = L^-1 e_c vanishes above row c
        for (int i = 0; i < c; ++i) y[i] = 0.0;
        for (int i = c; i < n; ++i) {
            double v = i == c ? 1.0 : 0.0;
            const double *Li = L.data() + (size_t)i * n;
            for (int k = c; k < i; ++k) v -= Li[k] * y[k];
            y[i] = v / Li[i];
        }
        for (int i = n - 1; i >= 0; --i) {
            double v = y[i];
            const double *Ui = LT.data() + (size_t)i * n;
            for (int k = i + 1; k < n; ++k) v -= Ui[k] * x[k];
            x[i] = v / Ui[i];
        }
    }
    return true;
}

}  // namespace nsfem
