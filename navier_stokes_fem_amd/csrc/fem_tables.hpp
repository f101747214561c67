// FEM tables and element geometry shared by the host assembly
// (assembly_host.cpp) and the device assembly (assembly.hip), so that both
// compute every local matrix entry with the same operations in the same
// order (both translation units are built without FMA contraction).
//
// Local matrices follow the reference functors exactly:
//   mass      LocalMassFunctor        (include/assembly.h:842-882) shapeSquared * |J|
//   stiffness LocalStiffnessFunctor   (:725-791) delPSq with B / |J|
//   divergence LocalDivergenceFunctor (:793-840)
// with the 8-point triangle rule of TriangleIntegrator (:259-304).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#ifndef NSFEM_HD
#ifdef __HIPCC__
#define NSFEM_HD __host__ __device__
#else
#define NSFEM_HD
#endif
#endif

namespace nsfem {
namespace fem {

const double QX[8] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0, 1.0 / 3.0, 0.0};
const double QY[8] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5, 1.0 / 3.0, 0.0};
const double QW[8] = {3.0 / 120.0, 3.0 / 120.0, 3.0 / 120.0, 8.0 / 120.0,
                      8.0 / 120.0, 8.0 / 120.0, 27.0 / 120.0, 0.0};

inline void p2_eval(double xi, double eta, double o[6]) {
    o[0] = 1 - 3 * xi - 3 * eta + 2 * xi * xi + 4 * xi * eta + 2 * eta * eta;
    o[1] = 2 * xi * xi - xi;
    o[2] = 2 * eta * eta - eta;
    o[3] = 4 * xi * eta;
    o[4] = 4 * eta - 4 * xi * eta - 4 * eta * eta;
    o[5] = 4 * xi - 4 * xi * xi - 4 * xi * eta;
}
inline void p2_del(double xi, double eta, double o[2][6]) {
    o[0][0] = -3 + 4 * eta + 4 * xi;
    o[0][1] = -1 + 4 * xi;
    o[0][2] = 0.0;
    o[0][3] = 4 * eta;
    o[0][4] = -4 * eta;
    o[0][5] = 4 - 4 * eta - 8 * xi;
    o[1][0] = -3 + 4 * eta + 4 * xi;
    o[1][1] = 0;
    o[1][2] = -1 + 4 * eta;
    o[1][3] = 4 * xi;
    o[1][4] = 4 - 8 * eta - 4 * xi;
    o[1][5] = -4 * xi;
}
inline void p1_eval(double xi, double eta, double o[3]) {
    o[0] = 1 - xi - eta;
    o[1] = xi;
    o[2] = eta;
}
inline void p1_del(double o[2][3]) {
    o[0][0] = -1.0; o[0][1] = 1.0; o[0][2] = 0.0;
    o[1][0] = -1.0; o[1][1] = 0.0; o[1][2] = 1.0;
}

struct Tables {
    double mass2[6][6];
    double delsq2[12][12];
    double delsq1[6][6];
    double comb_vd[3][12];
    double comb_pd[6][6];
    Tables() {
        std::memset(this, 0, sizeof(*this));
        for (int q = 0; q < 8; ++q) {
            const double x = QX[q], y = QY[q], w = QW[q];
            double f2[6], d2[2][6], f1[3], d1[2][3];
            p2_eval(x, y, f2);
            p2_del(x, y, d2);
            p1_eval(x, y, f1);
            p1_del(d1);
            const double *d2f = &d2[0][0];
            const double *d1f = &d1[0][0];
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) mass2[i][j] += (f2[i] * f2[j]) * w;
            for (int i = 0; i < 12; ++i)
                for (int j = 0; j < 12; ++j) delsq2[i][j] += (d2f[i] * d2f[j]) * w;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 6; ++j) delsq1[i][j] += (d1f[i] * d1f[j]) * w;
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 6; ++j)
                    for (int k = 0; k < 2; ++k) comb_vd[i][j + k * 6] += (f1[i] * d2[k][j]) * w;
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 2; ++k) comb_pd[i][j + k * 3] += (f2[i] * d1[k][j]) * w;
        }
    }
};

struct Geo {
    double B[2][2];
    double jinv;  // 1/|det B|
    double jabs;  // |linTriangleTmJacobian|
};

NSFEM_HD inline Geo geometry(const float *xy, const int32_t *el) {
    const double xk = xy[2 * el[0]], yk = xy[2 * el[0] + 1];
    const double xl = xy[2 * el[1]], yl = xy[2 * el[1] + 1];
    const double xm = xy[2 * el[2]], ym = xy[2 * el[2] + 1];
    Geo g;
    g.B[0][0] = ym - yk;
    g.B[0][1] = yk - yl;
    g.B[1][0] = xk - xm;
    g.B[1][1] = xl - xk;
    const double det = g.B[0][0] * g.B[1][1] - g.B[0][1] * g.B[1][0];
    g.jinv = 1.0 / fabs(det);
    const double a = xl - xk, b = yl - yk, c = xm - xk, d = ym - yk;
    g.jabs = fabs(a * d - b * c);
    return g;
}

// Row i of a local stiffness matrix (LocalStiffnessFunctor::operator()).
NSFEM_HD inline double stiff(const double *delsq, int S, const Geo &g, int i, int j) {
    const int W = 2 * S;
    const double sq0 = delsq[i * W + j];
    const double sq1 = delsq[i * W + S + j];
    const double sq2 = delsq[(S + i) * W + j];
    const double sq3 = delsq[(S + i) * W + S + j];
    double acc = 0.0;
    for (int k = 0; k < 2; ++k) {
        acc += sq0 * g.B[k][0] * g.B[k][0] + sq1 * g.B[k][0] * g.B[k][1] +
               sq2 * g.B[k][0] * g.B[k][1] + sq3 * g.B[k][1] * g.B[k][1];
    }
    return acc * g.jinv;
}

}  // namespace fem
}  // namespace nsfem
