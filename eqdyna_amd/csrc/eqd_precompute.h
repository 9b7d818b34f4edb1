// One-time element precompute of the reference host (the element part of
// src/assembleGlobalMass.f90): shape-function derivatives and Jacobian at the
// element centre, lumped mass, and the hourglass operators SS and Phi of one
// hexahedron (or wedge).  Plain arithmetic in the reference's operation
// order, compiled both for the host case builder (g++) and for the device
// precompute kernel (hipcc), each with -ffp-contract=off, so both give the
// reference's rounded results.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define EQD_HD __host__ __device__
#else
#define EQD_HD
#endif

namespace eqd {

// vlm, src/library.f90:60-93
EQD_HD inline double vlm(const double xl[3][8])
{
    constexpr int it[8][8] = {{1, 2, 3, 4, 5, 6, 7, 8}, {2, 3, 4, 1, 6, 7, 8, 5}, {3, 4, 1, 2, 7, 8, 5, 6},
                              {4, 1, 2, 3, 8, 5, 6, 7}, {5, 8, 7, 6, 1, 4, 3, 2}, {6, 5, 8, 7, 2, 1, 4, 3},
                              {7, 6, 5, 8, 3, 2, 1, 4}, {8, 7, 6, 5, 4, 3, 2, 1}};  // it[column][row]
    double bb[8];
    for (int i = 0; i < 8; i++) {
        auto Y = [&](int r) { return xl[1][it[i][r - 1] - 1]; };
        auto Z = [&](int r) { return xl[2][it[i][r - 1] - 1]; };
        bb[i] = Y(2) * (Z(6) - Z(3) + Z(5) - Z(4)) + Y(3) * (Z(2) - Z(4)) + Y(4) * (Z(3) - Z(8) + Z(2) - Z(5)) +
                Y(5) * (Z(8) - Z(6) + Z(4) - Z(2)) + Y(6) * (Z(5) - Z(2)) + Y(8) * (Z(4) - Z(5));
    }
    double v = 0.0;
    for (int i = 0; i < 8; i++) v = v + xl[0][i] * bb[i];
    return v / 12.0;
}

// calcGlobalShapeFunc + contm + calcSSPhi4Hrgls for one hexahedron,
// src/calcGlobalShapeFunc.f90:1-75, src/assembleGlobalMass.f90:328-406.
// xl[j][a]: coordinate j of local node a.  Returns false on a non-positive
// Jacobian (src/calcGlobalShapeFunc.f90:52-57).
EQD_HD inline bool elem_precompute(const double xl[3][8], bool wedge, double rho, double lam, double mu, double w,
                                   double shp[24], double *detp, double emass[24], double ss[6], double phi[32])
{
    constexpr double ac[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                 {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
    const double cst = 1.0 / 8.0;
    double g[4][8];
    for (int i = 0; i < 8; i++) {
        g[3][i] = cst;
        for (int j = 0; j < 3; j++) g[j][i] = cst * ac[i][j];
    }
    if (wedge) {
        // degenerate to a wedge: columns 3+4 and 7+8 collapse (src/calcGlobalShapeFunc.f90:19-28)
        for (int i = 0; i < 4; i++) {
            g[i][2] = g[i][2] + g[i][3];
            g[i][3] = 0.0;
            g[i][6] = g[i][6] + g[i][7];
            g[i][7] = 0.0;
        }
    }
    double xs[3][3];  // xs[j][i] = d x_j / d xi_i
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double t = 0.0;
            for (int k = 0; k < 8; k++) t = t + g[i][k] * xl[j][k];
            xs[j][i] = t;
        }
    const double cof11 = xs[1][1] * xs[2][2] - xs[1][2] * xs[2][1];
    const double cof12 = xs[1][2] * xs[2][0] - xs[1][0] * xs[2][2];
    const double cof13 = xs[1][0] * xs[2][1] - xs[1][1] * xs[2][0];
    const double cof21 = xs[2][1] * xs[0][2] - xs[2][2] * xs[0][1];
    const double cof22 = xs[2][2] * xs[0][0] - xs[2][0] * xs[0][2];
    const double cof23 = xs[2][0] * xs[0][1] - xs[2][1] * xs[0][0];
    const double cof31 = xs[0][1] * xs[1][2] - xs[0][2] * xs[1][1];
    const double cof32 = xs[0][2] * xs[1][0] - xs[0][0] * xs[1][2];
    const double cof33 = xs[0][0] * xs[1][1] - xs[0][1] * xs[1][0];
    const double det = xs[0][0] * cof11 + xs[0][1] * cof12 + xs[0][2] * cof13;
    if (det <= 0.0) return false;
    for (int i = 0; i < 8; i++) {
        const double t1 = g[0][i], t2 = g[1][i], t3 = g[2][i];
        shp[i * 3 + 0] = (t1 * cof11 + t2 * cof12 + t3 * cof13) / det;
        shp[i * 3 + 1] = (t1 * cof21 + t2 * cof22 + t3 * cof23) / det;
        shp[i * 3 + 2] = (t1 * cof31 + t2 * cof32 + t3 * cof33) / det;
    }
    // returned xs = reshape(cof)/det: xr[r][c] with columns (cof11,cof12,cof13), (cof21..), (cof31..)
    const double xr[3][3] = {{cof11 / det, cof21 / det, cof31 / det},
                             {cof12 / det, cof22 / det, cof32 / det},
                             {cof13 / det, cof23 / det, cof33 / det}};
    *detp = det;
    // contm
    const double totmas = rho * w * det;
    double dsum = 0.0, work[8];
    for (int j = 0; j < 8; j++) {
        const double t2 = totmas * (g[3][j] * g[3][j]);
        dsum = dsum + t2;
        work[j] = 0.0 + t2;
    }
    const double t1 = totmas / dsum;
    for (int j = 0; j < 8; j++) {
        const double t2 = t1 * work[j];
        for (int k = 0; k < 3; k++) emass[j * 3 + k] = t2;
    }
    // calcSSPhi4Hrgls
    const double vol = vlm(xl);
    double ce = mu * (3 * lam + 2 * mu) / (lam + mu);
    ce = 16.0 * ce / 15.0;
    const double co = ce * vol / 48.0;
    auto X = [&](int r, int c) { return xr[r - 1][c - 1]; };
    ss[0] = co * (X(1, 1) * X(1, 1) + X(2, 1) * X(2, 1) + X(3, 1) * X(3, 1));
    ss[1] = co * (X(1, 1) * X(1, 2) + X(2, 1) * X(2, 2) + X(3, 1) * X(3, 2));
    ss[2] = co * (X(1, 1) * X(1, 3) + X(2, 1) * X(2, 3) + X(3, 1) * X(3, 3));
    ss[3] = co * (X(1, 2) * X(1, 2) + X(2, 2) * X(2, 2) + X(3, 2) * X(3, 2));
    ss[4] = co * (X(1, 2) * X(1, 3) + X(2, 2) * X(2, 3) + X(3, 2) * X(3, 3));
    ss[5] = co * (X(1, 3) * X(1, 3) + X(2, 3) * X(2, 3) + X(3, 3) * X(3, 3));
    constexpr int ha[4][8] = {{1, 1, -1, -1, -1, -1, 1, 1}, {1, -1, -1, 1, -1, 1, 1, -1},
                              {1, -1, 1, -1, 1, -1, 1, -1}, {-1, 1, -1, 1, 1, -1, 1, -1}};
    for (int m = 0; m < 4; m++) {
        double ph[8];
        for (int j = 0; j < 8; j++) {
            double v = 0.0;
            for (int k = 0; k < 8; k++)
                v = v + (double)ha[m][k] * (xl[0][k] * shp[j * 3] + xl[1][k] * shp[j * 3 + 1] + xl[2][k] * shp[j * 3 + 2]);
            ph[j] = (double)ha[m][j] - v;
        }
        double v = 0.0;
        for (int j = 0; j < 8; j++) v = v + ph[j] * ph[j];
        v = sqrt(v / 8.0);
        for (int j = 0; j < 8; j++) phi[m * 8 + j] = ph[j] / v;
    }
    return true;
}

}  // namespace eqd
