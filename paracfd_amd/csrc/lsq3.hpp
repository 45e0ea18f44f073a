// Least-squares gradient support of the special cells (host, built once per solver).
#pragma once

#include <algorithm>
#include <cmath>

namespace pcfd {

// Pseudo-inverse of a symmetric 3x3 matrix: cyclic Jacobi rotations to diagonal form, then
// the sum over modes whose eigenvalue exceeds 1e-4 of the largest of v v^T / lambda (the
// rank rule of amr_advection.cu:822-831, which drops directions the stencil does not sample).
inline void symmetric_pinv3(const double* A, double* out) {
    double a[3][3], v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a[r][c] = A[3 * r + c];
    for (int sweep = 0; sweep < 64; ++sweep) {
        const double off = std::abs(a[0][1]) + std::abs(a[0][2]) + std::abs(a[1][2]);
        const double dia = std::abs(a[0][0]) + std::abs(a[1][1]) + std::abs(a[2][2]);
        if (!(off > 1e-300) || off <= 1e-17 * dia) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                if (a[p][q] == 0.0) continue;
                // rotation angle that annihilates a[p][q] (the smaller of the two roots)
                const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
                const double t = std::copysign(1.0, theta) / (std::abs(theta) + std::sqrt(theta * theta + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double kp = a[k][p], kq = a[k][q];
                    a[k][p] = c * kp - s * kq;
                    a[k][q] = s * kp + c * kq;
                }
                for (int k = 0; k < 3; ++k) {
                    const double pk = a[p][k], qk = a[q][k];
                    a[p][k] = c * pk - s * qk;
                    a[q][k] = s * pk + c * qk;
                }
                for (int k = 0; k < 3; ++k) {
                    const double kp = v[k][p], kq = v[k][q];
                    v[k][p] = c * kp - s * kq;
                    v[k][q] = s * kp + c * kq;
                }
            }
    }
    const double top = std::max({std::abs(a[0][0]), std::abs(a[1][1]), std::abs(a[2][2])});
    for (int k = 0; k < 9; ++k) out[k] = 0.0;
    for (int m = 0; m < 3; ++m) {
        const double l = a[m][m];
        if (!(l > 1e-4 * top)) continue;
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) out[3 * r + c] += v[r][m] * v[c][m] / l;
    }
}

}  // namespace pcfd
