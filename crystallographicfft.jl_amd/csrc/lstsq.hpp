// Host-side least squares for the Anderson mixing step (plain C++, no HIP):
// included by the SCFT driver and compiled on its own by the CPU tests.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace cfft {

// Minimum-norm least squares of U c = V for a symmetric positive
// semidefinite k x k matrix U (row-major): the eigenpairs of U by cyclic
// Jacobi rotations, then c = sum_j (v_j . V / lam_j) v_j over the eigenvalues
// above eps k max|lam| (the truncated-SVD solution that numpy.linalg.lstsq
// gives with rcond=None; for PSD U its singular values are |lam_j|).  False
// when no eigenvalue passes.
inline bool lstsq_sym(std::vector<double> U, const std::vector<double>& V, int k, std::vector<double>& c) {
  std::vector<double> Q(k * k, 0.0);
  for (int i = 0; i < k; ++i) Q[i * k + i] = 1.0;
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int i = 0; i < k; ++i) {
      diag += U[i * k + i] * U[i * k + i];
      for (int j = i + 1; j < k; ++j) off += U[i * k + j] * U[i * k + j];
    }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        const double apq = U[p * k + q];
        if (apq == 0.0) continue;
        const double theta = (U[q * k + q] - U[p * k + p]) / (2.0 * apq);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double cs = 1.0 / std::sqrt(t * t + 1.0), sn = t * cs;
        for (int r = 0; r < k; ++r) {  // columns p, q
          const double a = U[r * k + p], b = U[r * k + q];
          U[r * k + p] = cs * a - sn * b;
          U[r * k + q] = sn * a + cs * b;
        }
        for (int r = 0; r < k; ++r) {  // rows p, q
          const double a = U[p * k + r], b = U[q * k + r];
          U[p * k + r] = cs * a - sn * b;
          U[q * k + r] = sn * a + cs * b;
        }
        for (int r = 0; r < k; ++r) {  // eigenvector columns
          const double a = Q[r * k + p], b = Q[r * k + q];
          Q[r * k + p] = cs * a - sn * b;
          Q[r * k + q] = sn * a + cs * b;
        }
      }
  }
  double lmax = 0.0;
  for (int j = 0; j < k; ++j) lmax = std::max(lmax, std::fabs(U[j * k + j]));
  const double cut = std::numeric_limits<double>::epsilon() * k * lmax;
  c.assign(k, 0.0);
  bool any = false;
  for (int j = 0; j < k; ++j) {
    const double lam = std::fabs(U[j * k + j]);
    if (!(lam > cut)) continue;
    any = true;
    double pj = 0.0;
    for (int r = 0; r < k; ++r) pj += Q[r * k + j] * V[r];
    pj /= U[j * k + j];
    for (int r = 0; r < k; ++r) c[r] += pj * Q[r * k + j];
  }
  return any;
}

}  // namespace cfft
