// FP64 dense symmetric eigensolver for the small Davidson subspace problem
// (the reference uses Eigen::SelfAdjointEigenSolver at
// xtp/src/libxtp/davidsonsolver.cc:221-239). Householder reduction to tridiagonal
// form followed by the implicit symmetric QR iteration with Wilkinson shifts
// (Golub & Van Loan, "Matrix Computations", 4th ed., Alg. 8.3.1 / 8.3.3).
// Eigenvalues ascending; eigenvectors are the columns of a column-major matrix.
#pragma once
#include <algorithm>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace xg {

// A: n x n column-major symmetric (only read). On return w[n] ascending and
// Z (n x n column-major) with A Z = Z diag(w). Returns false if QR did not converge.
inline bool sym_eig(int n, const double* A, int lda, double* w, double* Z) {
  if (n <= 0) return true;
  std::vector<double> a((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[(size_t)j * n + i] = 0.5 * (A[(size_t)j * lda + i] + A[(size_t)i * lda + j]);
  // Q accumulates the Householder reflections
  std::vector<double> q((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
  std::vector<double> v(n), p(n), wv(n);
  auto at = [&](int i, int j) -> double& { return a[(size_t)j * n + i]; };
  for (int k = 0; k + 2 < n; ++k) {
    // x = a[k+1:n, k]
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += at(i, k) * at(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    double x0 = at(k + 1, k);
    if (x0 > 0) alpha = -alpha;  // reflect onto -sign(x0)*|x| e1 (no cancellation)
    // v = x - alpha e1
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = at(i, k);
    v[k + 1] -= alpha;
    double vnorm2 = 0.0;
    for (int i = k + 1; i < n; ++i) vnorm2 += v[i] * v[i];
    if (vnorm2 == 0.0) continue;
    double beta = 2.0 / vnorm2;
    // p = beta * A v (trailing block rows/cols k..n-1 where v is non-zero from k+1)
    for (int i = k; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += at(i, j) * v[j];
      p[i] = beta * s;
    }
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
    double c = 0.5 * beta * pv;
    for (int i = k; i < n; ++i) wv[i] = p[i] - c * v[i];
    // A <- A - v w^T - w v^T  (v[k] = 0)
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i) at(i, j) -= v[i] * wv[j] + wv[i] * v[j];
    // Q <- Q H, H = I - beta v v^T
    for (int i = 0; i < n; ++i) {
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += q[(size_t)j * n + i] * v[j];
      s *= beta;
      for (int j = k + 1; j < n; ++j) q[(size_t)j * n + i] -= s * v[j];
    }
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);

  const double eps = 2.220446049250313e-16;
  int h = n - 1;
  int iter = 0;
  const int maxiter = 60 * n + 100;
  while (h > 0) {
    if (std::fabs(e[h - 1]) <= eps * (std::fabs(d[h - 1]) + std::fabs(d[h]))) {
      e[h - 1] = 0.0;
      --h;
      continue;
    }
    int l = h - 1;
    while (l > 0 && std::fabs(e[l - 1]) > eps * (std::fabs(d[l - 1]) + std::fabs(d[l]))) --l;
    if (++iter > maxiter) return false;
    // Wilkinson shift from the trailing 2x2 block
    double t = 0.5 * (d[h - 1] - d[h]);
    double ee = e[h - 1] * e[h - 1];
    double den = t + (t >= 0 ? 1.0 : -1.0) * std::sqrt(t * t + ee);
    double mu = d[h] - (den != 0.0 ? ee / den : 0.0);
    double x = d[l] - mu;
    double z = e[l];
    for (int k = l; k < h; ++k) {
      double r = std::hypot(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) { c = x / r; s = z / r; }
      if (k > l) e[k - 1] = r;
      double a0 = d[k], b0 = e[k], c0 = d[k + 1];
      d[k] = c * c * a0 + 2.0 * c * s * b0 + s * s * c0;
      d[k + 1] = s * s * a0 - 2.0 * c * s * b0 + c * c * c0;
      e[k] = c * s * (c0 - a0) + (c * c - s * s) * b0;
      if (k + 1 < h) {
        z = s * e[k + 1];
        e[k + 1] = c * e[k + 1];
      }
      x = e[k];
      double* q0 = &q[(size_t)k * n];
      double* q1 = &q[(size_t)(k + 1) * n];
      for (int i = 0; i < n; ++i) {
        double u0 = q0[i], u1 = q1[i];
        q0[i] = c * u0 + s * u1;
        q1[i] = -s * u0 + c * u1;
      }
    }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int i, int j) { return d[i] < d[j]; });
  for (int j = 0; j < n; ++j) {
    w[j] = d[idx[j]];
    for (int i = 0; i < n; ++i) Z[(size_t)j * n + i] = q[(size_t)idx[j] * n + i];
  }
  return true;
}

}  // namespace xg
