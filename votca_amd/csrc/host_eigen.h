// FP64 dense symmetric eigensolver for the small Davidson subspace problem
// (the reference uses Eigen::SelfAdjointEigenSolver at
// xtp/src/libxtp/davidsonsolver.cc:221-239). Householder reduction to tridiagonal
// form followed by the implicit symmetric QR iteration with Wilkinson shifts
// (Golub & Van Loan, "Matrix Computations", 4th ed., Alg. 8.3.1 / 8.3.3).
// Eigenvalues ascending; eigenvectors are the columns of a column-major matrix.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
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
  std::vector<double> v(n), p(n), wv(n), tq(n);
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
    // p = beta * A v (trailing block rows/cols k..n-1 where v is non-zero from k+1), swept
    // column by column so every access is contiguous
    for (int i = k; i < n; ++i) p[i] = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      const double* col = &a[(size_t)j * n];
      for (int i = k; i < n; ++i) p[i] += col[i] * vj;
    }
    for (int i = k; i < n; ++i) p[i] *= beta;
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
    double c = 0.5 * beta * pv;
    for (int i = k; i < n; ++i) wv[i] = p[i] - c * v[i];
    // A <- A - v w^T - w v^T  (v[k] = 0)
    for (int j = k; j < n; ++j)
      for (int i = k; i < n; ++i) at(i, j) -= v[i] * wv[j] + wv[i] * v[j];
    // Q <- Q H, H = I - beta v v^T: t = Q v, then Q -= beta t v^T (column sweeps)
    for (int i = 0; i < n; ++i) tq[i] = 0.0;
    for (int j = k + 1; j < n; ++j) {
      const double vj = v[j];
      const double* col = &q[(size_t)j * n];
      for (int i = 0; i < n; ++i) tq[i] += col[i] * vj;
    }
    for (int j = k + 1; j < n; ++j) {
      const double f = beta * v[j];
      double* col = &q[(size_t)j * n];
      for (int i = 0; i < n; ++i) col[i] -= tq[i] * f;
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

// The `need` lowest eigenpairs only, for the Rayleigh-Ritz step of the Davidson loop:
// Householder tridiagonalisation with the reflectors kept (no explicit Q), all eigenvalues of
// the tridiagonal matrix by the implicit QR iteration without vectors, eigenvectors of the
// tridiagonal matrix for the wanted ones by inverse iteration with partial pivoting
// (Wilkinson's tinvit scheme: Golub & Van Loan 8.2.2 / Wilkinson & Reinsch II/18), vectors
// of close eigenvalues re-orthogonalised against each other, then the reflectors applied to
// the `need` vectors. O(n^3 / 3 + n^2 need) instead of O(8 n^3) for the full decomposition.
// w[need] ascending, Z (n x need column-major). Returns false if QR did not converge.
inline bool sym_eig_lowest(int n, const double* A, int lda, int need, double* w, double* Z) {
  if (n <= 0 || need <= 0) return true;
  if (need > n) need = n;
  std::vector<double> a((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) a[(size_t)j * n + i] = 0.5 * (A[(size_t)j * lda + i] + A[(size_t)i * lda + j]);
  auto at = [&](int i, int j) -> double& { return a[(size_t)j * n + i]; };
  std::vector<double> refl((size_t)n * n, 0.0), betas(n, 0.0);  // reflector k in column k
  std::vector<double> p(n), wv(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += at(i, k) * at(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (at(k + 1, k) > 0) alpha = -alpha;
    double* v = &refl[(size_t)k * n];
    for (int i = k + 1; i < n; ++i) v[i] = at(i, k);
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    betas[k] = beta;
    // p = beta * A v from the lower triangle only (A symmetric; v[k] = 0)
    for (int i = k; i < n; ++i) p[i] = 0.0;
    for (int j = k; j < n; ++j) {
      const double* col = &a[(size_t)j * n];
      const double vj = v[j];
      double d0 = 0.0, d1 = 0.0, d2 = 0.0, d3 = 0.0;  // four dot-product chains
      p[j] += col[j] * vj;
      int i = j + 1;
      for (; i + 3 < n; i += 4) {
        const double c0 = col[i], c1 = col[i + 1], c2 = col[i + 2], c3 = col[i + 3];
        p[i] += c0 * vj;
        p[i + 1] += c1 * vj;
        p[i + 2] += c2 * vj;
        p[i + 3] += c3 * vj;
        d0 += c0 * v[i];
        d1 += c1 * v[i + 1];
        d2 += c2 * v[i + 2];
        d3 += c3 * v[i + 3];
      }
      for (; i < n; ++i) {
        p[i] += col[i] * vj;
        d0 += col[i] * v[i];
      }
      p[j] += (d0 + d1) + (d2 + d3);
    }
    for (int i = k; i < n; ++i) p[i] *= beta;
    double pv = 0.0;
    for (int i = k + 1; i < n; ++i) pv += p[i] * v[i];
    const double c = 0.5 * beta * pv;
    // v[i] = 0 for i <= k (reflector storage starts zeroed)
    for (int i = k; i < n; ++i) wv[i] = p[i] - c * v[i];
    for (int j = k; j < n; ++j) {  // lower triangle only
      const double vj = v[j], wj = wv[j];
      double* col = &a[(size_t)j * n];
      for (int i = j; i < n; ++i) col[i] -= v[i] * wj + wv[i] * vj;
    }
  }
  std::vector<double> d(n), e(n, 0.0);
  for (int i = 0; i < n; ++i) d[i] = at(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = at(i + 1, i);
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i) tnorm = std::max(tnorm, std::fabs(d[i]) + std::fabs(e[i]) + (i ? std::fabs(e[i - 1]) : 0.0));
  if (tnorm == 0.0) tnorm = 1.0;

  // eigenvalues of the tridiagonal matrix (implicit QR, Wilkinson shift, no vectors)
  std::vector<double> dd(d), ee(e);
  const double eps = 2.220446049250313e-16;
  {
    int h = n - 1, iter = 0;
    const int maxiter = 60 * n + 100;
    while (h > 0) {
      if (std::fabs(ee[h - 1]) <= eps * (std::fabs(dd[h - 1]) + std::fabs(dd[h]))) {
        ee[h - 1] = 0.0;
        --h;
        continue;
      }
      int l = h - 1;
      while (l > 0 && std::fabs(ee[l - 1]) > eps * (std::fabs(dd[l - 1]) + std::fabs(dd[l]))) --l;
      if (++iter > maxiter) return false;
      double t = 0.5 * (dd[h - 1] - dd[h]);
      double e2 = ee[h - 1] * ee[h - 1];
      double den = t + (t >= 0 ? 1.0 : -1.0) * std::sqrt(t * t + e2);
      double mu = dd[h] - (den != 0.0 ? e2 / den : 0.0);
      double x = dd[l] - mu, z = ee[l];
      for (int k = l; k < h; ++k) {
        double r = std::sqrt(x * x + z * z);  // entries are O(|T|): no overflow to guard against
        double cc = 1.0, ss = 0.0;
        if (r != 0.0) { cc = x / r; ss = z / r; }
        if (k > l) ee[k - 1] = r;
        double a0 = dd[k], b0 = ee[k], c0 = dd[k + 1];
        dd[k] = cc * cc * a0 + 2.0 * cc * ss * b0 + ss * ss * c0;
        dd[k + 1] = ss * ss * a0 - 2.0 * cc * ss * b0 + cc * cc * c0;
        ee[k] = cc * ss * (c0 - a0) + (cc * cc - ss * ss) * b0;
        if (k + 1 < h) {
          z = ss * ee[k + 1];
          ee[k + 1] = cc * ee[k + 1];
        }
        x = ee[k];
      }
    }
  }
  std::sort(dd.begin(), dd.end());

  // inverse iteration on T - lambda I (tridiagonal LU with partial pivoting)
  const double eps3 = eps * tnorm;                  // replacement for a zero pivot
  const double ortol = 1e-3 * tnorm;                // eigenvalues closer than this form a cluster
  const double growth = 1.0 / (std::sqrt((double)n) * 1e3 * eps);
  std::vector<double> r1(n), r2(n), r3(n), mult(n), y(n), b(n);
  std::vector<char> swp(n);
  std::vector<double> Y((size_t)n * need);
  unsigned long long seed = 0x9E3779B97F4A7C15ull;
  double lam_prev = 0.0;
  int cluster_start = 0;
  for (int j = 0; j < need; ++j) {
    double lam = dd[j];
    if (j > 0) {
      if (lam - dd[j - 1] > ortol) cluster_start = j;
      if (lam - lam_prev < 10.0 * eps3) lam = lam_prev + 10.0 * eps3;  // separate equal shifts
    }
    lam_prev = lam;
    // factor T - lam I = P L U (U with two superdiagonals)
    double u = d[0] - lam, v = n > 1 ? e[0] : 0.0;
    for (int i = 1; i < n; ++i) {
      const double sub = e[i - 1], di = d[i] - lam, ei = i + 1 < n ? e[i] : 0.0;
      if (std::fabs(sub) > std::fabs(u)) {  // strict: sub != 0 here
        const double xu = u / sub;
        mult[i] = xu;
        swp[i] = 1;
        r1[i - 1] = sub;
        r2[i - 1] = di;
        r3[i - 1] = ei;
        u = v - xu * di;
        v = -xu * ei;
      } else {
        const double xu = sub / (u != 0.0 ? u : eps3);
        mult[i] = xu;
        swp[i] = 0;
        r1[i - 1] = u != 0.0 ? u : eps3;
        r2[i - 1] = v;
        r3[i - 1] = 0.0;
        u = di - xu * v;
        v = ei;
      }
    }
    r1[n - 1] = u != 0.0 ? u : eps3;
    r2[n - 1] = 0.0;
    r3[n - 1] = 0.0;
    for (int i = 0; i + 1 < n; ++i)
      if (r1[i] == 0.0) r1[i] = eps3;
    // start vector: pseudo-random, deterministic
    for (int i = 0; i < n; ++i) {
      seed = seed * 6364136223846793005ull + 1442695040888963407ull;
      b[i] = ((double)(seed >> 11) * (1.0 / 9007199254740992.0)) - 0.5;
    }
    double* yj = &Y[(size_t)j * n];
    for (int it = 0; it < 6; ++it) {
      if (it > 0) {  // forward elimination of the right-hand side with the recorded pivots
        for (int i = 1; i < n; ++i) {
          if (swp[i]) std::swap(b[i - 1], b[i]);
          b[i] -= mult[i] * b[i - 1];
        }
      }
      for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        if (i + 1 < n) s -= r2[i] * y[i + 1];
        if (i + 2 < n) s -= r3[i] * y[i + 2];
        y[i] = s / r1[i];
      }
      // orthogonalise against the earlier vectors of this cluster (twice)
      for (int rep = 0; rep < 2; ++rep)
        for (int c = cluster_start; c < j; ++c) {
          const double* yc = &Y[(size_t)c * n];
          double s = 0.0;
          for (int i = 0; i < n; ++i) s += yc[i] * y[i];
          for (int i = 0; i < n; ++i) y[i] -= s * yc[i];
        }
      double nrm = 0.0;
      for (int i = 0; i < n; ++i) nrm += y[i] * y[i];
      nrm = std::sqrt(nrm);
      if (nrm == 0.0) {  // restart from another vector
        for (int i = 0; i < n; ++i) b[i] = (i % (j + 2) == 0) ? 1.0 : 0.0;
        continue;
      }
      for (int i = 0; i < n; ++i) b[i] = y[i] / nrm;
      if (nrm >= growth && it >= 1) break;
    }
    for (int i = 0; i < n; ++i) yj[i] = b[i];
    w[j] = dd[j];
  }
  // back-transform: Z = H_0 H_1 ... H_{n-3} Y
  // four vectors at a time: each reflector is read once for four independent dot-product chains
  for (int j0 = 0; j0 < need; j0 += 4) {
    const int nj = std::min(4, need - j0);
    double* z[4];
    for (int t = 0; t < 4; ++t) z[t] = &Y[(size_t)(j0 + std::min(t, nj - 1)) * n];  // spare lanes alias the last
    for (int k = n - 3; k >= 0; --k) {
      if (betas[k] == 0.0) continue;
      const double* v = &refl[(size_t)k * n];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int i = k + 1; i < n; ++i) {
        const double vi = v[i];
        s0 += vi * z[0][i];
        s1 += vi * z[1][i];
        s2 += vi * z[2][i];
        s3 += vi * z[3][i];
      }
      const double ss[4] = {s0 * betas[k], s1 * betas[k], s2 * betas[k], s3 * betas[k]};
      if (nj == 4) {
        for (int i = k + 1; i < n; ++i) {
          const double vi = v[i];
          z[0][i] -= ss[0] * vi;
          z[1][i] -= ss[1] * vi;
          z[2][i] -= ss[2] * vi;
          z[3][i] -= ss[3] * vi;
        }
      } else {
        for (int t = 0; t < nj; ++t)
          for (int i = k + 1; i < n; ++i) z[t][i] -= ss[t] * v[i];
      }
    }
    for (int t = 0; t < nj; ++t)
      for (int i = 0; i < n; ++i) Z[(size_t)(j0 + t) * n + i] = z[t][i];
  }
  return true;
}

}  // namespace xg

namespace xg {

// Real nonsymmetric eigenvalues: Householder reduction to upper Hessenberg form, then the
// Francis implicit double-shift QR iteration (Golub & Van Loan, Alg. 7.4.2 and 7.5.1-7.5.2).
// A: n x n column-major (copied). wr/wi receive real and imaginary parts. Returns false if
// the iteration did not converge.
inline bool nonsym_eigvals(int n, const double* A, int lda, double* wr, double* wi) {
  std::vector<double> h((size_t)n * n);
  auto H = [&](int i, int j) -> double& { return h[(size_t)j * n + i]; };
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) H(i, j) = A[(size_t)j * lda + i];
  std::vector<double> v(n);
  for (int k = 0; k + 2 < n; ++k) {
    double alpha = 0.0;
    for (int i = k + 1; i < n; ++i) alpha += H(i, k) * H(i, k);
    alpha = std::sqrt(alpha);
    if (alpha == 0.0) continue;
    if (H(k + 1, k) > 0) alpha = -alpha;
    for (int i = 0; i < n; ++i) v[i] = 0.0;
    for (int i = k + 1; i < n; ++i) v[i] = H(i, k);
    v[k + 1] -= alpha;
    double vn = 0.0;
    for (int i = k + 1; i < n; ++i) vn += v[i] * v[i];
    if (vn == 0.0) continue;
    const double beta = 2.0 / vn;
    for (int j = k; j < n; ++j) {  // H <- (I - b v v^T) H
      double s = 0.0;
      for (int i = k + 1; i < n; ++i) s += v[i] * H(i, j);
      s *= beta;
      for (int i = k + 1; i < n; ++i) H(i, j) -= s * v[i];
    }
    for (int i = 0; i < n; ++i) {  // H <- H (I - b v v^T)
      double s = 0.0;
      for (int j = k + 1; j < n; ++j) s += H(i, j) * v[j];
      s *= beta;
      for (int j = k + 1; j < n; ++j) H(i, j) -= s * v[j];
    }
    for (int i = k + 2; i < n; ++i) H(i, k) = 0.0;
  }
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iter = 0;
  while (hi >= 0) {
    if (hi == 0) { wr[0] = H(0, 0); wi[0] = 0.0; --hi; continue; }
    // find the start of the active unreduced block
    int l = hi;
    while (l > 0) {
      double s = std::fabs(H(l - 1, l - 1)) + std::fabs(H(l, l));
      if (s == 0.0) s = 1.0;
      if (std::fabs(H(l, l - 1)) < eps * s) { H(l, l - 1) = 0.0; break; }
      --l;
    }
    if (l == hi) { wr[hi] = H(hi, hi); wi[hi] = 0.0; --hi; iter = 0; continue; }
    if (l == hi - 1) {  // 2x2 block
      double a = H(hi - 1, hi - 1), b = H(hi - 1, hi), c = H(hi, hi - 1), d = H(hi, hi);
      double tr = 0.5 * (a + d), det = a * d - b * c, disc = tr * tr - det;
      if (disc >= 0.0) {
        double r = std::sqrt(disc);
        double e1 = tr + (tr >= 0 ? r : -r);
        double e2 = e1 != 0.0 ? det / e1 : tr - r;
        wr[hi - 1] = e1; wi[hi - 1] = 0.0;
        wr[hi] = e2; wi[hi] = 0.0;
      } else {
        double r = std::sqrt(-disc);
        wr[hi - 1] = tr; wi[hi - 1] = r;
        wr[hi] = tr; wi[hi] = -r;
      }
      hi -= 2;
      iter = 0;
      continue;
    }
    if (++iter > 60 * n) return false;
    // Francis double shift on rows/cols l..hi
    double s, t;
    if (iter % 11 == 10) {
      // exceptional shift (EISPACK hqr): the roots of x^2 - 1.5 w x + w^2 about the trailing
      // diagonal entry, w the size of the last two subdiagonal entries; centred at zero instead,
      // it can miss a spectrum far from zero (a 283 x 283 harmonic-Ritz matrix with eigenvalues in
      // [0.27, 5.3] never deflated)
      const double w = std::fabs(H(hi, hi - 1)) + std::fabs(H(hi - 1, hi - 2));
      const double c = H(hi, hi);
      s = 2.0 * c + 1.5 * w;
      t = c * c + 1.5 * w * c + w * w;
    } else {
      s = H(hi - 1, hi - 1) + H(hi, hi);
      t = H(hi - 1, hi - 1) * H(hi, hi) - H(hi - 1, hi) * H(hi, hi - 1);
    }
    double x = H(l, l) * H(l, l) + H(l, l + 1) * H(l + 1, l) - s * H(l, l) + t;
    double y = H(l + 1, l) * (H(l, l) + H(l + 1, l + 1) - s);
    double z = (l + 2 <= hi) ? H(l + 2, l + 1) * H(l + 1, l) : 0.0;
    for (int k = l - 1; k <= hi - 3; ++k) {
      // reflector P for (x, y, z) acting on rows k+1..k+3
      double u[3] = {x, y, z};
      double al = std::sqrt(x * x + y * y + z * z);
      if (al != 0.0) {
        if (x > 0) al = -al;
        u[0] -= al;
        double un = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        if (un != 0.0) {
          double be = 2.0 / un;
          int q = std::max(k, l);
          for (int j = q; j < n; ++j) {
            double d0 = be * (u[0] * H(k + 1, j) + u[1] * H(k + 2, j) + u[2] * H(k + 3, j));
            H(k + 1, j) -= d0 * u[0]; H(k + 2, j) -= d0 * u[1]; H(k + 3, j) -= d0 * u[2];
          }
          int r = std::min(k + 4, hi);
          for (int i = 0; i <= r; ++i) {
            double d0 = be * (u[0] * H(i, k + 1) + u[1] * H(i, k + 2) + u[2] * H(i, k + 3));
            H(i, k + 1) -= d0 * u[0]; H(i, k + 2) -= d0 * u[1]; H(i, k + 3) -= d0 * u[2];
          }
        }
      }
      x = H(k + 2, k + 1);
      y = H(k + 3, k + 1);
      if (k < hi - 3) z = H(k + 4, k + 1);
    }
    {  // final 2-element reflector on rows hi-1, hi
      double u[2] = {x, y};
      double al = std::sqrt(x * x + y * y);
      if (al != 0.0) {
        if (x > 0) al = -al;
        u[0] -= al;
        double un = u[0] * u[0] + u[1] * u[1];
        if (un != 0.0) {
          double be = 2.0 / un;
          for (int j = hi - 2 >= 0 ? hi - 2 : 0; j < n; ++j) {
            double d0 = be * (u[0] * H(hi - 1, j) + u[1] * H(hi, j));
            H(hi - 1, j) -= d0 * u[0]; H(hi, j) -= d0 * u[1];
          }
          for (int i = 0; i <= hi; ++i) {
            double d0 = be * (u[0] * H(i, hi - 1) + u[1] * H(i, hi));
            H(i, hi - 1) -= d0 * u[0]; H(i, hi) -= d0 * u[1];
          }
        }
      }
    }
  }
  return true;
}

// Right eigenvector of A for a real eigenvalue lam by inverse iteration on (A - lam' I),
// lam' perturbed by a relative 1e-10 so the LU factorization (partial pivoting) stays finite.
inline void real_eigvec(int n, const double* A, int lda, double lam, double* x) {
  std::vector<double> M((size_t)n * n);
  const double shift = lam + 1e-10 * std::max(1.0, std::fabs(lam));
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) M[(size_t)j * n + i] = A[(size_t)j * lda + i] - (i == j ? shift : 0.0);
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(M[(size_t)k * n + i]) > std::fabs(M[(size_t)k * n + p])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[(size_t)j * n + k], M[(size_t)j * n + p]);
    double d = M[(size_t)k * n + k];
    if (d == 0.0) d = M[(size_t)k * n + k] = 1e-300;
    for (int i = k + 1; i < n; ++i) M[(size_t)k * n + i] /= d;
    for (int j = k + 1; j < n; ++j) {
      double f = M[(size_t)j * n + k];
      if (f != 0.0)
        for (int i = k + 1; i < n; ++i) M[(size_t)j * n + i] -= M[(size_t)k * n + i] * f;
    }
  }
  for (int i = 0; i < n; ++i) x[i] = 1.0 / std::sqrt((double)n) * (1.0 + 0.01 * (double)(i % 7));
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k < n; ++k) std::swap(x[k], x[piv[k]]);
    for (int j = 0; j < n; ++j)
      for (int i = j + 1; i < n; ++i) x[i] -= M[(size_t)j * n + i] * x[j];
    for (int j = n - 1; j >= 0; --j) {
      x[j] /= M[(size_t)j * n + j];
      for (int i = 0; i < j; ++i) x[i] -= M[(size_t)j * n + i] * x[j];
    }
    double nr = 0.0;
    for (int i = 0; i < n; ++i) nr += x[i] * x[i];
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) x[i] /= nr;
  }
}

// Real part of the right eigenvector of A for a complex eigenvalue wr + i wi (wi != 0), by
// complex inverse iteration on (A - mu I), mu perturbed by a relative 1e-10. The complex vector is
// taken with unit Euclidean norm and its largest-modulus component real (LAPACK xGEEV's
// convention); x receives its real part, normalized to unit length.
inline void complex_eigvec_real(int n, const double* A, int lda, double wr, double wi, double* x) {
  using cd = std::complex<double>;
  std::vector<cd> M((size_t)n * n);
  const double scale = std::max(1.0, std::hypot(wr, wi));
  const cd shift(wr + 1e-10 * scale, wi);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) M[(size_t)j * n + i] = cd(A[(size_t)j * lda + i]) - (i == j ? shift : cd(0.0));
  std::vector<int> piv(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    for (int i = k + 1; i < n; ++i)
      if (std::abs(M[(size_t)k * n + i]) > std::abs(M[(size_t)k * n + p])) p = i;
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(M[(size_t)j * n + k], M[(size_t)j * n + p]);
    cd d = M[(size_t)k * n + k];
    if (d == cd(0.0)) d = M[(size_t)k * n + k] = cd(1e-300);
    for (int i = k + 1; i < n; ++i) M[(size_t)k * n + i] /= d;
    for (int j = k + 1; j < n; ++j) {
      const cd f = M[(size_t)j * n + k];
      if (f != cd(0.0))
        for (int i = k + 1; i < n; ++i) M[(size_t)j * n + i] -= M[(size_t)k * n + i] * f;
    }
  }
  std::vector<cd> z(n);
  for (int i = 0; i < n; ++i) z[i] = cd(1.0 / std::sqrt((double)n) * (1.0 + 0.01 * (double)(i % 7)), 0.0);
  for (int it = 0; it < 3; ++it) {
    for (int k = 0; k < n; ++k) std::swap(z[k], z[piv[k]]);
    for (int j = 0; j < n; ++j)
      for (int i = j + 1; i < n; ++i) z[i] -= M[(size_t)j * n + i] * z[j];
    for (int j = n - 1; j >= 0; --j) {
      z[j] /= M[(size_t)j * n + j];
      for (int i = 0; i < j; ++i) z[i] -= M[(size_t)j * n + i] * z[j];
    }
    double nr = 0.0;
    for (int i = 0; i < n; ++i) nr += std::norm(z[i]);
    nr = std::sqrt(nr);
    for (int i = 0; i < n; ++i) z[i] /= nr;
  }
  int kmax = 0;
  for (int i = 1; i < n; ++i)
    if (std::abs(z[i]) > std::abs(z[kmax])) kmax = i;
  const cd phase = std::conj(z[kmax]) / std::abs(z[kmax]);
  double nr = 0.0;
  for (int i = 0; i < n; ++i) {
    x[i] = (z[i] * phase).real();
    nr += x[i] * x[i];
  }
  nr = std::sqrt(nr);
  if (nr > 0.0)
    for (int i = 0; i < n; ++i) x[i] /= nr;
}

}  // namespace xg
