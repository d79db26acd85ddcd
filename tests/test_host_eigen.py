"""Host FP64 eigensolvers used by the Davidson loops (votca_amd/csrc/host_eigen.h), compiled with
g++ and checked against numpy's LAPACK on the same matrices: the full symmetric solver, the
partial (lowest `need`) solver on Davidson-like, exactly degenerate, tightly clustered and random
matrices, and the nonsymmetric eigenvalue routine."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "host_eigen.h"
#include <cstdio>
#include <vector>
int main() {
  int n, need, kind;
  while (std::scanf("%d %d %d", &kind, &n, &need) == 3) {
    std::vector<double> A((size_t)n * n);
    for (auto& x : A) std::scanf("%lf", &x);
    if (kind == 0) {
      std::vector<double> w(n), Z((size_t)n * n);
      int ok = xg::sym_eig(n, A.data(), n, w.data(), Z.data());
      std::printf("%d\n", ok);
      for (int i = 0; i < n; ++i) std::printf("%.17g ", w[i]);
      std::printf("\n");
      for (size_t i = 0; i < Z.size(); ++i) std::printf("%.17g ", Z[i]);
    } else if (kind == 1) {
      std::vector<double> w(need), Z((size_t)n * need);
      int ok = xg::sym_eig_lowest(n, A.data(), n, need, w.data(), Z.data());
      std::printf("%d\n", ok);
      for (int i = 0; i < need; ++i) std::printf("%.17g ", w[i]);
      std::printf("\n");
      for (size_t i = 0; i < Z.size(); ++i) std::printf("%.17g ", Z[i]);
    } else if (kind == 2) {
      std::vector<double> wr(n), wi(n);
      int ok = xg::nonsym_eigvals(n, A.data(), n, wr.data(), wi.data());
      std::printf("%d\n", ok);
      for (int i = 0; i < n; ++i) std::printf("%.17g ", wr[i]);
      std::printf("\n");
      for (int i = 0; i < n; ++i) std::printf("%.17g ", wi[i]);
    } else {
      // eigenvalue `need` given as (wr, wi) after the matrix: real part of its eigenvector
      double wr, wi;
      std::scanf("%lf %lf", &wr, &wi);
      std::vector<double> x(n);
      if (wi == 0.0)
        xg::real_eigvec(n, A.data(), n, wr, x.data());
      else
        xg::complex_eigvec_real(n, A.data(), n, wr, wi, x.data());
      std::printf("1\n");
      for (int i = 0; i < n; ++i) std::printf("%.17g ", x[i]);
      std::printf("\n");
    }
    std::printf("\n");
  }
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("eig")
    src = d / "drv.cc"
    src.write_text(DRIVER)
    exe = d / "drv"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "votca_amd", "csrc"), str(src), "-o",
                    str(exe)], check=True)
    return str(exe)


def run(driver, kind, A, need=0, extra=""):
    n = A.shape[0]
    inp = f"{kind} {n} {need}\n" + " ".join(repr(float(x)) for x in A.flatten(order="F")) + "\n" + extra
    out = subprocess.run([driver], input=inp, capture_output=True, text=True, check=True).stdout.split("\n")
    ok = int(out[0])
    a = np.array(out[1].split(), dtype=float)
    b = np.array(out[2].split(), dtype=float)
    return ok, a, b


def cases():
    rng = np.random.default_rng(7)
    out = []
    for n in (1, 2, 5, 40, 130):
        B = rng.normal(scale=1e-3, size=(n, n))
        out.append(("davidson", 0.5 * (B + B.T) + np.diag(0.2 + 1e-3 * np.arange(n))))
        out.append(("degenerate", np.diag(0.1 * (np.arange(n) // 3))))
        C = rng.normal(scale=1e-9, size=(n, n))
        out.append(("cluster", 0.5 * (C + C.T) + np.diag(0.5 + 1e-12 * np.arange(n))))
        R = rng.normal(size=(n, n))
        out.append(("random", 0.5 * (R + R.T)))
    return out


@pytest.mark.parametrize("name,A", cases(), ids=lambda x: x if isinstance(x, str) else str(x.shape[0]))
def test_symmetric_solvers_match_lapack(driver, name, A):
    n = A.shape[0]
    ref = np.linalg.eigvalsh(A)
    scale = max(np.abs(A).max(), 1e-300)
    ok, w, Z = run(driver, 0, A)
    assert ok
    Z = Z.reshape(n, n, order="F")
    assert np.abs(w - ref).max() <= 1e-13 * scale * max(1, n)
    assert np.abs(A @ Z - Z * w).max() <= 1e-13 * scale * max(1, n)
    assert np.abs(Z.T @ Z - np.eye(n)).max() <= 1e-13 * max(1, n)
    need = min(n, 40)
    ok, wl, Zl = run(driver, 1, A, need)
    assert ok
    Zl = Zl.reshape(n, need, order="F")
    assert np.abs(wl - ref[:need]).max() <= 1e-13 * scale * max(1, n)
    assert np.abs(A @ Zl - Zl * wl).max() <= 1e-13 * scale * max(1, n)
    assert np.abs(Zl.T @ Zl - np.eye(need)).max() <= 1e-13 * max(1, n)


def test_nonsymmetric_eigenvalues_match_lapack(driver):
    rng = np.random.default_rng(8)
    for n in (1, 3, 12, 60):
        A = rng.normal(size=(n, n))
        ok, wr, wi = run(driver, 2, A)
        assert ok
        got = np.sort_complex(wr + 1j * wi)
        ref = np.sort_complex(np.linalg.eigvals(A))
        assert np.abs(got - ref).max() < 1e-10 * max(1.0, np.abs(ref).max())


def test_nonsymmetric_qr_exceptional_shift_regression(driver):
    """A 283 x 283 harmonic-Ritz matrix B^-1 T of the full-BSE Davidson (a perturbed solve of the
    single-segment adjoint checks, captured on the device) with eigenvalues in [0.27, 5.3] and
    close complex pairs: the Francis iteration stalls on it unless its exceptional shift is centred
    on the trailing diagonal entry (EISPACK hqr). Eigenvalues against LAPACK's."""
    A = np.load(os.path.join(ROOT, "tests", "golden", "regress", "harmonic_ritz_283.npy"))
    ok, wr, wi = run(driver, 2, A)
    assert ok
    got = wr + 1j * wi
    ref = np.linalg.eigvals(A)
    assert max(np.abs(got - x).min() for x in ref) < 1e-10
    assert max(np.abs(ref - x).min() for x in got) < 1e-10


def test_eigenvector_real_parts_match_lapack(driver):
    """Right eigenvectors for real eigenvalues and the real parts of those for complex ones (unit
    norm, largest-modulus component real: xGEEV's convention), as the harmonic Ritz step keeps
    them, against numpy's LAPACK eig up to sign. Includes a near-degenerate real triple."""
    rng = np.random.default_rng(9)
    mats = [rng.normal(size=(n, n)) for n in (2, 7, 30)]
    Q = np.linalg.qr(rng.normal(size=(6, 6)))[0]
    D = np.diag([1.0, 1.0 + 1e-6, 1.0 + 2e-6, 3.0, -2.0, 5.0])
    mats.append(Q @ D @ np.linalg.inv(Q + 0.1 * rng.normal(size=(6, 6))) @ (Q + 0.1 * rng.normal(size=(6, 6))) @ Q.T)
    ncomplex = 0
    for A in mats:
        w, V = np.linalg.eig(A)
        for i in range(len(w)):
            if w[i].imag < 0:
                continue
            ncomplex += w[i].imag != 0
            _, x, _ = run(driver, 3, A, 0, f"{float(w[i].real)!r} {float(w[i].imag)!r}\n")
            ref = V[:, i].real / np.linalg.norm(V[:, i].real)
            assert abs(abs(x @ ref) - 1.0) < 1e-8, (A.shape, w[i])
    assert ncomplex > 0
