"""The device Davidson on generic matrix-free operators (MatrixFreeOperator, matrixfreeoperator.h:45-98):
the reference's own operator tests (test_davidson.cc:127-230 TestOperator, :262-340
HermitianBlockOperator under HamiltonianOperator) through Python and C++ callbacks, the library's
BSE operator through the same interface, and mixed-precision iterations on a callback operator."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from votca_amd import pyxtp as px

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def is_approx(a, b, prec):
    """Eigen's isApprox: ||a - b|| <= prec * min(||a||, ||b||)."""
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return np.linalg.norm(a - b) <= prec * min(np.linalg.norm(a), np.linalg.norm(b))


def init_matrix(n, eps):  # test_davidson.cc:45-59
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    with np.errstate(divide="ignore"):
        A = eps / (j - i).astype(float) ** 2
    A[np.diag_indices(n)] = np.sqrt(1.0 + np.arange(n))
    return A


def symm_matrix(n, eps, rng):  # test_davidson.cc:37-43 (uniform [-1, 1] entries, symmetrised)
    m = eps * rng.uniform(-1.0, 1.0, size=(n, n))
    return m + m.T


class TestOperator(px.MatrixFreeOperator):
    """test_davidson.cc:127-165: rows built on the fly, diagonal sqrt(i + 1), off-diagonal 0.01/(j-i)^2."""

    def row(self, i):
        n = self.size()
        j = np.arange(n)
        with np.errstate(divide="ignore"):
            r = 0.01 / (j - i).astype(float) ** 2
        r[i] = np.sqrt(i + 1.0)
        return r

    def matmul(self, X):
        return np.vstack([self.row(i) @ X for i in range(self.size())])

    def diagonal(self):
        return np.array([self.row(i)[i] for i in range(self.size())])


class BlockOperator(px.MatrixFreeOperator):
    """HermitianBlockOperator (test_davidson.cc:262-280): a dense matrix behind the interface; the
    optional mixed form applies the matrix rounded to FP32."""

    def __init__(self, mat, precisions=("fp64",)):
        self.mat = np.asarray(mat, dtype=np.float64)
        self.precisions = precisions
        self.calls = {"fp64": 0, "bf16x3": 0}
        self.set_size(self.mat.shape[0])

    def matmul(self, X, precision="fp64"):
        self.calls[precision] = self.calls.get(precision, 0) + 1
        if precision == "fp64":
            return self.mat @ X
        return (self.mat.astype(np.float32) @ X.astype(np.float32)).astype(np.float64)

    def diagonal(self):
        return np.diag(self.mat).copy()


@pytest.mark.parametrize("size", [100, 400])
def test_matrix_free_test_operator(size):  # davidson_matrix_free(_large), test_davidson.cc:167-230
    op = TestOperator()
    op.set_size(size)
    ds = px.DavidsonSolver()
    ds.set_tolerance("normal")
    ds.set_size_update("safe")
    ev, vec = ds.solve(op, 10)
    assert ds.converged()
    A = op @ np.eye(size)
    ref = np.linalg.eigvalsh(A)[:10]
    assert is_approx(ev, ref, 1e-6), (ev, ref)
    assert np.allclose(A, init_matrix(size, 0.01), atol=0, rtol=1e-15)
    # eigenvectors: unit norm, A v = lambda v to the Davidson tolerance
    assert np.abs(np.linalg.norm(vec, axis=0) - 1).max() < 1e-12
    assert np.linalg.norm(A @ vec - vec * ev, axis=0).max() < 1e-4


def test_matrix_free_iter_max_one_does_not_converge():  # davidson_full_matrix_fail, test_davidson.cc:107-125
    op = TestOperator()
    op.set_size(100)
    ds = px.DavidsonSolver()
    ds.set_iter_max(1)
    ev, _ = ds.solve(op, 10)
    assert not ds.converged()
    assert not is_approx(ev, np.linalg.eigvalsh(op @ np.eye(100))[:10], 1e-6)


@pytest.mark.parametrize("size,neigen", [(60, 5), (120, 5)])
def test_matrix_free_hamiltonian(size, neigen):  # davidson_hamiltonian_matrix_free(_large), test_davidson.cc:282-400
    rng = np.random.default_rng(size)
    R = BlockOperator(init_matrix(size, 0.01))
    Cm = BlockOperator(symm_matrix(size, 0.01, rng))
    H = px.HamiltonianOperator(R, Cm)
    ds = px.DavidsonSolver()
    ds.set_tolerance("normal")
    ds.set_size_update("max")
    ds.set_matrix_type("HAM")
    ev, vec = ds.solve(H, neigen)
    assert ds.converged()
    Hd = H @ np.eye(2 * size)
    assert np.allclose(Hd, np.block([[R.mat, Cm.mat], [-Cm.mat, -R.mat]]), atol=1e-15)
    w, U = np.linalg.eig(Hd)
    pos = [i for i in np.argsort(w.real) if w[i].real > 0][:neigen]
    order = np.argsort(ev)
    assert is_approx(ev[order], w.real[pos], 1e-6)
    ref = U[:, pos].real
    ref /= np.linalg.norm(ref, axis=0)
    assert is_approx(np.abs(vec[:, order]) ** 2, np.abs(ref) ** 2, 1e-3)


def test_bse_operator_through_generic_interface(ctx=None):
    """The library's BSE operator solved as a BSE_OPERATOR and as a MatrixFreeOperator view of it
    (xtpgpu_mfop_from_bse): the same solver on the same operator, so the same roots and vectors."""
    ctx = px.default_context()
    o, v, P = 6, 30, 50
    n = o + v
    Mmn = px.TCMatrix_gwbse(ctx).Initialize(P, 0, n - 1, 0, n - 1)
    Mmn.fill_synthetic(3, 0.05, 3.0)
    e = np.concatenate([np.linspace(-0.8, -0.3, o), np.linspace(0.1, 1.5, v)])
    op = px.SingletOperator_TDA(np.full(P, 0.8), Mmn, np.diag(e)).configure(
        px.BSEOperator_Options(o - 1, 0, 0, 0, n - 1))

    class View(px.MatrixFreeOperator):  # a Python operator that forwards to the device operator
        def __init__(self, bse):
            self.bse = bse
            self.set_size(bse.size)

        def diagonal(self):
            return self.bse.diagonal()

        def matmul(self, X):
            return self.bse.matmul(X)

    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-9)
    ev1, v1 = ds.solve(op, 5)
    ev2, v2 = ds.solve(View(op), 5)
    assert np.abs(ev1 - ev2).max() < 1e-12
    assert np.abs(np.abs(np.sum(v1 * v2, axis=0)) - 1).max() < 1e-9
    # full BSE: HamiltonianOperator over the TDA operator and its coupling block through the generic
    # HAM path, started from the BSE path's X-ranked solution, against the dense spectrum
    B = px.SingletOperator_BTDA_B(np.full(P, 0.8), Mmn, np.diag(e)).configure(
        px.BSEOperator_Options(o - 1, 0, 0, 0, n - 1))
    dsf = px.DavidsonSolver()
    dsf.set_tolerance(1e-8)
    dsf.set_max_search_space(40)
    dsf.set_iter_max(100)
    evf, Xf, Yf = dsf.solve_full(op, B, 4)
    assert dsf.converged()
    ds2 = px.DavidsonSolver()
    ds2.set_tolerance(1e-8)
    ds2.set_size_update("max")
    ds2.set_matrix_type("HAM")
    evh, vh = ds2.solve(px.HamiltonianOperator(op, B), 4, initial_guess=np.vstack([Xf, Yf]))
    assert ds2.converged()
    Ad, Bd = op @ np.eye(op.size), B @ np.eye(op.size)
    H = np.block([[Ad, Bd], [-Bd, -Ad]])
    w = np.sort(np.linalg.eigvals(H).real)
    assert np.abs(np.sort(evh) - w[w > 0][:4]).max() < 1e-7
    assert np.abs(np.sort(evf) - w[w > 0][:4]).max() < 1e-7
    assert np.linalg.norm(H @ vh - vh * evh, axis=0).max() < 1e-7


def test_mixed_precision_on_callback_operator():
    """The precision ladder and FP64 refinement of the device Davidson driving a caller's operator
    that declares a mixed form (here the matrix rounded to FP32): the refined roots meet the FP64
    tolerance and both forms were called."""
    rng = np.random.default_rng(7)
    n = 300
    A = np.diag(np.linspace(0.5, 3.0, n)) + 0.01 * symm_matrix(n, 1.0, rng)
    op = BlockOperator(A, precisions=("fp64", "bf16x3"))
    ds = px.DavidsonSolver()
    ds.set_tolerance(1e-6)  # the mixed loop stops at what the FP32 matrix resolves
    ds.set_precision("bf16x3", refine_tolerance=1e-9, refine_max=8)
    ev, vec = ds.solve(op, 6)
    assert ds.converged()
    assert op.calls["bf16x3"] > 0 and op.calls["fp64"] > 0
    ref = np.linalg.eigvalsh(A)[:6]
    assert np.abs(ev - ref).max() < 1e-9
    assert np.linalg.norm(A @ vec - vec * ev, axis=0).max() < 1e-8


def test_callback_errors_propagate():
    class Bad(px.MatrixFreeOperator):
        def diagonal(self):
            return np.arange(1.0, 31.0)

        def matmul(self, X):
            raise ValueError("matmul failed")

    op = Bad()
    op.set_size(30)
    with pytest.raises(ValueError, match="matmul failed"):
        px.DavidsonSolver().solve(op, 3)
    # an operator without the asked-for precision is refused before any callback
    ok = BlockOperator(np.diag(np.arange(1.0, 31.0)))
    ds = px.DavidsonSolver()
    ds.set_precision("bf16x3")
    with pytest.raises(RuntimeError):
        ds.solve(ok, 3)


CPP_MF_DRIVER = r"""
#include "xtpgpu.hpp"
#include <cmath>
#include <cstdio>
#include <vector>
// The reference's TestOperator (test_davidson.cc:127-165) and HermitianBlockOperator pair as C++
// subclasses of xtpgpu::HostMatrixFreeOperator; prints 10 SYMM roots then 5 HAM roots.
struct TestOp : xtpgpu::HostMatrixFreeOperator {
  using HostMatrixFreeOperator::HostMatrixFreeOperator;
  double el(int64_t i, int64_t j) const {
    return i == j ? std::sqrt(double(i + 1)) : 0.01 / double((j - i) * (j - i));
  }
  std::vector<double> diagonal() const override {
    std::vector<double> d(size_);
    for (int64_t i = 0; i < size_; ++i) d[i] = el(i, i);
    return d;
  }
  void matmul(const double* X, int64_t k, double* Y) const override {
    for (int64_t c = 0; c < k; ++c)
      for (int64_t i = 0; i < size_; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < size_; ++j) s += el(i, j) * X[c * size_ + j];
        Y[c * size_ + i] = s;
      }
  }
};
struct BlockOp : xtpgpu::HostMatrixFreeOperator {
  std::vector<double> m;  // column-major
  BlockOp(xtpgpu::Context& c, std::vector<double> a, int64_t n) : HostMatrixFreeOperator(c), m(std::move(a)) {
    set_size(n);
  }
  std::vector<double> diagonal() const override {
    std::vector<double> d(size_);
    for (int64_t i = 0; i < size_; ++i) d[i] = m[i * size_ + i];
    return d;
  }
  void matmul(const double* X, int64_t k, double* Y) const override {
    for (int64_t c = 0; c < k; ++c)
      for (int64_t i = 0; i < size_; ++i) {
        double s = 0.0;
        for (int64_t j = 0; j < size_; ++j) s += m[j * size_ + i] * X[c * size_ + j];
        Y[c * size_ + i] = s;
      }
  }
};
int main(int argc, char** argv) {
  xtpgpu::Context c(0);
  TestOp op(c);
  op.set_size(100);
  xtpgpu::DavidsonSolver ds;
  ds.set_tolerance("normal");
  ds.set_size_update("safe");
  if (!ds.solve(op, 10)) return 2;
  for (double e : ds.eigenvalues()) std::printf("%.17g\n", e);
  const int64_t n = 60;
  std::vector<double> r(n * n), cm(n * n);
  std::FILE* f = std::fopen(argv[1], "rb");
  if (!f || std::fread(r.data(), 8, n * n, f) != (size_t)(n * n) || std::fread(cm.data(), 8, n * n, f) != (size_t)(n * n))
    return 4;
  std::fclose(f);
  BlockOp R(c, r, n), C(c, cm, n);
  xtpgpu::DavidsonSolver dh;
  dh.set_tolerance("normal");
  dh.set_size_update("max");
  if (!dh.solve_ham(R, C, 5)) return 3;
  for (double e : dh.eigenvalues()) std::printf("%.17g\n", e);
  return 0;
}
"""


def test_cpp_matrix_free_operator_gpu(tmp_path):
    """include/xtpgpu.hpp: a C++ HostMatrixFreeOperator subclass (the reference's TestOperator) and a
    HamiltonianOperator pair solved by DavidsonSolver::solve / solve_ham."""
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    n = 60
    rng = np.random.default_rng(11)
    R, Cm = init_matrix(n, 0.01), symm_matrix(n, 0.01, rng)
    f = tmp_path / "m.bin"
    f.write_bytes(np.asfortranarray(R).tobytes(order="F") + np.asfortranarray(Cm).tobytes(order="F"))
    src = tmp_path / "mf.cc"
    src.write_text(CPP_MF_DRIVER)
    exe = tmp_path / "mf"
    libdir = os.path.dirname(px.LIB_PATH)
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                        "-L", libdir, "-lxtpgpu", "-Wl,-rpath," + libdir], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("[cpp-mfop] running the C++ driver", flush=True)
    r = subprocess.run([str(exe), str(f)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stderr)
    vals = np.array([float(x) for x in r.stdout.split()])
    assert is_approx(vals[:10], np.linalg.eigvalsh(init_matrix(100, 0.01))[:10], 1e-6)
    w = np.linalg.eigvals(np.block([[R, Cm], [-Cm, -R]])).real
    assert is_approx(np.sort(vals[10:]), np.sort(w[w > 0])[:5], 1e-6)
