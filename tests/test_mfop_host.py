"""Host-side pieces of the matrix-free operator interface (no GPU): the HamiltonianOperator block
product of bseoperator_btda.h:120-148 on Python operators, and the abstract MatrixFreeOperator."""
import numpy as np
import pytest

from votca_amd import pyxtp as px


class Dense(px.MatrixFreeOperator):
    def __init__(self, m):
        self.m = np.asarray(m, dtype=float)
        self.set_size(self.m.shape[0])

    def diagonal(self):
        return np.diag(self.m).copy()

    def matmul(self, X):
        return self.m @ X


def test_hamiltonian_operator_block_product():
    rng = np.random.default_rng(0)
    a, b = rng.normal(size=(7, 7)), rng.normal(size=(7, 7))
    a, b = a + a.T, b + b.T
    H = px.HamiltonianOperator(Dense(a), Dense(b))
    X = rng.normal(size=(14, 3))
    assert H.rows() == H.cols() == 14
    assert np.allclose(H @ X, np.block([[a, b], [-b, -a]]) @ X, atol=1e-14)
    with pytest.raises(ValueError):
        px.HamiltonianOperator(Dense(a), Dense(np.eye(3)))


def test_matrix_free_operator_interface():
    op = Dense(np.diag([3.0, 1.0, 2.0]))
    assert (op.size(), op.rows(), op.cols()) == (3, 3, 3)
    assert np.array_equal(op @ np.eye(3), np.diag([3.0, 1.0, 2.0]))
    with pytest.raises(NotImplementedError):
        px.MatrixFreeOperator().diagonal()
    ds = px.DavidsonSolver()
    with pytest.raises(RuntimeError):
        ds.set_matrix_type("GEN")
    ds.set_matrix_type("HAM")
    with pytest.raises(RuntimeError, match="HamiltonianOperator"):
        ds.solve(op, 1)
