"""MatrixMarket dense-array reader/writer for the golden fixtures.

TEST INFRASTRUCTURE ONLY: part of the oracle; the product path never imports it.

Follows the reference's reader/writer, tools/src/libtools/eigenio_matrixmarket.cc:63-110
(header line, one "rows cols" line, then rows*cols values in column-major order).
"""
from __future__ import annotations

import numpy as np


def read_matrix(path: str) -> np.ndarray:
    """Read a '%%MatrixMarket matrix array real general' file (column-major)."""
    with open(path, "r") as fh:
        lines = fh.read().splitlines()
    i = 0
    while lines[i].startswith("%"):
        i += 1
    rows, cols = (int(t) for t in lines[i].split()[:2])
    vals = []
    for ln in lines[i + 1:]:
        ln = ln.strip()
        if ln:
            vals.append(float(ln.split()[0]))
        if len(vals) == rows * cols:
            break
    if len(vals) != rows * cols:
        raise ValueError(f"{path}: expected {rows*cols} values, got {len(vals)}")
    return np.asarray(vals, dtype=np.float64).reshape((cols, rows)).T.copy()


def read_vector(path: str) -> np.ndarray:
    m = read_matrix(path)
    return m.reshape(-1)


def write_matrix(path: str, m: np.ndarray) -> None:
    m = np.atleast_2d(np.asarray(m, dtype=np.float64))
    with open(path, "w") as fh:
        fh.write("%%MatrixMarket matrix array real general\n")
        fh.write(f"{m.shape[0]} {m.shape[1]}\n")
        for v in m.T.reshape(-1):
            fh.write(f"{v:.17e}\n")


def isapprox(a: np.ndarray, b: np.ndarray, prec: float) -> bool:
    """Eigen's isApprox: ||a-b||_F <= prec * min(||a||_F, ||b||_F)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b)) <= prec * min(float(np.linalg.norm(a)),
                                                      float(np.linalg.norm(b)))
