"""Gaussian quadratures for the imaginary-frequency axis.

Mirrors the reference's QuadratureFactory schemes (xtp/src/libxtp/factories/quadrature_factory.cc:31-36,
xtp/src/libxtp/gaussian_quadrature/gauss_legendre_quadrature.h:29-74,
gauss_laguerre_quadrature.h, gauss_hermite_quadrature.h): the reference tabulates nodes and
"adapted" weights for orders {8,10,12,14,16,18,20,40,100}; here they are computed by the
Golub-Welsch eigenvalue method (numpy.polynomial) for any order, and pinned against the
reference's integral fixtures (gaussian_quadratures/*.mm) in the tests.

Each scheme exposes points(), weights() (already mapped to the integration variable), and
use_symmetry() (the factor-2 convention of GaussianQuadratureBase::Integrate).
"""
from __future__ import annotations

import math

import numpy as np
from numpy.polynomial import hermite, laguerre, legendre

ORDERS = (8, 10, 12, 14, 16, 18, 20, 40, 100)


class _Base:
    def __init__(self, order: int):
        if order < 1:
            raise ValueError("quadrature order must be positive")
        self.order = order
        self._x, self._w = self._nodes(order)

    def points(self) -> np.ndarray:
        return np.array([self.scaled_point(i) for i in range(self.order)])

    def weights(self) -> np.ndarray:
        return np.array([self.scaled_weight(i) for i in range(self.order)])

    def integrate(self, f) -> float:
        """GaussianQuadratureBase::Integrate: sum_j w_j f(j, x_j, symmetry)."""
        return float(sum(self.scaled_weight(j) * f(j, self.scaled_point(j), self.use_symmetry())
                         for j in range(self.order)))


class ModifiedLegendre(_Base):
    """x' = 0.5 (1+x)/(1-x), w' = w/(1-x)^2 on [0, inf); symmetric (gauss_legendre_quadrature.h:37-54)."""

    def _nodes(self, n):
        return legendre.leggauss(n)

    def scaled_point(self, i):
        x = self._x[i]
        return 0.5 * (1.0 + x) / (1.0 - x)

    def scaled_weight(self, i):
        x = self._x[i]
        return self._w[i] / ((1.0 - x) * (1.0 - x))

    def use_symmetry(self):
        return True


class Legendre(_Base):
    """x' = tan(pi x/2), w' = w (pi/2)/cos^2(pi x/2) on (-inf, inf) (gauss_legendre_quadrature.h:56-72)."""

    def _nodes(self, n):
        return legendre.leggauss(n)

    def scaled_point(self, i):
        return math.tan(0.5 * math.pi * self._x[i])

    def scaled_weight(self, i):
        c = math.cos(0.5 * math.pi * self._x[i])
        return self._w[i] * 0.5 * math.pi / (c * c)

    def use_symmetry(self):
        return False


class Laguerre(_Base):
    """Gauss-Laguerre nodes with adapted weights w e^{x} on [0, inf); symmetric."""

    def _nodes(self, n):
        x, w = laguerre.laggauss(n)
        return x, w * np.exp(x)

    def scaled_point(self, i):
        return self._x[i]

    def scaled_weight(self, i):
        return self._w[i]

    def use_symmetry(self):
        return True


class Hermite(_Base):
    """Gauss-Hermite nodes with adapted weights w e^{x^2} on (-inf, inf)."""

    def _nodes(self, n):
        x, w = hermite.hermgauss(n)
        return x, w * np.exp(x * x)

    def scaled_point(self, i):
        return self._x[i]

    def scaled_weight(self, i):
        return self._w[i]

    def use_symmetry(self):
        return False


_SCHEMES = {"modified_legendre": ModifiedLegendre, "legendre": Legendre, "laguerre": Laguerre,
            "hermite": Hermite}


def create(scheme: str, order: int):
    """QuadratureFactory().Create(scheme) + configure(order)."""
    if scheme not in _SCHEMES:
        raise RuntimeError(f"unknown quadrature scheme {scheme}")
    return _SCHEMES[scheme](order)
