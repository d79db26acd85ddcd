"""Gaussian integrals (overlap, 2- and 3-centre Coulomb, kinetic, nuclear attraction) for s/p
basis sets.

TEST INFRASTRUCTURE ONLY: part of the oracle. The reference builds its three-centre
tensor with libint2 (xtp/src/libxtp/libint2_calls.cc:540-651), which is not present
here. To feed the reference's CH4 / 3-21G golden fixtures (threecenter_gwbse, rpa,
bse_operator, sigma_ppm, sigma_cda, gw, bse) with the same Mmn, this module evaluates
the integrals with the McMurchie-Davidson scheme (Hermite expansion coefficients E,
Hermite Coulomb integrals R, Boys function), written from the published method
(McMurchie & Davidson, J. Comput. Phys. 26, 218 (1978); Helgaker, Jorgensen, Olsen,
"Molecular Electronic-Structure Theory", ch. 9).

Conventions matched to the reference:
  * basis XML layout and "SP" shells split into an S then a P shell
    (xtp/src/libxtp/basisset.cc:170-190);
  * contracted shells normalised to unit self-overlap (xtp/src/libxtp/aoshell.cc:81-88);
  * p functions ordered (y, z, x) = real solid harmonics m=-1,0,+1
    (xtp/src/libxtp/aoshell.cc:118-131);
  * geometry read in Angstrom and converted with ang2bohr
    (tools/include/votca/tools/constants.h:48-54).
The result is pinned against threecenter_gwbse/ref{0,2,4}b.mm in the tests; the one-electron
integrals by closed forms (one s primitive, the hydrogen variational bound) and, through the
core-Hamiltonian MOs they produce, by the plasmon-pole known answers of test_ppm.cc.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass
from functools import lru_cache

import numpy as np
from scipy.special import hyp1f1

ANG2BOHR = 1.8897259886
HRT2EV = 27.21138602

# Cartesian exponent triples per angular momentum, in the reference's AO order.
_CART = {
    0: [(0, 0, 0)],
    1: [(0, 1, 0), (0, 0, 1), (1, 0, 0)],  # y, z, x
}


@dataclass
class Shell:
    l: int
    center: np.ndarray
    exps: np.ndarray
    coefs: np.ndarray  # normalised contraction coefficients (incl. primitive norms)

    @property
    def nfunc(self) -> int:
        return 2 * self.l + 1


def read_xyz(path: str):
    with open(path) as fh:
        lines = fh.read().splitlines()
    n = int(lines[0].split()[0])
    atoms = []
    for ln in lines[2:2 + n]:
        t = ln.split()
        atoms.append((t[0], np.array([float(x) for x in t[1:4]]) * ANG2BOHR))
    return atoms


def read_basis(path: str):
    root = ET.parse(path).getroot()
    basis = {}
    for el in root.findall("element"):
        shells = []
        for sh in el.findall("shell"):
            stype = sh.get("type")
            for sub in stype:
                exps, coefs = [], []
                for const in sh.findall("constant"):
                    for c in const.findall("contractions"):
                        if c.get("type") == sub:
                            exps.append(float(const.get("decay")))
                            coefs.append(float(c.get("factor")))
                shells.append(({"S": 0, "P": 1, "D": 2}[sub], exps, coefs))
        basis[el.get("name")] = shells
    return basis


def _prim_norm(a: float, l: int) -> float:
    return (2.0 * a / math.pi) ** 0.75 * (4.0 * a) ** (l / 2.0)


def build_shells(xyz_path: str, basis_path: str):
    atoms = read_xyz(xyz_path)
    basis = read_basis(basis_path)
    shells = []
    for sym, pos in atoms:
        for l, exps, coefs in basis[sym]:
            exps = np.array(exps)
            c = np.array(coefs) * np.array([_prim_norm(a, l) for a in exps])
            sh = Shell(l, pos.copy(), exps, c)
            # normalise the contracted function to unit self-overlap
            s = _shell_overlap(sh, sh)[0, 0]
            sh.coefs = c / math.sqrt(s)
            shells.append(sh)
    return shells


# ---------------------------------------------------------------- MD machinery
def _E(i: int, j: int, t: int, qx: float, a: float, b: float) -> float:
    """Hermite expansion coefficient E^{ij}_t (recursive)."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == 0 and j == 0 and t == 0:
        return math.exp(-q * qx * qx)
    if j == 0:
        return (_E(i - 1, j, t - 1, qx, a, b) / (2 * p)
                - (q * qx / a) * _E(i - 1, j, t, qx, a, b)
                + (t + 1) * _E(i - 1, j, t + 1, qx, a, b))
    return (_E(i, j - 1, t - 1, qx, a, b) / (2 * p)
            + (q * qx / b) * _E(i, j - 1, t, qx, a, b)
            + (t + 1) * _E(i, j - 1, t + 1, qx, a, b))


def _boys(n: int, T: float) -> float:
    return float(hyp1f1(n + 0.5, n + 1.5, -T)) / (2.0 * n + 1.0)


def _R(t: int, u: int, v: int, n: int, p: float, X: float, Y: float, Z: float,
       cache: dict) -> float:
    key = (t, u, v, n)
    if key in cache:
        return cache[key]
    if t < 0 or u < 0 or v < 0:
        return 0.0
    if t == 0 and u == 0 and v == 0:
        val = (-2.0 * p) ** n * _boys(n, p * (X * X + Y * Y + Z * Z))
    elif t > 0:
        val = ((t - 1) * _R(t - 2, u, v, n + 1, p, X, Y, Z, cache)
               + X * _R(t - 1, u, v, n + 1, p, X, Y, Z, cache))
    elif u > 0:
        val = ((u - 1) * _R(t, u - 2, v, n + 1, p, X, Y, Z, cache)
               + Y * _R(t, u - 1, v, n + 1, p, X, Y, Z, cache))
    else:
        val = ((v - 1) * _R(t, u, v - 2, n + 1, p, X, Y, Z, cache)
               + Z * _R(t, u, v - 1, n + 1, p, X, Y, Z, cache))
    cache[key] = val
    return val


def _herm_pair(la, A, a, lb, B, b):
    """Hermite coefficients for a primitive pair: list over (cart_a, cart_b) of
    dict {(t,u,v): E}. b == 0 means 'unit function' (lb must be 0)."""
    out = {}
    for ia, ca in enumerate(_CART[la]):
        for ib, cb in enumerate(_CART[lb]):
            coeffs = {}
            ex = [[_E(ca[d], cb[d], t, A[d] - B[d], a, b) if b > 0 else _E1(ca[d], t, a)
                   for t in range(ca[d] + cb[d] + 1)] for d in range(3)]
            for t in range(len(ex[0])):
                for u in range(len(ex[1])):
                    for v in range(len(ex[2])):
                        val = ex[0][t] * ex[1][u] * ex[2][v]
                        if val != 0.0:
                            coeffs[(t, u, v)] = val
            out[(ia, ib)] = coeffs
    return out


def _E1(i: int, t: int, a: float) -> float:
    """E coefficients of a single Gaussian x^i exp(-a x^2) about its own centre."""
    if t < 0 or t > i:
        return 0.0
    if i == 0:
        return 1.0
    return _E1(i - 1, t - 1, a) / (2 * a) + (t + 1) * _E1(i - 1, t + 1, a)


def _shell_overlap(s1: Shell, s2: Shell) -> np.ndarray:
    out = np.zeros((len(_CART[s1.l]), len(_CART[s2.l])))
    for a, ca in zip(s1.exps, s1.coefs):
        for b, cb in zip(s2.exps, s2.coefs):
            p = a + b
            for i, c1 in enumerate(_CART[s1.l]):
                for j, c2 in enumerate(_CART[s2.l]):
                    e = 1.0
                    for d in range(3):
                        e *= _E(c1[d], c2[d], 0, s1.center[d] - s2.center[d], a, b)
                    out[i, j] += ca * cb * e * (math.pi / p) ** 1.5
    return out


def _eri_prim(hab, p, Pc, hcd, q, Qc) -> float:
    alpha = p * q / (p + q)
    X, Y, Z = Pc - Qc
    cache: dict = {}
    s = 0.0
    for (t, u, v), e1 in hab.items():
        for (tt, uu, vv), e2 in hcd.items():
            sign = -1.0 if (tt + uu + vv) % 2 else 1.0
            s += e1 * e2 * sign * _R(t + tt, u + uu, v + vv, 0, alpha, X, Y, Z, cache)
    return 2.0 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * s


def _offsets(shells):
    off, n = [], 0
    for sh in shells:
        off.append(n)
        n += sh.nfunc
    return off, n


def overlap_matrix(shells) -> np.ndarray:
    off, n = _offsets(shells)
    S = np.zeros((n, n))
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            S[off[i]:off[i] + si.nfunc, off[j]:off[j] + sj.nfunc] = _shell_overlap(si, sj)
    return S


def coulomb2c(shells) -> np.ndarray:
    """(P|Q) two-centre Coulomb matrix."""
    off, n = _offsets(shells)
    V = np.zeros((n, n))
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            if j > i:
                continue
            blk = np.zeros((si.nfunc, sj.nfunc))
            for a, ca in zip(si.exps, si.coefs):
                h1 = _herm_pair(si.l, si.center, a, 0, si.center, 0.0)
                for b, cb in zip(sj.exps, sj.coefs):
                    h2 = _herm_pair(sj.l, sj.center, b, 0, sj.center, 0.0)
                    for x in range(si.nfunc):
                        for y in range(sj.nfunc):
                            blk[x, y] += ca * cb * _eri_prim(h1[(x, 0)], a, si.center,
                                                              h2[(y, 0)], b, sj.center)
            V[off[i]:off[i] + si.nfunc, off[j]:off[j] + sj.nfunc] = blk
            V[off[j]:off[j] + sj.nfunc, off[i]:off[i] + si.nfunc] = blk.T
    return V


def coulomb3c(auxshells, shells) -> np.ndarray:
    """(P|mu nu) three-centre Coulomb tensor, shape (naux, nbf, nbf)."""
    aoff, na = _offsets(auxshells)
    off, n = _offsets(shells)
    T = np.zeros((na, n, n))
    pairs = {}
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            if j > i:
                continue
            plist = []
            for a, ca in zip(si.exps, si.coefs):
                for b, cb in zip(sj.exps, sj.coefs):
                    p = a + b
                    Pc = (a * si.center + b * sj.center) / p
                    plist.append((ca * cb, p, Pc,
                                  _herm_pair(si.l, si.center, a, sj.l, sj.center, b)))
            pairs[(i, j)] = plist
    for k, sk in enumerate(auxshells):
        prims = [(c, a, _herm_pair(sk.l, sk.center, a, 0, sk.center, 0.0))
                 for a, c in zip(sk.exps, sk.coefs)]
        for (i, j), plist in pairs.items():
            si, sj = shells[i], shells[j]
            blk = np.zeros((sk.nfunc, si.nfunc, sj.nfunc))
            for cc, q, hk in prims:
                for cab, p, Pc, hab in plist:
                    for x in range(sk.nfunc):
                        for y in range(si.nfunc):
                            for z in range(sj.nfunc):
                                blk[x, y, z] += cc * cab * _eri_prim(
                                    hab[(y, z)], p, Pc, hk[(x, 0)], q, sk.center)
            T[aoff[k]:aoff[k] + sk.nfunc, off[i]:off[i] + si.nfunc, off[j]:off[j] + sj.nfunc] = blk
            T[aoff[k]:aoff[k] + sk.nfunc, off[j]:off[j] + sj.nfunc, off[i]:off[i] + si.nfunc] = \
                blk.transpose(0, 2, 1)
    return T


# nuclear charges of the elements the fixtures use (all-electron basis sets, no ECP)
NUCLEAR_CHARGE = {"H": 1.0, "He": 2.0, "C": 6.0, "N": 7.0, "O": 8.0, "F": 9.0}


def _overlap_1d(i: int, j: int, qx: float, a: float, b: float) -> float:
    return _E(i, j, 0, qx, a, b) * math.sqrt(math.pi / (a + b)) if j >= 0 else 0.0


def _kinetic_1d(i: int, j: int, qx: float, a: float, b: float) -> float:
    """<x^i e^{-a x^2}| -1/2 d^2/dx^2 |x^j e^{-b x^2}> from
    d^2/dx^2 x^j e^{-b x^2} = [j(j-1) x^{j-2} - 2b(2j+1) x^j + 4b^2 x^{j+2}] e^{-b x^2}."""
    t = -2.0 * b * (2 * j + 1) * _overlap_1d(i, j, qx, a, b) + 4.0 * b * b * _overlap_1d(i, j + 2, qx, a, b)
    if j >= 2:
        t += j * (j - 1) * _overlap_1d(i, j - 2, qx, a, b)
    return -0.5 * t


def kinetic_matrix(shells) -> np.ndarray:
    """T_{mu nu} = <mu| -1/2 nabla^2 |nu> (the reference's AOKinetic, aokinetic.cc)."""
    off, n = _offsets(shells)
    K = np.zeros((n, n))
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            blk = np.zeros((si.nfunc, sj.nfunc))
            Q = si.center - sj.center
            for a, ca in zip(si.exps, si.coefs):
                for b, cb in zip(sj.exps, sj.coefs):
                    for x, c1 in enumerate(_CART[si.l]):
                        for y, c2 in enumerate(_CART[sj.l]):
                            S1 = [_overlap_1d(c1[d], c2[d], Q[d], a, b) for d in range(3)]
                            T1 = [_kinetic_1d(c1[d], c2[d], Q[d], a, b) for d in range(3)]
                            blk[x, y] += ca * cb * (T1[0] * S1[1] * S1[2] + S1[0] * T1[1] * S1[2]
                                                    + S1[0] * S1[1] * T1[2])
            K[off[i]:off[i] + si.nfunc, off[j]:off[j] + sj.nfunc] = blk
    return K


def nuclear_attraction(shells, atoms) -> np.ndarray:
    """V_{mu nu} = -sum_A Z_A <mu| 1/|r - R_A| |nu> (AOMultipole::FillPotential with the
    nuclear charges, aomultipole.cc:2095-2105): Hermite expansion with R_{tuv}(p, P - C)."""
    off, n = _offsets(shells)
    V = np.zeros((n, n))
    for i, si in enumerate(shells):
        for j, sj in enumerate(shells):
            blk = np.zeros((si.nfunc, sj.nfunc))
            for a, ca in zip(si.exps, si.coefs):
                for b, cb in zip(sj.exps, sj.coefs):
                    p = a + b
                    Pc = (a * si.center + b * sj.center) / p
                    h = _herm_pair(si.l, si.center, a, sj.l, sj.center, b)
                    for sym, Cpos in atoms:
                        X, Y, Z = Pc - Cpos
                        cache: dict = {}
                        for x in range(si.nfunc):
                            for y in range(sj.nfunc):
                                acc = 0.0
                                for (t, u, v), e in h[(x, y)].items():
                                    acc += e * _R(t, u, v, 0, p, X, Y, Z, cache)
                                blk[x, y] -= NUCLEAR_CHARGE[sym] * ca * cb * 2.0 * math.pi / p * acc
            V[off[i]:off[i] + si.nfunc, off[j]:off[j] + sj.nfunc] = blk
    return V


def core_hamiltonian(xyz_path: str, basis_path: str) -> np.ndarray:
    """Kinetic plus nuclear attraction, the matrix test_ppm.cc diagonalises for its MOs."""
    shells = build_shells(xyz_path, basis_path)
    return kinetic_matrix(shells) + nuclear_attraction(shells, read_xyz(xyz_path))


@lru_cache(maxsize=8)
def ao_integrals(xyz_path: str, basis_path: str, aux_basis_path: str | None = None):
    """Return (S_dft, S_aux, V_aux, T3c) for a molecule and basis (aux defaults to dft)."""
    shells = build_shells(xyz_path, basis_path)
    aux = shells if aux_basis_path is None else build_shells(xyz_path, aux_basis_path)
    S = overlap_matrix(shells)
    Saux = overlap_matrix(aux)
    V = coulomb2c(aux)
    T = coulomb3c(aux, shells)
    return S, Saux, V, T
