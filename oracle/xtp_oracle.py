"""FP64 CPU restatement of the VOTCA-XTP GW-BSE hot path (the oracle).

TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this module, and only as the checker / CPU baseline.
The product path (votca_amd, libxtpgpu.so) never imports or calls it.

Every function cites the reference file:line whose loop it restates. Layout follows
the reference: an Mmn tensor is a list of mtotal slices, each (ntotal x P); here it is
one numpy array M[m, n, P] (m = slice index, n = row, P = aux column), so that
M[m] is exactly the reference's Mmn_[m] (xtp/include/votca/xtp/threecenter.h:125).

What pins each function (tests/test_oracle_golden.py): fill_mmn / multiply_right by the
threecenter_gwbse fixtures; rpa_epsilon by the rpa fixtures and brute force; rpa_update_energies
by its closed form; rpa_h2p by the rpa_h2p known answers and the RPA-matrix spectrum; ppm_construct
and sigma_ppm_* by the ppm known answers and sigma_ppm fixtures; sigma_x by x_ref; qp_* by gw/ref
and the single-pole closed form; bse_dense / BSEOperatorOracle by the bse_operator fixtures and
the bse energy / coefficient fixtures; DavidsonOracle.solve / solve_ham by dense eigensolvers and
the bse fixtures; screening_at / perturbative_dynamical_screening by the bse dynamic fixtures;
SigmaCDAOracle by the sigma_cda fixtures; pseudo_invsqrt_gwbse through the threecenter fixtures;
the synthetic generators by their distribution tests. Parity unpinned: the Davidson's iteration
counts and intermediate iterates (only converged eigenpairs are pinned), and the phase convention
of a complex harmonic-Ritz eigenvector before its real part is taken (LAPACK's here, Eigen's QZ
back-substitution in the reference).
"""
from __future__ import annotations

import math

import numpy as np

HRT2EV = 27.21138602  # tools/include/votca/tools/constants.h:53


# ---------------------------------------------------------------- three-centre
def pseudo_invsqrt_gwbse(V: np.ndarray, S: np.ndarray, etol: float = 5e-7) -> np.ndarray:
    """((S^-1/2 V S^-1/2)^-1/2 S^-1/2)^T -- xtp/src/libxtp/aomatrices/aomatrix.cc:53-86."""
    ev, U = np.linalg.eigh(S)
    d = np.where(ev < etol, 0.0, 1.0 / np.sqrt(np.maximum(ev, etol)))
    Ssqrt = (U * d) @ U.T
    ortho = Ssqrt @ V @ Ssqrt
    ev2, U2 = np.linalg.eigh(ortho)
    d2 = np.where(ev2 < etol, 0.0, 1.0 / np.sqrt(np.maximum(ev2, etol)))
    Vm1 = (U2 * d2) @ U2.T
    return (Vm1 @ Ssqrt).T


def fill_mmn(T3c: np.ndarray, C: np.ndarray, mmin: int, mmax: int, nmin: int, nmax: int,
             Saux: np.ndarray, Vaux: np.ndarray) -> np.ndarray:
    """TCMatrix_gwbse::Fill -- xtp/src/libxtp/threecenter.cc:72-90 with Fill3cMO
    (xtp/src/libxtp/libint2_calls.cc:595-651): M[m](n,Q) = sum_{mu,nu} C[mu,n] (Q|mu nu) C[nu,m],
    then right-multiplied by the aux inverse square root."""
    Cm = C[:, mmin:mmax + 1]
    Cn = C[:, nmin:nmax + 1]
    M = np.einsum("qab,an,bm->mnq", T3c, Cn, Cm, optimize=True)
    return multiply_right(M, pseudo_invsqrt_gwbse(Vaux, Saux, 5e-7))


def multiply_right(M: np.ndarray, A: np.ndarray) -> np.ndarray:
    """TCMatrix_gwbse::MultiplyRightWithAuxMatrix -- xtp/src/libxtp/threecenter.cc:54-65."""
    return np.einsum("mnp,pq->mnq", M, A, optimize=True)


# ---------------------------------------------------------------- RPA
def rpa_update_energies(dft: np.ndarray, gw: np.ndarray, qpmin: int, homo: int,
                        rpamin: int, rpamax: int) -> np.ndarray:
    """RPA::UpdateRPAInputEnergies / ShiftUncorrectedEnergies / getMaxCorrection --
    xtp/src/libxtp/gwbse/rpa.cc:32-73."""
    rpatotal = rpamax - rpamin + 1
    e = dft[rpamin:rpamin + rpatotal].copy()
    gwsize = len(gw)
    e[qpmin - rpamin:qpmin - rpamin + gwsize] = gw
    lumo = homo + 1
    qpmax = qpmin + gwsize - 1

    def maxcorr(lo, hi):
        r = hi - lo + 1
        return float(np.max(np.abs(e[lo:lo + r] - dft[lo - rpamin:lo - rpamin + r])))

    occ = maxcorr(qpmin, homo)
    virt = maxcorr(lumo, qpmax)
    e[:qpmin] -= occ
    if rpamax - qpmax > 0:
        e[len(e) - (rpamax - qpmax):] += virt
    return e


def rpa_weights(energies: np.ndarray, homo: int, rpamin: int, rpamax: int,
                freq, kind: str, eta: float = 1e-4) -> np.ndarray:
    """Per-transition weight d[v, c] of eps = 1 + sum_v M_v^T diag(d_v) M_v.

    kind 'imag'    : 4 D/(D^2+w^2)                       rpa.cc:117-118
    kind 'real'    : 2[(D-w)/((D-w)^2+eta^2) + (D+w)/((D+w)^2+eta^2)]   rpa.cc:119-125
    kind 'complex' : -2 chi, chi as rpa.cc:182-190 (sign folded in: rpa.cc:197)
    """
    lumo = homo + 1
    n_occ = lumo - rpamin
    n_unocc = rpamax - lumo + 1
    D = energies[n_occ:n_occ + n_unocc][None, :] - energies[:n_occ][:, None]
    if kind == "imag":
        w = float(freq)
        return 4.0 * D / (D * D + w * w)
    if kind == "real":
        w = float(freq)
        a = D - w
        b = D + w
        return 2.0 * (a / (a * a + eta * eta) + b / (b * b + eta * eta))
    if kind == "complex":
        wr, wi = complex(freq).real, complex(freq).imag
        dm = wr - D
        dp = wr + D
        s1 = (wi + eta) ** 2
        s2 = (wi - eta) ** 2
        chi = dm / (dm * dm + s1) - dp / (dp * dp + s2)
        return -2.0 * chi
    raise ValueError(kind)


def rpa_epsilon(M: np.ndarray, energies: np.ndarray, homo: int, rpamin: int, rpamax: int,
                freq, kind: str, eta: float = 1e-4) -> np.ndarray:
    """RPA::calculate_epsilon<imag> / calculate_epsilon_r(complex) --
    xtp/src/libxtp/gwbse/rpa.cc:75-140 and 145-202, with OpenMP_CUDA::A_TDA
    (xtp/src/libxtp/openmp_cuda.cc:236-256): eps = 1 + sum_v M_v^T diag(d_v) M_v with
    M_v = Mmn_[v].bottomRows(n_unocc)."""
    lumo = homo + 1
    n_occ = lumo - rpamin
    n_unocc = rpamax - lumo + 1
    d = rpa_weights(energies, homo, rpamin, rpamax, freq, kind, eta)
    P = M.shape[2]
    eps = np.zeros((P, P))
    for v in range(n_occ):
        Mv = M[v][M.shape[1] - n_unocc:, :]
        eps += Mv.T @ (d[v][:, None] * Mv)
    eps[np.diag_indices(P)] += 1.0
    return eps


# ---------------------------------------------------------------- BSE operator
def rpa_epsilon_columns(M, energies: np.ndarray, homo: int, rpamin: int, rpamax: int, freq, kind: str,
                        cols, eta: float = 1e-4) -> np.ndarray:
    """Columns `cols` of RPA::calculate_epsilon (xtp/src/libxtp/gwbse/rpa.cc:75-140): the per-level
    accumulation eps += M_v^T diag(d_v) M_v of rpa.cc:88-136 (M_v = Mmn_[v].bottomRows(n_unocc))
    restricted to the requested columns, plus the identity. M is indexable by slice (occupied
    levels v - rpamin) and returns the (ntotal x P) slice, so only the occupied slices are read."""
    d = rpa_weights(energies, homo, rpamin, rpamax, freq, kind, eta)
    n_occ = homo + 1 - rpamin
    n_unocc = rpamax - homo
    cols = np.asarray(cols, dtype=np.int64)
    out = None
    for v in range(n_occ):
        Mv = np.asarray(M[v])[-n_unocc:, :]
        blk = Mv.T @ (d[v][:, None] * Mv[:, cols])
        out = blk if out is None else out + blk
    out = out.astype(np.result_type(out.dtype, np.float64))
    out[cols, np.arange(len(cols))] += 1.0
    return out


class BSEOperatorOracle:
    """BSE_OPERATOR<cqp,cx,cd,cd2> -- xtp/src/libxtp/gwbse/bse_operator.cc:28-188,
    xtp/include/votca/xtp/bse_operator.h:40-87, with the OpenMP_CUDA CPU branches
    (xtp/src/libxtp/openmp_cuda.cc:259-478). Compound index I(v,c) = ctotal*v + c
    (xtp/include/votca/xtp/vc2index.h:36-54)."""

    def __init__(self, cqp, cx, cd, cd2, eps_inv, M, Hqp, homo, rpamin, qpmin, vmin, cmax):
        assert not (cd != 0 and cd2 != 0)
        self.c = (cqp, cx, cd, cd2)
        self.eps_inv = np.asarray(eps_inv, dtype=np.float64)
        self.M = M
        self.Hqp = Hqp
        self.homo, self.rpamin, self.qpmin, self.vmin, self.cmax = homo, rpamin, qpmin, vmin, cmax
        self.bse_vmax = homo
        self.bse_cmin = homo + 1
        self.vtotal = self.bse_vmax - vmin + 1
        self.ctotal = cmax - self.bse_cmin + 1
        self.size = self.vtotal * self.ctotal

    def I(self, v, c):
        return self.ctotal * v + c

    def hqp_row(self, v1, c1):
        """BSE_OPERATOR::Hqp_row -- bse_operator.cc:121-132."""
        cqp = self.c[0]
        R = np.zeros((self.ctotal, self.vtotal))
        cmin = self.vtotal
        R[:, v1] += cqp * self.Hqp[cmin:cmin + self.ctotal, c1 + cmin]
        R[c1, :] -= cqp * self.Hqp[:self.vtotal, v1]
        return R.reshape(-1, order="F")

    def matmul(self, X: np.ndarray) -> np.ndarray:
        """BSE_OPERATOR::matmul -- bse_operator.cc:39-119 (row construction then Hx blocks)."""
        cqp, cx, cd, cd2 = self.c
        X = np.asarray(X, dtype=np.float64)
        vmin = self.vmin - self.rpamin
        cmin = self.bse_cmin - self.rpamin
        ct, vt = self.ctotal, self.vtotal
        out = np.zeros((self.size, X.shape[1]))
        ei = self.eps_inv
        for c1 in range(ct):
            if cd != 0:
                T = -cd * self.M[c1 + cmin][cmin:cmin + ct, :] * ei[None, :]
            elif cd2 != 0:
                T = -cd2 * self.M[c1 + cmin][vmin:vmin + vt, :] * ei[None, :]
            for v1 in range(vt):
                row = np.zeros(self.size)
                if cd != 0:
                    mat = self.M[v1 + vmin][vmin:vmin + vt, :]
                    row += (T @ mat.T).reshape(-1, order="F")  # (ct x vt), c fast
                if cd2 != 0:
                    mat = self.M[v1 + vmin][cmin:cmin + ct, :]
                    row += (mat @ T.T).reshape(-1, order="F")  # (ct x vt)
                if cqp != 0:
                    row += self.hqp_row(v1, c1)
                out[self.I(v1, c1)] = row @ X
        if cx > 0:
            for v1 in range(vt):
                M1 = cx * self.M[v1 + vmin][cmin:cmin + ct, :]
                for v2 in range(v1, vt):
                    B = M1 @ self.M[v2 + vmin][cmin:cmin + ct, :].T
                    out[v1 * ct:(v1 + 1) * ct] += B @ X[v2 * ct:(v2 + 1) * ct]
                    if v1 != v2:
                        out[v2 * ct:(v2 + 1) * ct] += B.T @ X[v1 * ct:(v1 + 1) * ct]
        return out

    def matmul_rows(self, X: np.ndarray, rows) -> np.ndarray:
        """Rows `rows` (BSE indices I(v,c)) of BSE_OPERATOR::matmul: the row construction of
        bse_operator.cc:61-96 (PrepareMatrix1/2 for Hd / Hd2, Hqp_row) for those rows only, then
        the Hx block term of bse_operator.cc:98-116 regrouped as cx * V_row (V^T X) with
        V_{I(v,c),P} = M^P_{vc} (the same sum over v2, c2, P as MultiplyBlocks' v-block pairs).
        Only the slices the rows touch are read from self.M (any object indexable by slice,
        returning the (ntotal x P) slice), so full-size operators can be checked row by row."""
        cqp, cx, cd, cd2 = self.c
        X = np.asarray(X, dtype=np.float64)
        vmin = self.vmin - self.rpamin
        cmin = self.bse_cmin - self.rpamin
        ct, vt = self.ctotal, self.vtotal
        ei = self.eps_inv
        rows = [int(r) for r in rows]
        out = np.zeros((len(rows), X.shape[1]))
        for t, r in enumerate(rows):
            v1, c1 = divmod(r, ct)
            row = np.zeros(self.size)
            if cd != 0:
                T = -cd * self.M[c1 + cmin][cmin:cmin + ct, :] * ei[None, :]
                row += (T @ self.M[v1 + vmin][vmin:vmin + vt, :].T).reshape(-1, order="F")
            if cd2 != 0:
                T = -cd2 * self.M[c1 + cmin][vmin:vmin + vt, :] * ei[None, :]
                row += (self.M[v1 + vmin][cmin:cmin + ct, :] @ T.T).reshape(-1, order="F")
            if cqp != 0:
                row += self.hqp_row(v1, c1)
            out[t] = row @ X
        if cx != 0:
            W = None
            for v2 in range(vt):
                blk = self.M[v2 + vmin][cmin:cmin + ct, :].T @ X[v2 * ct:(v2 + 1) * ct]
                W = blk if W is None else W + blk
            for t, r in enumerate(rows):
                v1, c1 = divmod(r, ct)
                out[t] += cx * (self.M[v1 + vmin][cmin + c1, :] @ W)
        return out

    def diagonal(self) -> np.ndarray:
        """BSE_OPERATOR::diagonal -- bse_operator.cc:134-175."""
        cqp, cx, cd, cd2 = self.c
        vmin = self.vmin - self.rpamin
        cmin = self.bse_cmin - self.rpamin
        ct, vt = self.ctotal, self.vtotal
        ei = self.eps_inv
        res = np.zeros(self.size)
        for v in range(vt):
            for c in range(ct):
                e = 0.0
                if cx != 0:
                    e += cx * float(np.sum(self.M[v + vmin][cmin + c] ** 2))
                if cqp != 0:
                    cq = vt
                    e += cqp * (self.Hqp[c + cq, c + cq] - self.Hqp[v, v])
                if cd != 0:
                    e -= cd * float(np.sum(self.M[c + cmin][c + cmin] * ei * self.M[v + vmin][v + vmin]))
                if cd2 != 0:
                    e -= cd2 * float(np.sum(self.M[c + cmin][v + vmin] * ei * self.M[v + vmin][c + cmin]))
                res[self.I(v, c)] = e
        return res


def adjust_hqp_size(Hqp: np.ndarray, rpa_energies: np.ndarray, homo: int, rpamin: int, qpmin: int, qpmax: int,
                    vmin: int, cmax: int) -> np.ndarray:
    """BSE::AdjustHqpSize (xtp/src/libxtp/gwbse/bse.cc:147-186): the QP Hamiltonian of the GW window
    [qpmin, qpmax] cut or extended to the BSE window [vmin, cmax]; levels outside the GW window get
    their RPA input energy on the diagonal (RPA offset vmin - rpamin)."""
    vt = homo - vmin + 1
    ct = cmax - homo
    hs = vt + ct
    gw = qpmax - qpmin + 1
    off = vmin - rpamin
    out = np.zeros((hs, hs))
    if vmin >= qpmin:
        start = vmin - qpmin
        if cmax <= qpmax:
            out[:, :] = Hqp[start:start + hs, start:start + hs]
        else:
            vo = gw - start
            out[:vo, :vo] = Hqp[start:start + vo, start:start + vo]
            ve = cmax - qpmax
            out[np.arange(hs - ve, hs), np.arange(hs - ve, hs)] = rpa_energies[off + vo:off + vo + ve]
    if vmin < qpmin:
        oe = qpmin - vmin
        out[np.arange(oe), np.arange(oe)] = rpa_energies[off:off + oe]
        out[oe:oe + gw, oe:oe + gw] = Hqp
        if cmax > qpmax:
            vo = oe + gw
            ve = cmax - qpmax
            out[np.arange(hs - ve, hs), np.arange(hs - ve, hs)] = rpa_energies[off + vo:off + vo + ve]
    return out


def rpa_h2p(M: np.ndarray, energies: np.ndarray, homo: int, rpamin: int, rpamax: int):
    """RPA::Diagonalize_H2p / Calculate_H2p_AmB / Calculate_H2p_ApB / Diagonalize_H2p_C
    (xtp/src/libxtp/gwbse/rpa.cc:204-345, closed shell, no QSGW rotation):
    (A-B)_{vc} = e_c - e_v, A+B = diag(A-B) + 4 V V^T with V_{I(v,c),P} = M^P_{vc};
    C = (A-B)^1/2 (A+B) (A-B)^1/2 = U diag(omega^2) U^T;
    E_c = -1/4 (tr(A+B) + sum(A-B)) + 1/2 sum(omega); (X+Y)_s = omega_s^-1/2 (A-B)^1/2 U_s.
    Returns (omega ascending, XpY, E_c)."""
    lumo = homo + 1
    n_occ = lumo - rpamin
    n_unocc = rpamax - lumo + 1
    e = np.asarray(energies, dtype=np.float64)
    amb = (e[n_occ:n_occ + n_unocc][None, :] - e[:n_occ][:, None]).reshape(-1)
    V = M[:n_occ, n_occ:n_occ + n_unocc, :].reshape(n_occ * n_unocc, -1)
    apb = 4.0 * (V @ V.T) + np.diag(amb)
    erpa = -0.25 * (np.trace(apb) + amb.sum())
    r = np.sqrt(amb)
    lam, U = np.linalg.eigh(r[:, None] * apb * r[None, :])
    if lam.min() <= 0.0:
        raise RuntimeError("Detected non-positive eigenvalue.")
    omega = np.sqrt(lam)
    erpa += 0.5 * omega.sum()
    xpy = r[:, None] * U / np.sqrt(omega)[None, :]
    return omega, xpy, erpa


def bse_dense(cqp, cx, cd, cd2, eps_inv, M, Hqp, homo, rpamin, vmin, cmax) -> np.ndarray:
    """Explicit BSE matrix from the closed-form matrix elements the reference's
    matmul contracts (bse_operator.cc:61-116): H[(v,c),(v',c')] =
      cqp (d_vv' Hqp_cc' - d_cc' Hqp_v'v) + cx sum_P M^P_vc M^P_v'c'
      - cd sum_P eps^-1_P M^P_cc' M^P_vv' - cd2 sum_P eps^-1_P M^P_cv' M^P_vc'.
    Index independent of the row loop, so it pins the loop restatement above."""
    vm = vmin - rpamin
    cm = homo + 1 - rpamin
    vt = homo - vmin + 1
    ct = cmax - homo
    ei = np.asarray(eps_inv)
    Mvc = M[vm:vm + vt, cm:cm + ct, :]          # M^P_{v c}
    Mcc = M[cm:cm + ct, cm:cm + ct, :]          # M^P_{c c'}
    Mvv = M[vm:vm + vt, vm:vm + vt, :]          # M^P_{v v'}
    Mcv = M[cm:cm + ct, vm:vm + vt, :]          # M^P_{c v'}
    H = np.zeros((vt, ct, vt, ct))
    if cx:
        H += cx * np.einsum("acp,bdp->acbd", Mvc, Mvc)
    if cd:
        H -= cd * np.einsum("cdp,abp,p->acbd", Mcc, Mvv, ei)
    if cd2:
        H -= cd2 * np.einsum("cbp,adp,p->acbd", Mcv, Mvc, ei)
    if cqp:
        Hc = Hqp[vt:vt + ct, vt:vt + ct]
        Hv = Hqp[:vt, :vt]
        H += cqp * np.einsum("ab,cd->acbd", np.eye(vt), Hc)
        H -= cqp * np.einsum("ba,cd->acbd", Hv, np.eye(ct))
    return H.reshape(vt * ct, vt * ct)


# ---------------------------------------------------------------- Davidson
class DavidsonOracle:
    """DavidsonSolver -- xtp/include/votca/xtp/davidsonsolver.h:64-280 and
    xtp/src/libxtp/davidsonsolver.cc:29-558 (SYMM matrix type: Ritz, DPR/Olsen,
    'min'/'safe'/'max' size update, CGS twice, restart)."""

    TOL = {"loose": 1e-3, "normal": 1e-4, "strict": 1e-5, "lapack": 1e-9}

    def __init__(self, tol=1e-4, correction="DPR", update="safe", iter_max=50,
                 max_search_space=0):
        self.tol = self.TOL.get(tol, tol) if isinstance(tol, str) else float(tol)
        self.correction = correction
        self.update = update
        self.iter_max = iter_max
        self.max_search_space = max_search_space
        self.info_converged = False
        self.num_iterations = 0

    def size_update(self, neigen):  # davidsonsolver.cc:149-170
        if self.update == "min":
            return neigen
        if self.update == "safe":
            return int(1.5 * neigen) if neigen < 20 else neigen + 10
        return 2 * neigen

    @staticmethod
    def gramschmidt(Q, nstart):  # davidsonsolver.cc:456-497
        nupdate = Q.shape[1] - nstart
        norms = np.linalg.norm(Q[:, nstart:], axis=0)
        for rep in range(2):
            if nstart > 0:
                Q[:, nstart:] -= Q[:, :nstart] @ (Q[:, :nstart].T @ Q[:, nstart:])
                Q[:, nstart:] /= np.linalg.norm(Q[:, nstart:], axis=0)
            for j in range(nstart + 1, Q.shape[1]):
                Qb = Q[:, nstart:j]
                Q[:, j] -= Qb @ (Qb.T @ Q[:, j])
                nrm = np.linalg.norm(Q[:, j])
                if rep == 1 and nrm <= 1e-12 * norms[j - nstart]:
                    raise RuntimeError("Linear dependencies in Gram-Schmidt.")
                Q[:, j] /= nrm
        return Q

    def solve(self, matmul, diag, neigen):
        diag = np.asarray(diag, dtype=np.float64)
        n = diag.size
        if self.max_search_space < neigen:
            self.max_search_space = neigen * 5
        if self.max_search_space > n:
            self.max_search_space = n
        size_guess = 2 * neigen
        restart_size = size_guess
        idx = np.argsort(diag, kind="stable")
        V = np.zeros((n, size_guess))
        for j in range(size_guess):
            V[idx[j], j] = 1.0
        size_update = self.size_update(neigen)
        AV = None
        T = None
        self.info_converged = False
        self.history = []
        for it in range(self.iter_max):
            self.num_iterations = it
            if it == 0:
                AV = matmul(V)
                T = V.T @ AV
            else:
                old = AV.shape[1]
                new = V[:, old:]
                AVn = matmul(new)
                AV = np.hstack([AV, AVn])
                Tn = np.zeros((V.shape[1], V.shape[1]))
                Tn[:old, :old] = T
                Tn[:, old:] = V.T @ AVn
                Tn[old:, :old] = Tn[:old, old:].T
                T = Tn
            lam, U = np.linalg.eigh(T)
            need = min(T.shape[1], max(restart_size, size_update))
            lam = lam[:need]
            U = U[:, :need]
            q = V @ U
            res = AV @ U - q * lam[None, :]
            rn = np.linalg.norm(res, axis=0)
            self.history.append(float(rn[:neigen].max()))
            conv = rn[:size_update] < self.tol
            if np.all(conv[:neigen]):
                self.info_converged = True
                break
            if it == self.iter_max - 1:
                lam = lam.copy()
                q = q.copy()
                for i in range(neigen):
                    if not conv[i]:
                        lam[i] = 0.0
                        q[:, i] = 0.0
                break
            newv = []
            for j in range(size_update):
                if conv[j]:
                    continue
                if self.correction == "DPR":
                    w = -res[:, j] / (diag - lam[j])
                else:
                    d = -res[:, j] / (diag - lam[j])
                    xq = q[:, j]
                    num = -xq @ d
                    den = -xq @ (-xq / (diag - lam[j]))
                    w = d + (num / den) * xq
                w = np.where(np.isfinite(w), w, 0.0)
                newv.append(w / np.linalg.norm(w))
            nup = len(newv)
            V = np.hstack([V, np.array(newv).T])
            V = self.gramschmidt(V, V.shape[1] - nup)
            if V.shape[1] > self.max_search_space:  # restart, davidsonsolver.cc:499-522
                newV = np.hstack([q[:, :restart_size], V[:, -nup:]])
                AV = AV @ U[:, :restart_size]
                T = newV[:, :restart_size].T @ AV
                V = newV
        self.eigenvalues = lam[:neigen]
        vecs = q[:, :neigen]
        nrm = np.linalg.norm(vecs, axis=0)
        nrm[nrm == 0] = 1.0
        self.eigenvectors = vecs / nrm
        return self.eigenvalues, self.eigenvectors


    def solve_ham(self, matmul, diag_a, diag_b, neigen, guess=None):
        """HAM matrix type as called by BSE::Solve_nonhermitian_Davidson (bse.cc:320-358): guess
        from BuildFullBSEXRankedInitialGuess (bse_initialization.h:47-93), diagonal [Ad; -Ad]
        (bseoperator_btda.h:85-91), projections T = V^T H V and B = V^T H H V (davidsonsolver.h:
        239-280), harmonic Ritz pairs (davidsonsolver.cc:241-332) and the QR restart
        (davidsonsolver.cc:499-518). `matmul` applies H = [[A, B], [-B, -A]] to 2N x k blocks.
        `guess` (2N x k) replaces the ranked guess: the initial-guess overload of
        DavidsonSolver::solve (davidsonsolver.h:123-152), restart size min(k, max space)."""
        import scipy.linalg as sla
        diag_a = np.asarray(diag_a, dtype=np.float64)
        diag_b = np.asarray(diag_b, dtype=np.float64)
        nh = diag_a.size
        n = 2 * nh
        diag = np.concatenate([diag_a, -diag_a])
        if self.max_search_space < neigen:
            self.max_search_space = neigen * 5
        if self.max_search_space > n:
            self.max_search_space = n
        if guess is None:
            nguess = min(nh, max(4 * neigen, 8))
            omega = np.sqrt(np.maximum(0.0, (diag_a - diag_b) * (diag_a + diag_b)))
            rank = sorted(range(nh), key=lambda i: (omega[i], diag_a[i]))
            V = np.zeros((n, nguess))
            for j in range(nguess):
                V[rank[j], j] = 1.0
        else:  # unit columns: gramschmidt (as the reference's) never normalizes column 0
            V = np.array(guess, dtype=np.float64)
            V = V / np.linalg.norm(V, axis=0)
            nguess = V.shape[1]
        V = self.gramschmidt(V, 0)
        restart_size = min(nguess, self.max_search_space)
        size_update = self.size_update(neigen)
        self.info_converged = False
        self.history = []
        AV = AAV = None
        for it in range(self.iter_max):
            self.num_iterations = it
            old = 0 if AV is None else AV.shape[1]
            AVn = matmul(V[:, old:])
            AAVn = matmul(AVn)
            AV = AVn if AV is None else np.hstack([AV, AVn])
            AAV = AAVn if AAV is None else np.hstack([AAV, AAVn])
            T = V.T @ AV
            B = V.T @ AAV
            w, vr = sla.eig(T, B)
            keep, taken = [], set()
            for i in range(len(w)):  # first member of each complex pair kept (real part)
                if w[i].imag != 0:
                    if i in taken:
                        continue
                    for k in range(i + 1, len(w)):
                        if k not in taken and abs(w[k].real - w[i].real) < 1e-9 and abs(w[k].imag + w[i].imag) < 1e-9:
                            taken.add(k)
                            break
                keep.append(i)
            ev = w.real[keep]
            U = vr.real[:, keep]
            U = U / np.linalg.norm(U, axis=0)
            need = min(T.shape[1], max(restart_size, size_update))
            order = np.argsort(ev, kind="stable")[::-1][:need]
            U = U[:, order]
            lam = np.einsum("ij,ik,kj->j", U, T, U)
            q = V @ U
            res = AV @ U - q * lam[None, :]
            rn = np.linalg.norm(res, axis=0)
            self.history.append(float(rn[:neigen].max()))
            # a guess narrower than the size update leaves fewer pairs than size_update
            conv = rn[:size_update] < self.tol
            if np.all(conv[:neigen]):
                self.info_converged = True
                break
            if it == self.iter_max - 1:
                lam = lam.copy()
                q = q.copy()
                for i in range(neigen):
                    if not conv[i]:
                        lam[i] = 0.0
                        q[:, i] = 0.0
                break
            newv = []
            for j in range(len(conv)):
                if conv[j]:
                    continue
                if self.correction == "DPR":
                    wv = -res[:, j] / (diag - lam[j])
                else:
                    d = -res[:, j] / (diag - lam[j])
                    xq = q[:, j]
                    wv = d + ((-xq @ d) / (-xq @ (-xq / (diag - lam[j])))) * xq
                wv = np.where(np.isfinite(wv), wv, 0.0)
                newv.append(wv / np.linalg.norm(wv))
            nup = len(newv)
            V = np.hstack([V, np.array(newv).T])
            V = self.gramschmidt(V, V.shape[1] - nup)
            if V.shape[1] > self.max_search_space:
                Q, _ = np.linalg.qr(U[:, :restart_size])
                V = np.hstack([V[:, :V.shape[1] - nup] @ Q, V[:, -nup:]])
                AV = AV @ Q
                AAV = AAV @ Q
        self.eigenvalues = lam[:neigen]
        vecs = q[:, :neigen]
        nrm = np.linalg.norm(vecs, axis=0)
        nrm[nrm == 0] = 1.0
        vecs = vecs / nrm
        nx = np.sum(vecs[:nh] ** 2, axis=0)
        ny = np.sum(vecs[nh:] ** 2, axis=0)
        with np.errstate(divide="ignore", invalid="ignore"):
            s = np.sqrt(1.0 / (nx - ny))
        self.X = vecs[:nh] * s
        self.Y = vecs[nh:] * s
        return self.eigenvalues, self.X, self.Y


# ---------------------------------------------------------------- synthetic inputs
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        x = (x + np.uint64(0x9E3779B97F4A7C15))
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def synth_normal(seed: int, p, i, j) -> np.ndarray:
    """Host twin of the device generator (votca_amd/csrc/kernels_misc.h synth_normal):
    splitmix64 hash of (seed, P, i, j), Irwin-Hall sum of 12 uniform 16-bit integers."""
    p = np.asarray(p, dtype=np.uint64)
    i = np.asarray(i, dtype=np.uint64)
    j = np.asarray(j, dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = _splitmix64(np.uint64(seed) + np.zeros(np.broadcast(p, i, j).shape, dtype=np.uint64))
        h = _splitmix64(h ^ (p * np.uint64(0x100000001B3)))
        h = _splitmix64(h ^ (i * np.uint64(0xC2B2AE3D27D4EB4F)))
        h = _splitmix64(h ^ (j * np.uint64(0x165667B19E3779F9)))
        s = np.zeros(h.shape, dtype=np.uint64)
        f = np.uint64(0xFFFF)
        for _ in range(3):
            h = _splitmix64(h)
            s += (h & f) + ((h >> np.uint64(16)) & f) + ((h >> np.uint64(32)) & f) + ((h >> np.uint64(48)) & f)
    return (s.astype(np.float64) + 6.0) * (1.0 / 65536.0) - 6.0


def synthetic_mmn(seed: int, mmin: int, mmax: int, nmin: int, nmax: int, aux: int,
                  sigma0: float = 0.3, decay: float = 3.0) -> np.ndarray:
    """M[m, n, P] = sigma_P * synth_normal(seed, P, min(m,n), max(m,n)),
    sigma_P = sigma0 * exp(-decay * P / aux)."""
    m = np.arange(mmin, mmax + 1)[:, None, None]
    n = np.arange(nmin, nmax + 1)[None, :, None]
    P = np.arange(aux)[None, None, :]
    a = np.minimum(m, n)
    b = np.maximum(m, n)
    # math.exp: the same libm exp the library's host side uses for its sigma table
    sig = np.array([sigma0 * math.exp(-decay * float(q) / float(aux)) for q in range(aux)])
    return synth_normal(seed, P, a, b) * sig[None, None, :]


def synthetic_energies(seed: int, n_occ: int, n_virt: int):
    """Occupied levels sorted uniform in [-20,-5] eV, virtual sorted uniform in [0,30] eV, in Ha."""
    rng = np.random.default_rng(seed)
    eo = np.sort(rng.uniform(-20.0, -5.0, n_occ))
    ev = np.sort(rng.uniform(0.0, 30.0, n_virt))
    return np.concatenate([eo, ev]) / HRT2EV


# ---------------------------------------------------------------- GW: exchange, PPM, Sigma_c
def sigma_x(M: np.ndarray, homo: int, rpamin: int, qpmin: int, qpmax: int) -> np.ndarray:
    """Sigma_base::CalcExchangeMatrix -- xtp/src/libxtp/gwbse/sigma_base.cc:36-52:
    Sx[l1,l2] = -sum_{n<occ} sum_P M^P_{l1 n} M^P_{l2 n}."""
    occ = homo - rpamin + 1
    q0 = qpmin - rpamin
    A = M[q0:q0 + qpmax - qpmin + 1, :occ, :].reshape(qpmax - qpmin + 1, -1)
    return -A @ A.T


def ppm_construct(M: np.ndarray, energies: np.ndarray, homo: int, rpamin: int, rpamax: int,
                  screening_r: float = 0.0, screening_i: float = 0.5):
    """PPM::PPM_construct_parameters -- xtp/src/libxtp/gwbse/ppm.cc:31-59 (defaults ppm.h:34).
    Returns (phi, weight, freq)."""
    er = rpa_epsilon(M, energies, homo, rpamin, rpamax, screening_r, "real")
    lam, phi = np.linalg.eigh(er)
    weight = 1.0 - 1.0 / lam
    ei = rpa_epsilon(M, energies, homo, rpamin, rpamax, screening_i, "imag")
    inv = np.linalg.inv(phi.T @ ei @ phi)
    freq = np.empty_like(weight)
    for i in range(len(weight)):
        if weight[i] < 1e-5:
            weight[i] = 0.0
            freq[i] = 0.5
        else:
            nom = inv[i, i] - 1.0
            frac = -1.0 * nom / (nom + weight[i]) * screening_i * screening_i
            freq[i] = np.sqrt(abs(frac))
    return phi, weight, freq


def sigma_ppm_diag(M, energies, weight, freq, homo, rpamin, qpmin, eta, level, omega):
    """Sigma_PPM::CalcCorrelationDiagElement and ...Derivative -- sigma_ppm.cc:37-89.
    M is the PPM-rotated tensor (Mmn * phi, sigma_ppm.cc:32-35). Returns (sigma, dsigma/domega)."""
    lumo = homo + 1
    eta2 = eta * eta
    m = M[level + qpmin - rpamin]            # (nsize x P)
    keep = weight >= 1e-9
    W, F = weight[keep], freq[keep]
    M2 = m[:, keep] ** 2                     # (n x P')
    t = omega - energies[:, None] + np.where(np.arange(m.shape[0])[:, None] < lumo, F[None, :], -F[None, :])
    den = t * t + eta2
    fac = 0.5 * W * F
    s = float(np.sum(fac[None, :] * M2 * t / den))
    ds = float(np.sum(fac[None, :] * (eta2 - t * t) * M2 / (den * den)))
    return s, ds


def sigma_ppm_offdiag(M, energies, weight, freq, homo, rpamin, qpmin, eta, l1, l2, w1, w2):
    """Sigma_PPM::CalcCorrelationOffDiagElement -- sigma_ppm.cc:91-126."""
    lumo = homo + 1
    eta2 = eta * eta
    a = M[l1 + qpmin - rpamin]
    b = M[l2 + qpmin - rpamin]
    keep = weight >= 1e-9
    W, F = weight[keep], freq[keep]
    ab = a[:, keep] * b[:, keep]
    occ = np.arange(a.shape[0])[:, None] < lumo
    temp = energies[:, None] + np.where(occ, -F[None, :], F[None, :])
    t1 = w1 - temp
    t2 = w2 - temp
    fac = 0.25 * W * F
    return float(np.sum(fac[None, :] * (t1 / (t1 * t1 + eta2) + t2 / (t2 * t2 + eta2)) * ab))


def qp_fixed_point(sigma_fn, intercept, omega0, alpha=0.75, tol=1e-5, max_iter=50):
    """GW::SolveQP_FixedPoint (gw.cc:741-757) with NewtonRapson (newton_rapson.h:49-70) on
    f(w) = Sigma_c(w) + intercept - w. Returns the root or None."""
    x = omega0
    for _ in range(max_iter):
        s, ds = sigma_fn(x)
        f = s + intercept - x
        df = ds - 1.0
        if abs(df) < 1e-12:
            return None
        step = -alpha * f / df
        if abs(step) < tol:
            return x
        x += step
    return None


def qp_grid(sigma_fn, intercept, omega0, occupied, half_width=0.75, spacing=0.002, restrict=True,
            zero_margin=1e-6, virtual_min=-0.1, tol=1e-5, min_Z=0.05, max_Z=1.5, bis_max=200):
    """Dense grid QP search (GW::SolveQP_Grid, gw.cc:677-739; candidate acceptance and score
    qp_solver_utils.h:330-402): sign changes of f(w) = Sigma_c(w) + intercept - w on
    [w0-hw, w0+hw] (restricted to w < 0 for occupied / w > virtual_min for virtual levels),
    bisection to tol, Z = 1/(1 - dSigma/dw), accepted for min_Z <= Z <= max_Z, best score
    Z - 0.1|w - w0|. Returns the root or None."""
    left, right = omega0 - half_width, omega0 + half_width
    if restrict:
        if occupied:
            right = min(right, -zero_margin)
        else:
            left = max(left, virtual_min)
    npts = int(np.floor(2.0 * half_width / spacing + 0.5)) + 1
    xs = [omega0 - half_width + spacing * k for k in range(npts)]
    xs = [x for x in xs if left - 1e-14 <= x <= right + 1e-14]
    fs = [sigma_fn(x)[0] + intercept - x for x in xs]
    best, best_score = None, -np.inf
    for k in range(len(xs) - 1):
        f0, f1 = fs[k], fs[k + 1]
        if not ((f0 < 0) != (f1 < 0) or f0 == 0.0):
            continue
        a, b, fa = xs[k], xs[k + 1], f0
        for _ in range(bis_max):
            if b - a <= tol:
                break
            m = 0.5 * (a + b)
            fm = sigma_fn(m)[0] + intercept - m
            if (fm < 0) == (fa < 0):
                a, fa = m, fm
            else:
                b = m
        w = 0.5 * (a + b)
        s, ds = sigma_fn(w)
        deriv = ds - 1.0
        Z = -1.0 / deriv if deriv != 0 else np.inf
        res = s + intercept - w
        if not np.isfinite(Z) or Z <= 0 or Z < min_Z or Z > max_Z or abs(res) > 10 * tol:
            continue
        score = Z - 0.1 * abs(w - omega0)
        if score > best_score:
            best, best_score = w, score
    return best


# ---------------------------------------------------------------- Sigma_c by contour deformation
class SigmaCDAOracle:
    """Sigma_CDA (xtp/src/libxtp/self_energy_evaluators/sigma_cda.cc:30-141) with
    ImaginaryAxisIntegration (xtp/src/libxtp/ImaginaryAxisIntegration.cc:90-176)."""

    def __init__(self, M, energies, homo, rpamin, rpamax, qpmin, eta, alpha, points, weights, symmetry):
        self.M, self.e = M, np.asarray(energies, dtype=np.float64)
        self.homo, self.rpamin, self.rpamax, self.qpmin = homo, rpamin, rpamax, qpmin
        self.eta, self.alpha = eta, alpha
        self.points, self.weights, self.symmetry = np.asarray(points), np.asarray(weights), symmetry
        P = M.shape[2]
        # PrepareScreening: kappa0 = eps_r(0+0i)^-1 - 1  (sigma_cda.cc:30-45)
        self.k0 = np.linalg.inv(rpa_epsilon(M, self.e, homo, rpamin, rpamax, 0.0 + 0.0j, "complex")) - np.eye(P)
        # CalcDielInvVector (ImaginaryAxisIntegration.cc:90-101)
        self.D = []
        for x in self.points:
            einv = np.linalg.inv(rpa_epsilon(M, self.e, homo, rpamin, rpamax, x, "imag")) - np.eye(P)
            self.D.append(-einv + self.k0 * np.exp(-(alpha * x) ** 2))

    def sigma_gq(self, omega, level, rpa_eta=1e-4):
        """ImaginaryAxisIntegration::SigmaGQDiag (ImaginaryAxisIntegration.cc:140-176). The
        broadening is the RPA's fixed eta (rpa.h:90), which Sigma_CDA passes in
        (sigma_cda.cc:125: gq_.SigmaGQDiag(frequency, gw_level, rpa_.getEta())); the GW eta of the
        options does not enter Sigma_c^CDA."""
        occ = self.homo + 1 - self.rpamin
        Imx = self.M[level + self.qpmin - self.rpamin]
        dE = (omega - self.e).astype(np.complex128)
        dE[:occ] += 1j * rpa_eta
        dE[occ:] -= 1j * rpa_eta
        tot = 0.0
        for j, (x, w) in enumerate(zip(self.points, self.weights)):
            if self.symmetry:
                den = 1.0 / (dE + 1j * x) + 1.0 / (dE - 1j * x)
                f = 2.0
            else:
                den = 1.0 / (dE + 1j * x)
                f = 1.0
            S = np.einsum("np,pq,nq->n", Imx, self.D[j], Imx)
            # GaussianQuadratureBase::Integrate multiplies by the symmetry factor only through the
            # integrand; FunctionEvaluation itself returns 0.5/pi Re sum(...)
            tot += w * 0.5 / np.pi * float(np.sum(S * den).real)
        return tot

    @staticmethod
    def prefactor(e_f, e_m, w, tol=1e-10):
        """Sigma_CDA::CalcResiduePrefactor (sigma_cda.cc:62-77)."""
        if e_f < e_m < w:
            return 1.0
        if e_f > e_m > w:
            return -1.0
        if abs(e_m - w) < tol and e_f > e_m:
            return -0.5
        if abs(e_m - w) < tol and e_f < e_m:
            return 0.5
        return 0.0

    def residues(self, omega, level, rpa_eta=1e-4):
        """Sigma_CDA::CalcResidueContribution + tail (sigma_cda.cc:79-141)."""
        homo = self.homo - self.rpamin
        fermi = 0.5 * (self.e[homo + 1] + self.e[homo])
        Imx = self.M[level + self.qpmin - self.rpamin]
        s, tail = 0.0, 0.0
        for i in range(len(self.e)):
            d = self.e[i] - omega
            fac = self.prefactor(fermi, self.e[i], omega)
            if abs(fac) > 1e-10:
                eps = rpa_epsilon(self.M, self.e, self.homo, self.rpamin, self.rpamax, abs(d) + 1j * rpa_eta, "complex")
                x = np.linalg.solve(eps, Imx[i]) - Imx[i]
                s += fac * float(x @ Imx[i])
            if abs(d) > 1e-10:
                from scipy.special import erfc
                ef = 0.5 * np.copysign(1.0, d) * np.exp((self.alpha * d) ** 2) * erfc(abs(self.alpha * d))
                tail += float(Imx[i] @ self.k0 @ Imx[i]) * ef
        return s + tail

    def diag(self, omega, level):
        return self.residues(omega, level) + self.sigma_gq(omega, level)


def evgw_ppm(M, dft, vxc, homo, rpamin, rpamax, qpmin, qpmax, eta, qp_opts, max_iter, sc_limit, mix_order=0,
             mix_alpha=0.7, ScaHFX=0.0):
    """GW::CalculateGWPerturbation with Sigma_PPM (gw.cc:218-310): Sigma_x once; per iteration the
    PPM screening from the current RPA input energies (ppm.cc:31-59) with the tensor rotated
    cumulatively (sigma_ppm.cc:32-35), SolveQP at gw_sc_iteration i, Anderson mixing
    (anderson_mixing.cc) and RPA::UpdateRPAInputEnergies (rpa.cc:32-73) until the RPA energies
    change by at most sc_limit; then Sigma_c at the final frequencies with the last PPM parameters
    and the RPA energies as they stand (Sigma_PPM reads them live, sigma_ppm.cc:52). The
    three-centre rebuild (reset_3c) is left out: every quantity is invariant under the orthogonal
    aux rotations it undoes. Returns (GWA energies, final RPA energies, iterations)."""
    from oracle import qp_solver as Q
    dft = np.asarray(dft, dtype=np.float64)
    sx = (1.0 - ScaHFX) * sigma_x(M, homo, rpamin, qpmin, qpmax)
    e_rpa = dft[rpamin:rpamax + 1].copy()
    freqs = dft[qpmin:qpmax + 1].copy()
    icpt = freqs + np.diag(sx) - np.diag(vxc)
    mix = Q.Anderson(mix_order, mix_alpha)
    Mr = M
    it = 0
    w = f = None
    for i in range(max_iter):
        it = i + 1
        phi, w, f = ppm_construct(Mr, e_rpa, homo, rpamin, rpamax)
        Mr = multiply_right(Mr, phi)
        if mix_order > 0 and i > 0:
            mix.update_input(freqs)
        e_now = e_rpa

        def sig(l, om, d=0):
            return sigma_ppm_diag(Mr, e_now, w, f, homo, rpamin, qpmin, eta, l, om)[d]

        out, _ = Q.solve_qp(sig, lambda l, om: sig(l, om, 1), icpt, freqs, homo, qpmin, qp_opts, gw_sc_iteration=i)
        freqs = np.array(out)
        if max_iter > 1:
            old = e_rpa.copy()
            if mix_order > 0 and i > 0:
                mix.update_output(freqs)
                freqs = mix.mix()
            e_rpa = rpa_update_energies(dft, freqs, qpmin, homo, rpamin, rpamax)
            if np.max(np.abs(e_rpa - old)) <= sc_limit:
                break
    sc = np.array([sigma_ppm_diag(Mr, e_rpa, w, f, homo, rpamin, qpmin, eta, l, freqs[l])[0]
                   for l in range(qpmax - qpmin + 1)])
    return np.diag(sx) + sc - np.diag(vxc) + dft[qpmin:qpmax + 1], e_rpa, it


def screening_at(M, rpa_energies, homo, rpamin, rpamax, omega):
    """BSE::SetupDirectInteractionOperator(energies, omega) (bse.cc:188-204): eps_r(omega) of the
    RPA energies, M rotated to its eigenvectors, eps^-1 = 1/lambda (lambda > 1e-8, else 0)."""
    eps = rpa_epsilon(M, rpa_energies, homo, rpamin, rpamax, omega, "real")
    lam, U = np.linalg.eigh(eps)
    return np.einsum("mnp,pq->mnq", M, U), np.where(lam > 1e-8, 1.0 / lam, 0.0)


def perturbative_dynamical_screening(M, rpa_energies, homo, rpamin, rpamax, Hqp, vmin, cmax, energies, X, Y=None,
                                     max_iter=10, tol=1e-5):
    """BSE::Perturbative_DynamicalScreening (bse.cc:610-700): per excitation, iterate
    Omega_dyn = Omega + <Hd>(W(0)) - <Hd>(W(Omega_dyn)) from Omega_dyn = Omega until the change is
    below `tol` (at most `max_iter` passes). <Hd> is ExpectationValue_Operator(_State)
    (bse.cc:500-533): X^T Hd X, plus Y^T Hd Y + 2 Y^T Hd2 X for the full BSE."""
    def direct(Mr, epsinv, Xs, Ys):
        Hd = bse_dense(0, 0, 1, 0, epsinv, Mr, Hqp, homo, rpamin, vmin, cmax)
        d = np.sum(Xs * (Hd @ Xs), axis=0)
        if Ys is not None:
            Hd2 = bse_dense(0, 0, 0, 1, epsinv, Mr, Hqp, homo, rpamin, vmin, cmax)
            d = d + np.sum(Ys * (Hd @ Ys), axis=0) + 2.0 * np.sum(Ys * (Hd2 @ Xs), axis=0)
        return d

    static = direct(*screening_at(M, rpa_energies, homo, rpamin, rpamax, 0.0), X, Y)
    out = np.array(energies, dtype=float).copy()
    for i in range(len(out)):
        for _ in range(max_iter):
            old = out[i]
            Mr, epsinv = screening_at(M, rpa_energies, homo, rpamin, rpamax, old)
            dyn = direct(Mr, epsinv, X[:, i:i + 1], None if Y is None else Y[:, i:i + 1])[0]
            out[i] = energies[i] + static[i] - dyn
            if abs(out[i] - old) < tol:
                break
    return out
