"""torch.autograd binding for BSE excitation energies (pyxtp style over the C-ABI).

BSEExcitationEnergies.apply(qp_energies, m_ia, problem) returns the lowest `nroots` BSE
excitation energies as a torch tensor; its backward uses the hand-written Hellmann-Feynman
kernels of libxtpgpu (xtpgpu_bse_adjoint) instead of the finite differences the reference's
Python layer uses for excited-state forces (xtp/src/pyxtp/pyxtp/xtp.py:592-682).

Derivatives are at fixed screening: eps0^-1 (and the c c' / v v' blocks used by the direct
term) belong to `problem` and are held constant; m_ia is the symmetric transition block
M^P_{vc} = M^P_{cv}.
"""
from __future__ import annotations

import numpy as np
import torch

from . import pyxtp as px


class BSEProblem:
    """Fixed parts of a BSE solve: the resident tensor, screening and window."""

    def __init__(self, Mmn: px.TCMatrix_gwbse, eps0inv, homo: int, rpamin: int, vmin: int, cmax: int,
                 nroots: int, full: bool = False, tol: float = 1e-9, singlet: bool = True):
        self.Mmn, self.eps0inv = Mmn, np.asarray(eps0inv, dtype=np.float64)
        self.homo, self.rpamin, self.vmin, self.cmax = homo, rpamin, vmin, cmax
        self.vt, self.ct = homo - vmin + 1, cmax - homo
        self.nroots, self.full, self.tol, self.singlet = nroots, full, tol, singlet

    def solve(self, qp_energies: np.ndarray, m_ia: np.ndarray):
        self.Mmn.set_block(self.vmin - self.rpamin, self.homo + 1 - self.rpamin, m_ia, symmetric=True)
        Hqp = np.diag(qp_energies)
        opt = px.BSEOperator_Options(self.homo, self.rpamin, 0, self.vmin, self.cmax)
        cx = 2 if self.singlet else 0
        A = px.BSE_OPERATOR(1, cx, 1, 0, self.eps0inv, self.Mmn, Hqp).configure(opt)
        ds = px.DavidsonSolver()
        ds.set_tolerance(self.tol)
        ds.set_iter_max(200)
        if self.full:
            B = px.BSE_OPERATOR(0, cx, 0, 1, self.eps0inv, self.Mmn, Hqp).configure(opt)
            ev, X, Y = ds.solve_full(A, B, self.nroots)
        else:
            B, Y = None, None
            ev, X = ds.solve(A, self.nroots)
        if not ds.converged():
            raise RuntimeError("BSE Davidson did not converge")
        return ev, A, B, X, Y


class BSEExcitationEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qp_energies: torch.Tensor, m_ia: torch.Tensor, problem: BSEProblem):
        e = qp_energies.detach().to("cpu", torch.float64).numpy()
        m = m_ia.detach().to("cpu", torch.float64).numpy()
        ev, A, B, X, Y = problem.solve(e, m)
        ctx.problem = problem
        ctx.ops = (A, B)
        ctx.vecs = (X, Y)
        ctx.shapes = (qp_energies.device, qp_energies.dtype, m_ia.device, m_ia.dtype)
        return torch.as_tensor(ev, dtype=qp_energies.dtype, device=qp_energies.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        A, B = ctx.ops
        X, Y = ctx.vecs
        dev_e, dt_e, dev_m, dt_m = ctx.shapes
        need_e, need_m = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        d_eps, d_m = px.bse_adjoint(A, X, B, Y, want_m=need_m)
        g = grad_out.detach().to("cpu", torch.float64).numpy()
        ge = torch.as_tensor(d_eps @ g, dtype=dt_e, device=dev_e) if need_e else None
        gm = torch.as_tensor(np.tensordot(g, d_m, axes=(0, 0)), dtype=dt_m, device=dev_m) if need_m else None
        return ge, gm, None


def excitation_energies(qp_energies: torch.Tensor, m_ia: torch.Tensor, problem: BSEProblem) -> torch.Tensor:
    return BSEExcitationEnergies.apply(qp_energies, m_ia, problem)
