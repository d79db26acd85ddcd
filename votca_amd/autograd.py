"""torch.autograd binding for BSE excitation energies (pyxtp style over the C-ABI).

BSEExcitationEnergies.apply(qp_energies, m_ia, problem) returns the lowest `nroots` BSE
excitation energies as a torch tensor; its backward uses the hand-written Hellmann-Feynman
kernels of libxtpgpu (xtpgpu_bse_adjoint_weighted_device) instead of the finite differences the
reference's Python layer uses for excited-state forces (xtp/src/pyxtp/pyxtp/xtp.py:592-682).

Derivatives are at fixed screening: eps0^-1 (and the c c' / v v' blocks used by the direct
term) belong to `problem` and are held constant; m_ia is the symmetric transition block
M^P_{vc} = M^P_{cv}, shape (aux, vtotal, ctotal).

The operators are packed once per problem. A forward with new QP energies only replaces the
operators' Hqp (xtpgpu_bse_set_hqp); a forward with a new m_ia writes it into the resident tensor
and repacks the transition-block layouts (xtpgpu_bse_refresh) — the (c c') block, the largest,
is never repacked. The backward contracts the incoming gradient into one aux x vtotal x ctotal
derivative on the device; when m_ia lives on the GPU, its gradient is that device array (no host
round trip).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import pyxtp as px


class BSEProblem:
    """Fixed parts of a BSE solve: the resident tensor, screening and window; the packed operators."""

    def __init__(self, Mmn: px.TCMatrix_gwbse, eps0inv, homo: int, rpamin: int, vmin: int, cmax: int,
                 nroots: int, full: bool = False, tol: float = 1e-9, singlet: bool = True):
        self.Mmn, self.eps0inv = Mmn, np.asarray(eps0inv, dtype=np.float64)
        self.homo, self.rpamin, self.vmin, self.cmax = homo, rpamin, vmin, cmax
        self.vt, self.ct = homo - vmin + 1, cmax - homo
        self.nroots, self.full, self.tol, self.singlet = nroots, full, tol, singlet
        self.A = self.B = None
        self._m_key = None
        self.packs = 0      # full operator packings
        self.refreshes = 0  # transition-block repacks

    def _operators(self, Hqp):
        if self.A is None:
            opt = px.BSEOperator_Options(self.homo, self.rpamin, 0, self.vmin, self.cmax)
            cx = 2 if self.singlet else 0
            self.A = px.BSE_OPERATOR(1, cx, 1, 0, self.eps0inv, self.Mmn, Hqp).configure(opt)
            if self.full:
                self.B = px.BSE_OPERATOR(0, cx, 0, 1, self.eps0inv, self.Mmn, Hqp).configure(opt)
            self.packs += 1
        else:
            self.A.set_hqp(Hqp)  # B (cqp = 0) has no Hqp part
        return self.A, self.B

    def _set_transitions(self, m_ia, key):
        if key is not None and key == self._m_key:
            return False
        m = m_ia.detach().to("cpu", torch.float64).numpy() if isinstance(m_ia, torch.Tensor) else m_ia
        self.Mmn.set_block(self.vmin - self.rpamin, self.homo + 1 - self.rpamin, m, symmetric=True)
        self._m_key = key
        # holding the tensor keeps its storage alive, so no later tensor can reuse its address (and
        # key) while the tensor's values are the ones packed; writes through any view of it bump the
        # shared version counter
        self._m_hold = m_ia if key is not None else None
        return True

    def ensure_transitions(self, m_ia, m_key):
        """Make the packed transition block the given one again (a backward after another forward)."""
        if self._set_transitions(m_ia, m_key) and self.A is not None:
            self.A.refresh_transitions()
            if self.B is not None:
                self.B.refresh_transitions()
            self.refreshes += 1

    def solve(self, qp_energies: np.ndarray, m_ia, m_key=None):
        changed = self._set_transitions(m_ia, m_key)
        fresh = self.A is None
        A, B = self._operators(np.diag(qp_energies))
        if changed and not fresh:
            A.refresh_transitions()
            if B is not None:
                B.refresh_transitions()
            self.refreshes += 1
        ds = px.DavidsonSolver()
        ds.set_tolerance(self.tol)
        ds.set_iter_max(200)
        if self.full:
            ev, X, Y = ds.solve_full(A, B, self.nroots)
        else:
            Y = None
            ev, X = ds.solve(A, self.nroots)
        if not ds.converged():
            raise RuntimeError("BSE Davidson did not converge")
        return ev, A, B, X, Y


def _key(t: torch.Tensor):
    """Identity of a tensor's current contents: storage, offset, version counter."""
    return (t.data_ptr(), t.storage_offset(), t._version, tuple(t.shape), str(t.device))


class BSEExcitationEnergies(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qp_energies: torch.Tensor, m_ia: torch.Tensor, problem: BSEProblem):
        e = qp_energies.detach().to("cpu", torch.float64).numpy()
        key = _key(m_ia)
        ev, A, B, X, Y = problem.solve(e, m_ia, key)
        ctx.problem = problem
        ctx.m_ia, ctx.m_key = m_ia.detach(), key
        ctx.ops = (A, B)
        ctx.vecs = (X, Y)
        ctx.shapes = (qp_energies.device, qp_energies.dtype, m_ia.device, m_ia.dtype, tuple(m_ia.shape))
        return torch.as_tensor(ev, dtype=qp_energies.dtype, device=qp_energies.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        A, B = ctx.ops
        X, Y = ctx.vecs
        # dOmega/dM reads the packed transition block: the one of this forward
        ctx.problem.ensure_transitions(ctx.m_ia, ctx.m_key)
        dev_e, dt_e, dev_m, dt_m, shape_m = ctx.shapes
        need_e, need_m = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        g = np.ascontiguousarray(grad_out.detach().to("cpu", torch.float64).numpy())
        Xf = px._f64(X)
        Yf = px._f64(Y) if Y is not None else None
        d_eps = np.zeros(A.vt + A.ct) if need_e else None
        gm = None
        if need_m:
            # the weighted derivative lands in a device array; torch owns it when a GPU is there
            on_gpu = torch.cuda.is_available()
            if on_gpu:
                buf = torch.empty(shape_m, dtype=torch.float64, device="cuda")
                torch.cuda.synchronize()
                ptr = C.c_void_p(buf.data_ptr())
            else:
                dbuf = px.DeviceBuffer(A.Mmn.ctx, int(np.prod(shape_m)) * 8)
                ptr = dbuf.ptr
        else:
            ptr = None
        px._check(px.lib().xtpgpu_bse_adjoint_weighted_device(
            A.h, B.h if B is not None else None, Xf.shape[1], px._dptr(Xf),
            px._dptr(Yf) if Yf is not None else None, Xf.shape[0], px._dptr(g),
            px._dptr(d_eps) if need_e else None, ptr))
        if need_m:
            if on_gpu:
                gm = buf.to(device=dev_m, dtype=dt_m)
            else:
                gm = torch.as_tensor(dbuf.download(shape_m, order="C"), dtype=dt_m, device=dev_m)
        ge = torch.as_tensor(d_eps, dtype=dt_e, device=dev_e) if need_e else None
        return ge, gm, None


def excitation_energies(qp_energies: torch.Tensor, m_ia: torch.Tensor, problem: BSEProblem) -> torch.Tensor:
    return BSEExcitationEnergies.apply(qp_energies, m_ia, problem)
