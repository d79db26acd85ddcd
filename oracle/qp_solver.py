"""Oracle restatement of the reference's QP-equation solver (test infrastructure only).

Solves E = intercept_l + Sigma_c(l, E) per QP level the way GW::SolveQP does
(xtp/src/libxtp/gwbse/gw.cc:323-410): Newton fixed point (gw.cc:741-757 with
xtp/include/votca/xtp/newton_rapson.h:49-70), then the grid search (gw.cc:677-739: a restricted
window first, a full dense window when it finds no accepted root), then linearisation
(gw.cc:412-431). The grid search modes (gw.cc:433-675) and the shared utilities -- the adaptive
outward shell scan with its local 12-step mini-scan, bisection, Brent, root acceptance and score --
follow xtp/include/votca/xtp/qp_solver_utils.h:154-712. Option normalisation from the legacy
qp_grid_steps / qp_grid_spacing pair follows qp_solver_utils.h:115-170.

Sigma enters through two callables, sigma(level, omega) and dsigma(level, omega), as QPFunc
(gw.h:221-293) calls Sigma_base::CalcCorrelationDiagElement(Derivative).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace


class QPError(RuntimeError):
    pass


@dataclass
class QPOptions:
    qp_solver: str = "grid"                 # "grid" or "fixedpoint" (gw.h:69)
    qp_solver_alpha: float = 0.75           # gw.h:70
    g_sc_limit: float = 1e-5
    g_sc_max_iterations: int = 50
    qp_full_window_half_width: float = 0.75
    qp_dense_spacing: float = 0.002
    qp_adaptive_shell_width: float = 0.025
    qp_adaptive_shell_count: int = 0
    qp_restrict_search: bool = True         # gw.h:95-97
    qp_zero_margin: float = 1e-6
    qp_virtual_min_energy: float = -0.1
    qp_root_finder: str = "bisection"       # gw.h:98
    qp_grid_search_mode: str = "adaptive_with_dense_fallback"  # gw.h:99
    min_accepted_Z: float = 0.05            # qp_solver_utils.h:110-111
    max_accepted_Z: float = 1.5


def from_legacy_grid(steps: int, spacing: float, **kw) -> QPOptions:
    """NormalizeGridSearchOptions with only the legacy pair set (qp_solver_utils.h:115-170):
    half width 0.5*spacing*(steps-1), dense spacing = spacing, adaptive shell width =
    full width / (max(21, steps // 4) - 1)."""
    full = spacing * float(steps - 1)
    coarse = max(21, steps // 4)
    return QPOptions(qp_full_window_half_width=0.5 * full, qp_dense_spacing=spacing,
                     qp_adaptive_shell_width=full / float(coarse - 1) if coarse > 1 else 4.0 * spacing, **kw)


class QPFunc:
    """gw.h:221-293: value(w) = Sigma_c(w) + intercept - w, deriv(w) = dSigma_c/dw - 1."""

    def __init__(self, level, sigma, dsigma, intercept):
        self.level, self.sigma, self.dsigma, self.offset = level, sigma, dsigma, intercept
        self.calls = 0

    def value(self, w):
        self.calls += 1
        return self.sigma(self.level, w) + self.offset - w

    def deriv(self, w):
        self.calls += 1
        return self.dsigma(self.level, w) - 1.0


@dataclass
class Root:
    omega: float = 0.0
    residual: float = 0.0
    deriv: float = 0.0
    Z: float = 0.0
    distance_to_ref: float = 0.0
    accepted: bool = False


def shell_width(o: QPOptions) -> float:  # EffectiveAdaptiveShellWidth
    if o.qp_adaptive_shell_count > 0:
        return o.qp_full_window_half_width / float(o.qp_adaptive_shell_count)
    return o.qp_adaptive_shell_width


def bisection(lo, flo, hi, fhi, f: QPFunc, o: QPOptions):  # qp_solver_utils.h:199-226
    if flo * fhi > 0.0:
        raise QPError("Bisection needs a positive and negative function value")
    while True:
        c = 0.5 * (lo + hi)
        if abs(hi - lo) < o.g_sc_limit:
            return c
        yc = f.value(c)
        if abs(yc) < o.g_sc_limit:
            return c
        if yc * flo > 0.0:
            lo, flo = c, yc
        else:
            hi, fhi = c, yc


def brent(lo, flo, hi, fhi, f: QPFunc, o: QPOptions):  # qp_solver_utils.h:228-318
    if flo * fhi > 0.0:
        raise QPError("Brent needs a positive and negative function value")
    a, b, fa, fb = lo, hi, flo, fhi
    c, fc = a, fa
    d = b - a
    e = d
    tol = o.g_sc_limit
    for _ in range(o.g_sc_max_iterations):
        if (fb > 0.0 and fc > 0.0) or (fb < 0.0 and fc < 0.0):
            c, fc = a, fa
            d = b - a
            e = d
        if abs(fc) < abs(fb):
            a, b, c = b, c, b
            fa, fb, fc = fb, fc, fb
        m = 0.5 * (c - b)
        if abs(m) < tol or abs(fb) < o.g_sc_limit:
            return b
        if abs(e) >= tol and abs(fa) > abs(fb):
            s = fb / fa
            if a == c:
                p = 2.0 * m * s
                q = 1.0 - s
            else:
                q1 = fa / fc
                r = fb / fc
                p = s * (2.0 * m * q1 * (q1 - r) - (b - a) * (r - 1.0))
                q = (q1 - 1.0) * (r - 1.0) * (s - 1.0)
            if p > 0.0:
                q = -q
            p = abs(p)
            if q != 0.0 and 2.0 * p < min(3.0 * m * q - abs(tol * q), abs(e * q)):
                e = d
                d = p / q
            else:
                d = m
                e = m
        else:
            d = m
            e = m
        a, fa = b, fb
        b += d if abs(d) > tol else (tol if m > 0.0 else -tol)
        fb = f.value(b)
    raise QPError("Brent did not converge within qp_bisection_max_iter")


def accept_root(r: Root, o: QPOptions) -> bool:  # qp_solver_utils.h:320-344
    if not math.isfinite(r.omega) or not math.isfinite(r.Z):
        return False
    if abs(r.residual) > o.g_sc_limit:
        return False
    return 0.0 < r.Z and o.min_accepted_Z <= r.Z <= o.max_accepted_Z


def score(r: Root) -> float:  # qp_solver_utils.h:346-348
    return r.Z - 0.1 * r.distance_to_ref


def refine_interval(lo, flo, hi, fhi, f: QPFunc, reference, o: QPOptions, use_brent):  # :350-402
    r = Root()
    left0 = abs(flo) <= o.g_sc_limit
    right0 = abs(fhi) <= o.g_sc_limit
    if flo * fhi > 0.0:
        if left0 or right0:
            r.omega = lo if abs(flo) <= abs(fhi) else hi
        else:
            return None
    else:
        r.omega = (brent if use_brent else bisection)(lo, flo, hi, fhi, f, o)
    r.residual = f.value(r.omega)
    r.deriv = f.deriv(r.omega)
    r.Z = -1.0 / r.deriv if abs(r.deriv) > 1e-14 else math.inf
    r.distance_to_ref = abs(r.omega - reference)
    r.accepted = accept_root(r, o)
    return r


def _best(roots):
    best = None
    for r in roots:  # std::max_element: the first of equal scores
        if best is None or score(best) < score(r):
            best = r
    return best


def grid_windowed_adaptive_core(f: QPFunc, w0, left, right, gw_sc_iteration, o: QPOptions, use_brent):
    """qp_solver::SolveQP_Grid_Windowed (qp_solver_utils.h:404-708). Returns (omega or None,
    accepted roots, rejected roots)."""
    acc, rej = [], []
    if left >= right:
        return None, acc, rej
    sw = shell_width(o)
    center = w0
    if gw_sc_iteration == 0:
        f0 = f.value(w0)
        df0 = f.deriv(w0)
        if math.isfinite(f0) and math.isfinite(df0) and abs(df0) > 1e-6:
            wl = w0 - f0 / df0
            if math.isfinite(wl) and left <= wl <= right:
                center = wl
    center = max(left, min(right, center))
    reach = max(center - left, right - center)
    n_shells = int(math.ceil(reach / sw))

    def refine_and_store(a, fa, b, fb):
        if b < a:
            a, b, fa, fb = b, a, fb, fa
        nsub = 12
        brackets = []
        if b > a:
            dx = (b - a) / float(nsub)
            xp, fp = a, fa
            for i in range(1, nsub + 1):
                xc = b if i == nsub else a + float(i) * dx
                fc = fb if i == nsub else f.value(xc)
                if (fp < 0.0 and fc > 0.0) or (fp > 0.0 and fc < 0.0):
                    brackets.append((xp, fp, xc, fc))
                if abs(fp) <= o.g_sc_limit and xp < xc:
                    brackets.append((xp, fp, xc, fc))
                if abs(fc) <= o.g_sc_limit and xp < xc:
                    brackets.append((xp, fp, xc, fc))
                xp, fp = xc, fc
        if brackets:
            best = brackets[0]
            bd = abs(0.5 * (best[0] + best[2]) - center)
            for br in brackets[1:]:
                dist = abs(0.5 * (br[0] + br[2]) - center)
                if dist < bd - 1e-14 or (abs(dist - bd) <= 1e-14 and br[0] < best[0]):
                    best, bd = br, dist
            a, fa, b, fb = best
        r = refine_interval(a, fa, b, fb, f, w0, o, use_brent)
        if r is None:
            return
        (acc if r.accepted else rej).append(r)

    cval = f.value(center)
    lp = (center, cval)
    rp = (center, cval)
    la = ra = True
    for shell in range(1, n_shells + 1):
        added = False
        delta = float(shell) * sw
        if la:
            wl = center - delta
            if wl >= left:
                cur = (wl, f.value(wl))
                added = True
                if lp[1] * cur[1] < 0.0:
                    refine_and_store(cur[0], cur[1], lp[0], lp[1])
                lp = cur
            else:
                la = False
        if ra:
            wr = center + delta
            if wr <= right:
                cur = (wr, f.value(wr))
                added = True
                if rp[1] * cur[1] < 0.0:
                    refine_and_store(rp[0], rp[1], cur[0], cur[1])
                rp = cur
            else:
                ra = False
        if not added and not la and not ra:
            break
    if lp[0] > left + 1e-12:
        end = (left, f.value(left))
        if end[1] * lp[1] < 0.0:
            refine_and_store(end[0], end[1], lp[0], lp[1])
    if rp[0] < right - 1e-12:
        end = (right, f.value(right))
        if rp[1] * end[1] < 0.0:
            refine_and_store(rp[0], rp[1], end[0], end[1])
    if acc:
        return _best(acc).omega, acc, rej
    if rej:
        return _best(rej).omega, acc, rej
    return None, acc, rej


def windowed_adaptive(f, w0, left, right, allow_rejected, gw_sc_iteration, o):  # gw.cc:433-500
    res, acc, rej = grid_windowed_adaptive_core(f, w0, left, right, gw_sc_iteration, o,
                                                o.qp_root_finder == "brent")
    if acc:
        return res
    if rej and not allow_rejected:
        return None
    return res


def windowed_dense(f, w0, left, right, allow_rejected, o):  # gw.cc:502-614
    use_brent = o.qp_root_finder == "brent"
    acc, rej = [], []
    if left < right:
        fp_x, fp_v = left, f.value(left)
        n_steps = max(2, int(math.ceil((right - left) / o.qp_dense_spacing)) + 1)
        for i in range(1, n_steps):
            x = right if i == n_steps - 1 else min(right, left + float(i) * o.qp_dense_spacing)
            v = f.value(x)
            if fp_v * v < 0.0:
                r = refine_interval(fp_x, fp_v, x, v, f, w0, o, use_brent)
                if r is not None:
                    (acc if r.accepted else rej).append(r)
            fp_x, fp_v = x, v
    if acc:
        return _best(acc).omega
    if rej:
        return _best(rej).omega if allow_rejected else None
    return None


def windowed(f, w0, left, right, allow_rejected, gw_sc_iteration, o):  # gw.cc:616-675
    if o.qp_grid_search_mode == "adaptive":
        return windowed_adaptive(f, w0, left, right, allow_rejected, gw_sc_iteration, o)
    if o.qp_grid_search_mode == "dense":
        return windowed_dense(f, w0, left, right, allow_rejected, o)
    if o.qp_grid_search_mode == "adaptive_with_dense_fallback":
        r = windowed_adaptive(f, w0, left, right, allow_rejected, gw_sc_iteration, o)
        if r is not None:
            return r
        return windowed_dense(f, w0, left, right, allow_rejected, o)
    raise QPError(f"Unknown gw.qp_grid_search_mode '{o.qp_grid_search_mode}'")


def solve_grid(f, w0, occupied, gw_sc_iteration, o):  # gw.cc:677-739
    rng = o.qp_full_window_half_width
    fl, fr = w0 - rng, w0 + rng
    rl, rr = fl, fr
    restricted = False
    if o.qp_restrict_search:
        if occupied:
            rr = min(fr, -o.qp_zero_margin)
        else:
            rl = max(fl, o.qp_virtual_min_energy)
        restricted = abs(rl - fl) > 1e-12 or abs(rr - fr) > 1e-12
    if restricted and rl < rr:
        r = windowed(f, w0, rl, rr, False, gw_sc_iteration, o)
        if r is not None:
            return r
        return windowed_dense(f, w0, fl, fr, True, o)
    return windowed(f, w0, fl, fr, True, gw_sc_iteration, o)


def fixed_point(f: QPFunc, w0, o: QPOptions):  # gw.cc:741-757, newton_rapson.h:49-70
    x = w0
    for _ in range(o.g_sc_max_iterations):
        v, dv = f.value(x), f.deriv(x)
        if abs(dv) < 1e-12:
            return None
        step = -o.qp_solver_alpha * v / dv
        if abs(step) < o.g_sc_limit:
            return x
        x += step
    return None


def linearisation(f: QPFunc, w0):  # gw.cc:412-431
    """As the reference computes it: Z = 1 - QPFunc::deriv(w0), where QPFunc::deriv already
    returns dSigma/dw - 1 (gw.h:251-256), so the step divides by 2 - dSigma/dw."""
    s = f.value(w0) + w0 - f.offset  # QPFunc::sigma(w0)
    Z = 1.0 - f.deriv(w0)
    if abs(Z) > 1e-9:
        return w0 + (f.offset - w0 + s) / Z
    return None


def solve_qp(sigma, dsigma, intercepts, frequencies, homo, qpmin, o: QPOptions, gw_sc_iteration=0):
    """GW::SolveQP (gw.cc:323-410): per level fixed point (if asked) -> grid -> linearisation.
    Returns (energies, converged flags)."""
    out, conv = [], []
    for l, (icpt, w0) in enumerate(zip(intercepts, frequencies)):
        f = QPFunc(l, sigma, dsigma, icpt)
        r = fixed_point(f, w0, o) if o.qp_solver == "fixedpoint" else None
        if r is not None:
            out.append(r)
            conv.append(True)
            continue
        r = solve_grid(f, w0, l + qpmin <= homo, gw_sc_iteration, o)
        if r is not None:
            out.append(r)
            conv.append(True)
            continue
        r = linearisation(f, w0)
        out.append(w0 if r is None else r)
        conv.append(False)
    return out, conv


class Anderson:
    """xtp/src/libxtp/anderson_mixing.cc:27-99: input/output histories of order+1 vectors (the
    input history trimmed by the OUTPUT history's length, as the reference does), coefficients
    from the (used x used) normal equations of the residual differences, result
    alpha * mixed output + (1 - alpha) * mixed input."""

    def __init__(self, order, alpha):
        self.order, self.alpha = order + 1, alpha
        self.inp, self.out = [], []

    def update_input(self, v):
        if len(self.out) > self.order - 1:
            self.inp.pop(0)
        self.inp.append([float(x) for x in v])

    def update_output(self, v):
        if len(self.out) > self.order - 1:
            self.out.pop(0)
        self.out.append([float(x) for x in v])

    def mix(self):
        import numpy as np
        it = len(self.out)
        used = it - 1
        om = np.array(self.out[-1])
        im = np.array(self.inp[-1])
        if it > 1 and self.order > 1:
            dn = om - im
            r = [dn - np.array(self.out[used - m]) + np.array(self.inp[used - m]) for m in range(1, it)]
            A = np.array([[float(a @ b) for b in r] for a in r])
            c = np.array([float(a @ dn) for a in r])
            coef = np.linalg.lstsq(A, c, rcond=None)[0]
            for n in range(1, it):
                om = om + coef[n - 1] * (np.array(self.out[used - n]) - np.array(self.out[used]))
                im = im + coef[n - 1] * (np.array(self.inp[used - n]) - np.array(self.inp[used]))
        return self.alpha * om + (1.0 - self.alpha) * im


__all__ = ["Anderson", "QPOptions", "QPError", "from_legacy_grid", "solve_qp", "QPFunc", "refine_interval", "bisection", "brent",
           "windowed_dense", "grid_windowed_adaptive_core", "replace"]
