"""Solvers of the FWI drivers (scripts/fwi.py:46-131, efwi.py:208-256) over
the library's objectives, and shot sharding across processes.

The reference drives its problems with python-solver (branch rustam-dev of
ettore88/python-solver, docker/Dockerfile:26; not part of the reference
tree): LBFGSsolver / NLCGsolver on ProblemL2NonLinear(Reg) with optional
bounds, and an LCGsolver fit of the preconditioned model.  Restated here from
the published algorithms:

  LBFGS   two-loop recursion with m_steps pairs (Nocedal, 1980), strong-Wolfe
          line search; the first trial step is step0 |x| / |g| along -g (a
          fraction of the model's size, so that models in any unit -- slowness
          squared in s^2/m^2 is ~1e-7 -- start sensibly; 1/|g| if x = 0)
  NLCG    beta = FR, PRP, HS or HZ (Hager-Zhang), restart when beta < 0 or
          the direction is not a descent direction, strong-Wolfe line search
  line    bracketing + zoom with cubic interpolation under the strong Wolfe
  search  conditions c1 = 1e-4, c2 = 0.9 (L-BFGS) / 0.1 (NLCG)
          (Nocedal & Wright, Algorithms 3.5-3.6)
  CGLS    conjugate gradients on the normal equations of 1/2 |L x - y|^2
  bounds  the real part of the model clipped to [min, max] after each step
          (ProblemL2NonLinear's minBound / maxBound)

Models are lists of tensors; the inner product is Re sum a conj(b) in double
precision (complexHyper::dot).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import torch

Vec = List[torch.Tensor]


def dot(a: Vec, b: Vec) -> float:
    s = 0.0
    for x, y in zip(a, b):
        if x.is_complex():
            s += float((x.to(torch.complex128) * y.to(torch.complex128).conj()).real.sum())
        else:
            s += float((x.double() * y.double()).sum())
    return s


def axpy(a: float, x: Vec, y: Vec) -> Vec:
    return [yy + a * xx for xx, yy in zip(x, y)]


def scale(a: float, x: Vec) -> Vec:
    return [a * xx for xx in x]


@dataclass
class History:
    objective: List[float] = field(default_factory=list)
    gradient_norm: List[float] = field(default_factory=list)
    steps: List[float] = field(default_factory=list)
    evaluations: int = 0


class Problem:
    """f(x) -> (value, gradient); bounds clip the real part."""

    def __init__(self, fg: Callable[[Vec], Tuple[float, Vec]], min_bound: Optional[float] = None,
                 max_bound: Optional[float] = None):
        self.fg = fg
        self.min_bound, self.max_bound = min_bound, max_bound

    def project(self, x: Vec) -> Vec:
        if self.min_bound is None and self.max_bound is None:
            return x
        out = []
        for t in x:
            if t.is_complex():
                re = t.real.clamp(self.min_bound, self.max_bound)
                out.append(torch.complex(re, t.imag))
            else:
                out.append(t.clamp(self.min_bound, self.max_bound))
        return out


def _cubic_min(a, fa, ga, b, fb, gb):
    """Minimiser of the cubic interpolating (a, fa, ga), (b, fb, gb), or the
    bisection if it is not inside the interval."""
    d1 = ga + gb - 3 * (fa - fb) / (a - b)
    disc = d1 * d1 - ga * gb
    if disc >= 0:
        d2 = math.copysign(math.sqrt(disc), b - a)
        den = gb - ga + 2 * d2
        if den != 0:
            x = b - (b - a) * (gb + d2 - d1) / den
            lo, hi = min(a, b), max(a, b)
            if lo + 0.05 * (hi - lo) < x < hi - 0.05 * (hi - lo):
                return x
    return 0.5 * (a + b)


def line_search(problem: Problem, x: Vec, f0: float, g0: Vec, p: Vec, alpha0: float, c1: float, c2: float,
                hist: History, max_iter: int = 20):
    """Strong-Wolfe line search along p from x; returns (alpha, x1, f1, g1)."""
    dphi0 = dot(g0, p)
    if dphi0 >= 0:
        raise ValueError("line search: not a descent direction")

    def phi(a):
        xa = problem.project(axpy(a, p, x))
        fa, ga = problem.fg(xa)
        hist.evaluations += 1
        return xa, fa, ga, dot(ga, p)

    a_prev, f_prev, d_prev = 0.0, f0, dphi0
    a = alpha0
    best = None
    for it in range(max_iter):
        xa, fa, ga, da = phi(a)
        if fa > f0 + c1 * a * dphi0 or (it > 0 and fa >= f_prev):
            return _zoom(problem, x, f0, dphi0, p, a_prev, f_prev, d_prev, a, fa, da, c1, c2, hist,
                         best=(a_prev, None, f_prev, None) if best is None else best, phi=phi)
        if abs(da) <= -c2 * dphi0:
            return a, xa, fa, ga
        best = (a, xa, fa, ga)
        if da >= 0:
            return _zoom(problem, x, f0, dphi0, p, a, fa, da, a_prev, f_prev, d_prev, c1, c2, hist, best=best,
                         phi=phi)
        a_prev, f_prev, d_prev = a, fa, da
        a *= 2.0
    return best[0], best[1], best[2], best[3]


def _zoom(problem, x, f0, dphi0, p, alo, flo, dlo, ahi, fhi, dhi, c1, c2, hist, best, phi, max_iter=40):
    for _ in range(max_iter):
        a = _cubic_min(alo, flo, dlo, ahi, fhi, dhi)
        xa, fa, ga, da = phi(a)
        if fa > f0 + c1 * a * dphi0 or fa >= flo:
            ahi, fhi, dhi = a, fa, da
        else:
            if abs(da) <= -c2 * dphi0:
                return a, xa, fa, ga
            if da * (ahi - alo) >= 0:
                ahi, fhi, dhi = alo, flo, dlo
            alo, flo, dlo = a, fa, da
            best = (a, xa, fa, ga)
        if abs(ahi - alo) < 1e-12 * max(1.0, abs(alo)):
            break
    if best[1] is None:          # no point with sufficient decrease was found
        xa, fa, ga, _ = phi(alo) if alo > 0 else (None, None, None, None)
        if xa is None:
            raise RuntimeError("line search failed")
        return alo, xa, fa, ga
    return best


def _first_step(x: Vec, g: Vec, step0: float) -> float:
    gn = math.sqrt(dot(g, g))
    xn = math.sqrt(dot(x, x))
    return step0 * xn / gn if xn > 0 else 1.0 / max(gn, 1e-300)


class LBFGS:
    def __init__(self, niter: int, m_steps: int = 10, c1: float = 1e-4, c2: float = 0.9, step0: float = 0.05):
        self.niter, self.m, self.c1, self.c2, self.step0 = niter, m_steps, c1, c2, step0

    def run(self, problem: Problem, x0: Vec) -> Tuple[Vec, History]:
        hist = History()
        x = problem.project([t.clone() for t in x0])
        f, g = problem.fg(x)
        hist.evaluations += 1
        hist.objective.append(f)
        hist.gradient_norm.append(math.sqrt(dot(g, g)))
        S, Y, R = [], [], []
        for k in range(self.niter):
            if hist.gradient_norm[-1] <= 1e-300:
                break
            q = [t.clone() for t in g]
            al = []
            for s, y, r in reversed(list(zip(S, Y, R))):
                a = r * dot(s, q)
                al.append(a)
                q = axpy(-a, y, q)
            gamma = dot(S[-1], Y[-1]) / dot(Y[-1], Y[-1]) if S else _first_step(x, g, self.step0)
            z = scale(gamma, q)
            for (s, y, r), a in zip(zip(S, Y, R), reversed(al)):
                b = r * dot(y, z)
                z = axpy(a - b, s, z)
            p = scale(-1.0, z)
            if dot(g, p) >= 0:              # not a descent direction: restart
                S, Y, R = [], [], []
                p = scale(-_first_step(x, g, self.step0), g)
            try:
                a, x1, f1, g1 = line_search(problem, x, f, g, p, 1.0, self.c1, self.c2, hist)
            except (ValueError, RuntimeError):
                break                       # no decrease left along any direction
            if f1 >= f and dot(g1, g1) >= dot(g, g):
                break
            s = axpy(-1.0, x, x1)
            y = axpy(-1.0, g, g1)
            sy = dot(s, y)
            if sy > 1e-12 * math.sqrt(dot(s, s) * dot(y, y)):
                S.append(s)
                Y.append(y)
                R.append(1.0 / sy)
                if len(S) > self.m:
                    S.pop(0), Y.pop(0), R.pop(0)
            x, f, g = x1, f1, g1
            hist.objective.append(f)
            hist.gradient_norm.append(math.sqrt(dot(g, g)))
            hist.steps.append(a)
        return x, hist


class NLCG:
    def __init__(self, niter: int, beta_type: str = "FR", c1: float = 1e-4, c2: float = 0.1, step0: float = 0.05):
        if beta_type not in ("FR", "PRP", "HS", "HZ"):
            raise ValueError("beta_type must be FR, PRP, HS or HZ")
        self.niter, self.beta_type, self.c1, self.c2, self.step0 = niter, beta_type, c1, c2, step0

    def _beta(self, g1: Vec, g0: Vec, p: Vec) -> float:
        y = axpy(-1.0, g0, g1)
        if self.beta_type == "FR":
            return dot(g1, g1) / dot(g0, g0)
        if self.beta_type == "PRP":
            return dot(g1, y) / dot(g0, g0)
        if self.beta_type == "HS":
            return dot(g1, y) / dot(p, y)
        py = dot(p, y)
        t = axpy(-2.0 * dot(y, y) / py, p, y)
        return dot(t, g1) / py

    def run(self, problem: Problem, x0: Vec) -> Tuple[Vec, History]:
        hist = History()
        x = problem.project([t.clone() for t in x0])
        f, g = problem.fg(x)
        hist.evaluations += 1
        hist.objective.append(f)
        hist.gradient_norm.append(math.sqrt(dot(g, g)))
        p = scale(-1.0, g)
        alpha = _first_step(x, g, self.step0)
        for k in range(self.niter):
            if hist.gradient_norm[-1] <= 1e-300:
                break
            try:
                a, x1, f1, g1 = line_search(problem, x, f, g, p, alpha, self.c1, self.c2, hist)
            except (ValueError, RuntimeError):
                break
            beta = max(0.0, self._beta(g1, g, p))
            pn = axpy(beta, p, scale(-1.0, g1))
            if dot(pn, g1) >= 0:
                pn = scale(-1.0, g1)
            alpha = a * max(dot(g, p) / min(dot(g1, pn), -1e-300), 1e-3) if dot(g1, pn) < 0 else a
            x, f, g, p = x1, f1, g1, pn
            hist.objective.append(f)
            hist.gradient_norm.append(math.sqrt(dot(g, g)))
            hist.steps.append(a)
        return x, hist


def cgls(forward: Callable[[Vec], Vec], adjoint: Callable[[Vec], Vec], x0: Vec, y: Vec, niter: int) -> Vec:
    """CG on the normal equations of 1/2 |L x - y|^2 from x0 (LCGsolver on
    ProblemL2Linear, fwi.py:58-68)."""
    x = [t.clone() for t in x0]
    r = axpy(-1.0, forward(x), y)
    s = adjoint(r)
    p = [t.clone() for t in s]
    gam = dot(s, s)
    for _ in range(niter):
        if gam == 0:
            break
        q = forward(p)
        a = gam / dot(q, q)
        x = axpy(a, p, x)
        r = axpy(-a, q, r)
        s = adjoint(r)
        g1 = dot(s, s)
        p = axpy(g1 / gam, p, s)
        gam = g1
    return x


# ------------------------------------------------------------------ sharding

def shot_range(ns: int, rank: int, world: int) -> Tuple[int, int]:
    """Contiguous shots [s0, s1) of a rank, sizes differing by at most one."""
    base, extra = divmod(ns, world)
    s0 = rank * base + min(rank, extra)
    return s0, s0 + base + (1 if rank < extra else 0)


def allreduce_value_and_grad(f: float, g: Vec) -> Tuple[float, Vec]:
    """Sum of every rank's objective and gradient (the one exchange of the
    shot-sharded problem), in double precision (tensors are summed as
    complex128 / float64 and rounded back once).  gloo reduces host copies; nccl (RCCL)
    reduces the device tensors.  Without an initialised process group it is
    the identity."""
    import torch.distributed as dist
    if not (dist.is_available() and dist.is_initialized()) or dist.get_world_size() == 1:
        return f, g
    host = dist.get_backend() == "gloo"
    out = []
    for t in g:
        w = t.detach().to("cpu", copy=True) if host else t.detach().clone()
        if w.is_complex():
            w = torch.view_as_real(w.to(torch.complex128)).contiguous()
            dist.all_reduce(w)
            w = torch.view_as_complex(w).to(t.dtype)
        else:
            w = w.to(torch.float64)
            dist.all_reduce(w)
            w = w.to(t.dtype)
        out.append(w.to(t.device) if host else w)
    fv = torch.tensor([float(f)], dtype=torch.float64, device="cpu" if host else g[0].device)
    dist.all_reduce(fv)
    return float(fv.item()), out
