"""The L2 data-misfit objective of the reference's FWI scripts as a
torch.autograd.Function (scripts/fwi.py:70-131 with WEM.py:52-81):

    J(S, rho) = 1/2 || WEM([S, rho]) - d ||^2
    grad J    = Born([S, rho])^H (WEM([S, rho]) - d)

computed by the library's WEM and Born operators.  The misfit is accumulated
in double precision.  Data may be complex frequency data [ns][nr][nbins] (J
over the survey bins) or float traces [ns][nr][nt] (J in time, the residual
transformed with the unitary FFT for the adjoint, as BornFull::adjoint does).
mode='real' zeroes the imaginary part of the gradient (WEM.py:59-67), and the
exact adjoint of the scatterer is used by default so that the gradient is the
adjoint of the linearised operator.

checkpoint=k keeps the background and receiver wavefields only every k
depth steps during the gradient and recomputes them segment by segment: the
gradient is bitwise the same, the device memory per (shot, frequency) row
drops from 4 nz nx samples to about (4 k + 2 nz / k) nx, so larger batches
of rows fit one launch.

fwi_problem / fit_preconditioned / shard_par assemble the driver of
scripts/fwi.py:46-131 around it (solvers in wemfwi.solver, interpolation in
wemfwi.interp).
"""

from __future__ import annotations

from typing import Mapping, Optional

import torch

from .api import Born, Hypercube, SuperVector, Vector, WEM, cdot


class WEMMisfit:
    """Objective J(S, rho) for a fixed observed data vector; call it on
    [nz][nx] complex64 CUDA tensors to get a differentiable scalar."""

    def __init__(self, obs: Vector, wave: Vector, par: Mapping, hyper: Hypercube, mode: Optional[str] = "real",
                 adjoint_mode: str = "exact", checkpoint: int = 0):
        self.obs = obs
        self.checkpoint = int(checkpoint)
        self.wave = wave
        self.par = dict(par)
        self.hyper = hyper
        self.mode = mode
        self.adjoint_mode = adjoint_mode
        self._wem: Optional[WEM] = None
        self._born: Optional[Born] = None

    def _model(self, slow: torch.Tensor, den: torch.Tensor) -> SuperVector:
        return SuperVector(Vector(self.hyper, arr=slow.detach().contiguous()),
                           Vector(self.hyper, arr=den.detach().contiguous()))

    def residual(self, slow: torch.Tensor, den: torch.Tensor) -> Vector:
        m = self._model(slow, den)
        if self._wem is None:
            self._wem = WEM(m, self.obs, self.wave, self.par)
        pred = self.obs.clone().zero()
        self._wem.forward(False, m, pred)
        pred.arr.sub_(self.obs.arr)
        return pred

    def gradient(self, slow: torch.Tensor, den: torch.Tensor, res: Vector):
        m = self._model(slow, den)
        if self._born is None:
            self._born = Born(m, self.obs, self.wave, self.par, mode=None, adjoint_mode=self.adjoint_mode,
                              checkpoint=self.checkpoint)
        else:
            self._born.set_background(m)
        g = m.clone().zero()
        self._born.adjoint(False, g, res)
        if self.mode == "real":
            for v in g.vecs:
                v.arr.imag.zero_()
        return g[0].arr, g[1].arr

    def __call__(self, slow: torch.Tensor, den: torch.Tensor) -> torch.Tensor:
        return _Misfit.apply(slow, den, self)


class _Misfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, slow, den, obj: WEMMisfit):
        res = obj.residual(slow, den)
        ctx.obj = obj
        ctx.res = res
        ctx.save_for_backward(slow, den)
        return torch.tensor(0.5 * cdot(res.arr, res.arr), dtype=torch.float64, device=slow.device)

    @staticmethod
    def backward(ctx, grad_out):
        slow, den = ctx.saved_tensors
        gS, gR = ctx.obj.gradient(slow, den, ctx.res)
        s = grad_out.to(torch.float32)
        return gS * s, gR * s, None


# ------------------------------------------------------------------ driver

def shard_par(par: Mapping, s0: int, s1: int) -> dict:
    """The parameters of shots [s0, s1) of a regular source line
    (n/o/d s keys of Device::createRegGeom): one process's share of the shots."""
    p = dict(par)
    if str(p.get("coord", "") or ""):
        raise ValueError("shard_par: irregular (JSON) geometry is sharded by its own coordinate lists")
    p["ns"] = s1 - s0
    p["osx"] = float(p.get("osx", 0.0)) + s0 * float(p.get("dsx", 0.0))
    p["osz"] = float(p.get("osz", 0.0)) + s0 * float(p.get("dsz", 0.0))
    return p


def fwi_problem(misfit: WEMMisfit, den: torch.Tensor, precond=None, coarse_shape=None, vmin: Optional[float] = None,
                vmax: Optional[float] = None, mask: Optional[torch.Tensor] = None):
    """ProblemL2NonLinear of scripts/fwi.py:97-110 on the slowness: the model
    is the slowness itself or, with a preconditioner P (an interpolation
    operator from a coarse grid), its coarse coefficients m = P m_pre
    (fwi.py:46-68, 95-96); bounds 1/vmax^2 <= Re m <= 1/vmin^2 on the solver's
    variable (fwi.py:98-104); mask multiplies the gradient (the Born input,
    fwi.py:73-76).  The objective and gradient are summed over the processes of
    an initialised torch.distributed group, each holding its shots."""
    from . import solver

    def fg(x):
        m = x[0]
        if precond is not None:
            full = torch.zeros_like(den)
            precond.forward(False, m, full)
            m = full
        res = misfit.residual(m, den)
        J = 0.5 * float((res.arr.to(torch.complex128).abs() ** 2).sum())
        gS, _ = misfit.gradient(m, den, res)
        if mask is not None:
            gS = gS * mask
        if precond is not None:
            gp = torch.zeros(coarse_shape, dtype=torch.complex64, device=gS.device)
            precond.adjoint(False, gp, gS.contiguous())
            gS = gp
        return solver.allreduce_value_and_grad(J, [gS])

    lo = None if vmax is None else 1.0 / vmax ** 2
    hi = None if vmin is None else 1.0 / vmin ** 2
    return solver.Problem(fg, lo, hi)


def fit_preconditioned(precond, model: torch.Tensor, coarse_shape, niter: int) -> torch.Tensor:
    """The coarse coefficients that best reproduce a model, by CG on
    1/2 |P m_pre - m|^2 (the LCGsolver fit of fwi.py:58-68)."""
    from . import solver

    def fwd(v):
        out = torch.zeros_like(model)
        precond.forward(False, v[0], out)
        return [out]

    def adj(r):
        out = torch.zeros(coarse_shape, dtype=torch.complex64, device=model.device)
        precond.adjoint(False, out, r[0].contiguous())
        return [out]

    x0 = [torch.zeros(coarse_shape, dtype=torch.complex64, device=model.device)]
    return solver.cgls(fwd, adj, x0, [model], niter)[0]
