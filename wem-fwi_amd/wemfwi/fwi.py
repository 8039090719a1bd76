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

Background wavefields are kept on the device for all rows of a call rather
than checkpointed: HBM holds the Down/Up fields of the whole survey for the
model sizes of this path, so recomputation would only cost time.
"""

from __future__ import annotations

from typing import Mapping, Optional

import torch

from .api import Born, Hypercube, SuperVector, Vector, WEM, cdot


class WEMMisfit:
    """Objective J(S, rho) for a fixed observed data vector; call it on
    [nz][nx] complex64 CUDA tensors to get a differentiable scalar."""

    def __init__(self, obs: Vector, wave: Vector, par: Mapping, hyper: Hypercube, mode: Optional[str] = "real",
                 adjoint_mode: str = "exact"):
        self.obs = obs
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
            self._born = Born(m, self.obs, self.wave, self.par, mode=None, adjoint_mode=self.adjoint_mode)
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
