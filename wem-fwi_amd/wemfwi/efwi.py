"""The frequency-extended FWI path of scripts/efwi.py on the library.

Operators of operator/src/python/Operator.py used by that script, on device
SuperVectors of [nw][nz][nx] models (slowness, density):

  HyperbolicPenalty / Softclip   Operator.py:619-671  (slowness: Im -> h(Im))
  Softmax, Softmin, dSoftmax,    Operator.py:674-775  (slowness: Re soft-clipped
  dSoftmin, SoftMinMax                                 to [xmin, xmax])
  Derivative                     Operator.py:228-292  (DSO filter along frequency)
  Mask                           Operator.py:533-546

and the problem efwi.py builds (efwi.py:154-230): with P = HyperbolicPenalty
after the optional SoftMinMax and D the Derivative,

  J(m) = 1/2 || extWEM(P(m)) - d ||^2 + 1/2 || eps D m ||^2
  grad = P'(m)^H Mask extBorn(P(m))^H (extWEM(P(m)) - d) + eps^2 D^H D m

(python-solver's ProblemL2NonLinearReg, branch rustam-dev of
ettore88/python-solver, docker/Dockerfile:26 -- not part of the reference
tree; its residual vector is [f(m) - d, eps g(m)] and its objective half the
squared norm).  The exact adjoint of the scatterer is used by default, so the
gradient is the adjoint of the linearised operator.

The soft clips and the frequency filter run as library kernels
(wemfwi_soft_clip, wemfwi_freq_filter); the filter matrix, including scipy's
boundary modes, is built by wemfwi_filter_matrix.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from ._lib import check, lib
from .api import Hypercube, Operator, SuperVector, Vector, cdot, extBorn, extWEM
from .core import _stream_ptr

_MODES = {"reflect": 0, "constant": 1, "nearest": 2, "mirror": 3, "wrap": 4}


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _vecs(v):
    return v.vecs if isinstance(v, SuperVector) else list(v)


def _soft_clip(inp: torch.Tensor, out: torch.Tensor, comp: int, sign: int, x0: float, l: float, tau: float,
               bg: Optional[torch.Tensor] = None, add: bool = False):
    for t in (inp, out) + ((bg,) if bg is not None else ()):
        if t.dtype != torch.complex64 or not t.is_cuda or not t.is_contiguous():
            raise TypeError("soft clip operands must be contiguous complex64 CUDA tensors")
    if inp.shape != out.shape or (bg is not None and bg.shape != inp.shape):
        raise ValueError("soft clip: shapes differ")
    check(lib.wemfwi_soft_clip(inp.numel(), comp, sign, float(x0), float(l), float(tau), int(bg is not None),
                               None if bg is None else _ptr(bg), _ptr(inp), _ptr(out), int(add),
                               _stream_ptr(None)))


def _pass(add, m: torch.Tensor, d: torch.Tensor):
    if add:
        d.add_(m)
    else:
        d.copy_(m)


# ------------------------------------------------------------------ penalties

class HyperbolicPenalty(Operator):
    """Im S -> h(Im S) with h(x) = l x - sqrt(l^2 x^2 + tau^2), l = l_in/2;
    Re S and the density pass through (Operator.py:625-643)."""

    def __init__(self, model, data, l=1, tau=.001):
        super().__init__(model, data)
        self.l = l / 2
        self.tau = tau

    def forward(self, add, model, data):
        m, d = _vecs(model), _vecs(data)
        _soft_clip(m[0].arr, d[0].arr, 1, 1, 0.0, self.l, self.tau, add=add)
        _pass(add, m[1].arr, d[1].arr)


class _Jacobian(Operator):
    """Self-adjoint diagonal Jacobian of a soft clip at a background."""

    comp = 1
    sign = 1

    def __init__(self, model, data, l, tau, x0=0.0):
        super().__init__(model, data)
        self.l = l / 2
        self.tau = tau
        self.x0 = x0
        self.bg = _vecs(model)[0].arr.detach().clone()

    def set_background(self, bg):
        self.bg.copy_(_vecs(bg)[0].arr)

    setBg = set_background  # noqa: N815 - the reference's test calls setBg

    def forward(self, add, model, data):
        m, d = _vecs(model), _vecs(data)
        _soft_clip(m[0].arr, d[0].arr, self.comp, self.sign, self.x0, self.l, self.tau, bg=self.bg, add=add)
        _pass(add, m[1].arr, d[1].arr)

    def adjoint(self, add, model, data):
        self.forward(add, data, model)


class Softclip(_Jacobian):
    """Jacobian of HyperbolicPenalty: Im scaled by h'(Im bg) (Operator.py:645-671)."""

    def __init__(self, model, data, l=1, tau=.001):
        super().__init__(model, data, l, tau)


class Softmax(Operator):
    """Re S -> h(Re S - xmax) + xmax (Operator.py:674-691)."""

    def __init__(self, model, data, xmax, l=1, tau=.001):
        super().__init__(model, data)
        self.l, self.tau, self.xmax = l / 2, tau, xmax

    def forward(self, add, model, data):
        m, d = _vecs(model), _vecs(data)
        _soft_clip(m[0].arr, d[0].arr, 0, 1, self.xmax, self.l, self.tau, add=add)
        _pass(add, m[1].arr, d[1].arr)


class Softmin(Operator):
    """Re S -> -h(-(Re S - xmin)) + xmin (Operator.py:693-710)."""

    def __init__(self, model, data, xmin, l=1, tau=.001):
        super().__init__(model, data)
        self.l, self.tau, self.xmin = l / 2, tau, xmin

    def forward(self, add, model, data):
        m, d = _vecs(model), _vecs(data)
        _soft_clip(m[0].arr, d[0].arr, 0, -1, self.xmin, self.l, self.tau, add=add)
        _pass(add, m[1].arr, d[1].arr)


class dSoftmax(_Jacobian):  # noqa: N801 - reference name
    """Jacobian of Softmax: Re scaled by h'(Re bg - xmax) (Operator.py:712-739)."""

    comp = 0
    sign = 1

    def __init__(self, model, data, xmax, l=1, tau=.001):
        super().__init__(model, data, l, tau, x0=xmax)


class dSoftmin(_Jacobian):  # noqa: N801 - reference name
    """Jacobian of Softmin: Re scaled by h'(-(Re bg - xmin)) (Operator.py:741-768)."""

    comp = 0
    sign = -1

    def __init__(self, model, data, xmin, l=1, tau=.001):
        super().__init__(model, data, l, tau, x0=xmin)


class NonLinearOperator:
    """Pair of a non-linear map and its Jacobian (python-solver's
    pyOperator.NonLinearOperator): set_background re-linearises."""

    def __init__(self, nl_op, lin_op, set_background_func=None):
        self.nl_op = nl_op
        self.lin_op = lin_op
        self._set = set_background_func or getattr(lin_op, "set_background", None)

    def forward(self, add, model, data):
        self.nl_op.forward(add, model, data)

    def set_background(self, model):
        if self._set is not None:
            self._set(model)


class CombNonlinearOp:
    """f(g(m)) of two NonLinearOperators g = op1, f = op2 (python-solver's
    CombNonlinearOp); Jacobian F'(g(m)) G'(m)."""

    def __init__(self, op1, op2, tmp):
        self.op1, self.op2 = op1, op2
        self.tmp = tmp          # a vector of g's range

    def forward(self, add, model, data):
        self.op1.forward(False, model, self.tmp)
        self.op2.forward(add, self.tmp, data)

    def set_background(self, model):
        self.op1.set_background(model)
        self.op1.forward(False, model, self.tmp)
        self.op2.set_background(self.tmp)


def SoftMinMax(model, data, xmin, xmax, l=1, tau=(0.001, 0.001)):  # noqa: N802 - reference name
    """Soft clip of Re S to [xmin, xmax]: Softmin(Softmax(m)) with its Jacobian
    (Operator.py:771-775)."""
    mx = NonLinearOperator(Softmax(model, data, xmax, l=l, tau=tau[0]), dSoftmax(model, data, xmax, l=l, tau=tau[0]))
    mn = NonLinearOperator(Softmin(model, data, xmin, l=l, tau=tau[1]), dSoftmin(model, data, xmin, l=l, tau=tau[1]))
    return CombNonlinearOp(mx, mn, data.clone())


# ------------------------------------------------------------------ DSO

def derivative_filter(which: int = 1, order=4, dt: Optional[float] = None, tmax: Optional[float] = None):
    """Taps of Derivative (Operator.py:241-268), float64 as there.  The 'exact'
    filter is f(w) = tmax cos(w tmax/2)/w - 2 sin(w tmax/2)/w^2 (0 at w = 0) on
    the reversed angular-frequency grid of t = arange(-tmax, tmax, dt), every
    second sample: f(w) = i int_{-tmax/2}^{tmax/2} t e^{i w t} dt, a lag
    weighting in the frequency domain."""
    if which == 1:
        if order == 1:
            return np.array([-1.0, 1.0])
        if order == 2:
            return np.array([-1 / 2, 0, 1 / 2])
        if order == 4:
            return np.array([1 / 12, -2 / 3, 0, 2 / 3, -1 / 12])
        if order == 8:
            return np.array([1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280])
        if order == "exact":
            if dt is None:
                raise ValueError("Must provide dt!")
            t = np.arange(-tmax, tmax, dt)
            w = 2 * np.pi * np.linspace(-1 / (2 * dt), 1 / (2 * dt), t.size)
            w = np.flip(w, axis=0)
            with np.errstate(divide="ignore", invalid="ignore"):
                f1 = (np.cos(w * tmax / 2) / w) * tmax
                f2 = -2 * (np.sin(w * tmax / 2) / w ** 2)
            f1 = np.where(w == 0, 0, f1)
            f2 = np.where(w == 0, 0, f2)
            return (f1 + f2)[::2]
    if which == 2:
        if order == 2:
            return np.array([1.0, -2.0, 1.0])
        if order == 4:
            return np.array([-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12])
    raise ValueError(f"unsupported derivative which={which} order={order}")


class Derivative(Operator):
    """DSO regularisation along the frequency axis (axis 0 of each [nw][nz][nx]
    vector): forward d += i beta corr(m) + i alpha mean_w(Im m); adjoint
    m += -i beta conv(d) + i alpha mean_w(Im d), with scipy.ndimage's
    correlate1d / convolve1d semantics and boundary mode (Operator.py:228-292;
    the adjoint is exact in mode 'constant')."""

    def __init__(self, model, data, which=1, order=4, f0=None, dt=None, tmax=None, alpha=0, beta=None,
                 mode="reflect"):
        if tmax is None:
            raise ValueError("Must provide tmax!")
        if f0 is None:
            raise ValueError("Must provide f0!")
        super().__init__(model, data)
        self.alpha = float(alpha)
        self.beta = 1.0 if beta is None else float(beta)
        if mode not in _MODES:
            raise ValueError(f"unsupported mode {mode}")
        self.mode = mode
        self.filter = derivative_filter(which, order, dt, tmax)
        nw = _vecs(model)[0].arr.shape[0]
        dev = _vecs(model)[0].arr.device
        self.M = {}
        for conv in (0, 1):
            M = np.zeros((nw, nw))
            w = np.ascontiguousarray(self.filter, dtype=np.float64)
            check(lib.wemfwi_filter_matrix(nw, w.size, w.ctypes.data_as(C.POINTER(C.c_double)), _MODES[mode], conv,
                                           M.ctypes.data_as(C.POINTER(C.c_double))))
            self.M[conv] = torch.from_numpy(M.astype(np.float32)).to(dev)

    def _apply(self, conv, c, inp: torch.Tensor, out: torch.Tensor, add: bool):
        nw = inp.shape[0]
        if out.shape != inp.shape or self.M[conv].shape[0] != nw:
            raise ValueError("Derivative: shapes")
        src = inp if inp.data_ptr() != out.data_ptr() else inp.clone()
        check(lib.wemfwi_freq_filter(nw, inp.numel() // nw, _ptr(self.M[conv]), float(c.real), float(c.imag),
                                     self.alpha, _ptr(src), _ptr(out), int(add), _stream_ptr(None)))

    def forward(self, add, model, data):
        for m, d in zip(_vecs(model), _vecs(data)):
            self._apply(0, 1j * self.beta, m.arr, d.arr, add)

    def adjoint(self, add, model, data):
        for m, d in zip(_vecs(model), _vecs(data)):
            self._apply(1, -1j * self.beta, d.arr, m.arr, add)


class Mask(Operator):
    """m * mask per vector (Operator.py:533-546), self-adjoint."""

    def __init__(self, model, data, mask):
        super().__init__(model, data)
        self.mask = [t.arr if isinstance(t, Vector) else t for t in _vecs(mask)]

    def forward(self, add, model, data):
        for m, d, k in zip(_vecs(model), _vecs(data), self.mask):
            v = m.arr * k
            if add:
                d.arr.add_(v)
            else:
                d.arr.copy_(v)

    def adjoint(self, add, model, data):
        self.forward(add, data, model)


# ------------------------------------------------------------------ problem

class ExtFWIObjective:
    """The efwi.py problem on [nw][nz][nx] slowness and density models:
    value(m) = J and gradient(m) = grad J as above.  par carries the
    operator parameters (the geometry JSON of efwi.py), tau the
    HyperbolicPenalty width (geom['tau']), epsilon / alpha / beta / der_mode the
    regularisation (efwi.py:194-196,204-206); vmin / vmax add SoftMinMax on the
    slowness with tau = (0.01, 0.001) (efwi.py:168-171); mask multiplies the
    Born input (efwi.py:145-149); checkpoint as for extBorn."""

    def __init__(self, obs: Vector, wave: Vector, par, hyper: Hypercube, epsilon: float, alpha: float = 0.0,
                 beta: float = 1.0, der_mode: str = "reflect", tau: float = 0.001, vmin: Optional[float] = None,
                 vmax: Optional[float] = None, mask: Optional[Sequence[torch.Tensor]] = None,
                 adjoint_mode: str = "exact", checkpoint: int = 0):
        self.obs = obs
        self.checkpoint = int(checkpoint)
        self.wave = wave
        self.par = dict(par)
        self.hyper = hyper
        self.eps = float(epsilon)
        proto = SuperVector(Vector(hyper), Vector(hyper))
        self._proto = proto
        self.hyp = NonLinearOperator(HyperbolicPenalty(proto, proto, l=1, tau=tau), Softclip(proto, proto, l=1, tau=tau))
        self.clip = None
        if vmin is not None and vmax is not None:
            self.clip = SoftMinMax(proto, proto.clone(), 1 / vmax ** 2, 1 / vmin ** 2, tau=(0.01, 0.001))
        wa = wave.getHyper().getAxis(1)
        dt = wa.d
        tmax = (wa.n - 1) * dt + wa.o
        self.der = Derivative(proto, proto, which=1, order="exact", f0=float(self.par.get("fmin", 1.0)),
                              tmax=tmax / 2, dt=dt, alpha=alpha, beta=beta, mode=der_mode)
        self.mask = None if mask is None else Mask(proto, proto, list(mask))
        self.adjoint_mode = adjoint_mode
        self._wem = extWEM(proto, obs, wave, self.par)

    def _param(self, m: SuperVector):
        """p = P(m) and the linearisation state of P."""
        q = m
        if self.clip is not None:
            self.clip.set_background(m)
            q = m.clone()
            self.clip.forward(False, m, q)
        self.hyp.set_background(q)
        p = q.clone()
        self.hyp.forward(False, q, p)
        return p

    def residual(self, m: SuperVector):
        p = self._param(m)
        pred = self.obs.clone().zero()
        self._wem.forward(False, p, pred)
        pred.arr.sub_(self.obs.arr)
        return p, pred

    def value(self, m: SuperVector) -> float:
        _, res = self.residual(m)
        dm = m.clone().zero()
        self.der.forward(False, m, dm)
        return 0.5 * cdot(res.arr, res.arr) + 0.5 * self.eps ** 2 * dm.dot(dm)

    def value_and_gradient(self, m: SuperVector):
        p, res = self.residual(m)
        born = extBorn(p, self.obs, self.wave, self.par, adjoint_mode=self.adjoint_mode, checkpoint=self.checkpoint)
        g = m.clone().zero()
        born.adjoint(False, g, res)
        if self.mask is not None:
            self.mask.adjoint(False, g, g.clone())
        g2 = g.clone()
        self.hyp.lin_op.adjoint(False, g2, g)            # Softclip at q
        if self.clip is not None:                        # dSoftmin at Softmax(m), then dSoftmax at m
            t = g2.clone()
            self.clip.op2.lin_op.adjoint(False, t, g2)
            self.clip.op1.lin_op.adjoint(False, g2, t)
        dm = m.clone().zero()
        self.der.forward(False, m, dm)
        dd = m.clone().zero()
        self.der.adjoint(False, dd, dm)
        for a, b in zip(g2.vecs, dd.vecs):
            a.arr.add_(b.arr, alpha=self.eps ** 2)
        J = 0.5 * cdot(res.arr, res.arr) + 0.5 * self.eps ** 2 * dm.dot(dm)
        return J, g2
