"""Model interpolation operators of the FWI preconditioning (scripts/fwi.py:46-68,
efwi.py:48-80) on the library: a coarse model on a regular grid mapped to the
propagation grid by separable filters, one axis at a time on the device.

  LanczosInterpolation2D / 3D   operator/src/LanczosInterpolation2D.h,
                                LanczosInterpolation3D.h, Operator.py:413-445
  Spline2D / Spline3D           operator/src/Spline2D.h, Spline3D.h,
                                Operator.py:447-494

Filters follow the reference's construction (wemfwi_interp_filter); axis 1 is
the fastest (x), as in the SEP hypercubes.  Spline2D's adjoint copies the
model's second row/column into the first and the next-to-last into the last
after the transpose (Spline2D.h:84-92; not an exact adjoint, as the reference
notes).  The spline `type` selects (a, b) = (0.5, 0) CR, (0, 1) B, (1/3, 1/3)
MN -- the values the reference's operators carry after pickling, which is how
its Dask drivers run them; its unpickled objects use the a, b arguments.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np
import torch

from ._lib import check, lib
from .api import Operator
from .core import _stream_ptr

_TYPES = {"CR-spline": (0.5, 0.0), "B-spline": (0.0, 1.0), "MN-spline": (1 / 3, 1 / 3)}


def interp_filter(kind: int, nd: int, dd: float, nm: int, dm: float, a: float, b: float = 0.0,
                  taper: float = 0.0) -> np.ndarray:
    """Host filter matrix [nd][nm] (kind 0 Lanczos, 1 spline)."""
    F = np.zeros((nd, nm), dtype=np.float32)
    check(lib.wemfwi_interp_filter(int(kind), int(nd), float(dd), int(nm), float(dm), float(a), float(b),
                                   float(taper), F.ctypes.data_as(C.POINTER(C.c_float))))
    return F


def _arr(v):
    return v.arr if hasattr(v, "arr") else v


class _Separable(Operator):
    """Separable interpolation: data = F_k ... F_1 model over the axes."""

    def __init__(self, model, data, filters):
        super().__init__(model, data)
        m, d = _arr(model), _arr(data)
        self.ms, self.ds = tuple(m.shape), tuple(d.shape)
        # filters[k] acts on SEP axis k+1 = numpy axis -(k+1)
        self.F = [torch.from_numpy(np.ascontiguousarray(f)).to(m.device) for f in filters]

    def _chain(self, x: torch.Tensor, adjoint: bool) -> torch.Tensor:
        nd = len(self.F)
        order = range(nd) if not adjoint else reversed(range(nd))
        shape = list(x.shape)
        for k in order:
            ax = len(shape) - 1 - k
            n_in = shape[ax]
            F = self.F[k]
            n_out = F.shape[1] if adjoint else F.shape[0]
            outer = int(np.prod(shape[:ax])) if ax > 0 else 1
            inner = int(np.prod(shape[ax + 1:])) if ax + 1 < len(shape) else 1
            shape[ax] = n_out
            y = torch.empty(shape, dtype=torch.complex64, device=x.device)
            xc = x.contiguous().to(torch.complex64)   # kept alive across the call
            check(lib.wemfwi_axis_apply(C.c_void_p(F.data_ptr()), n_out, n_in, outer, inner, int(adjoint),
                                        C.c_void_p(xc.data_ptr()), C.c_void_p(y.data_ptr()), 0,
                                        _stream_ptr(None)))
            x = y
        return x

    def forward(self, add, model, data):
        y = self._chain(_arr(model), False)
        d = _arr(data)
        d.add_(y) if add else d.copy_(y)

    def adjoint(self, add, model, data):
        y = self._chain(_arr(data), True)
        self._post_adjoint(y)
        m = _arr(model)
        m.add_(y) if add else m.copy_(y)

    def _post_adjoint(self, y):
        pass


def _axes(v, n):
    h = v.getHyper()
    return [h.getAxis(i + 1) for i in range(n)]


class LanczosInterpolation2D(_Separable):
    def __init__(self, model, data, a: Sequence[float] = (3, 3, 3)):
        am, ad = _axes(model, 2), _axes(data, 2)
        super().__init__(model, data, [interp_filter(0, ad[k].n, ad[k].d, am[k].n, am[k].d, a[k]) for k in range(2)])


class LanczosInterpolation3D(_Separable):
    def __init__(self, model, data, a: Sequence[float] = (3, 3, 3), taper: Sequence[float] = (0, 0, 0)):
        am, ad = _axes(model, 3), _axes(data, 3)
        super().__init__(model, data, [interp_filter(0, ad[k].n, ad[k].d, am[k].n, am[k].d, a[k], 0.0, taper[k])
                                       for k in range(3)])


class Spline2D(_Separable):
    def __init__(self, model, data, a=1, b=0, type=None):  # noqa: A002 - reference name
        a, b = _TYPES.get(type, (a, b))
        self.a, self.b = a, b
        am, ad = _axes(model, 2), _axes(data, 2)
        super().__init__(model, data, [interp_filter(1, ad[k].n, ad[k].d, am[k].n, am[k].d, a, b) for k in range(2)])

    def _post_adjoint(self, y):
        # Spline2D.h:84-92
        y[:, 0] = y[:, 1]
        y[:, -1] = y[:, -2]
        y[0, :] = y[1, :]
        y[-1, :] = y[-2, :]


class Spline3D(_Separable):
    def __init__(self, model, data, a=1, b=0, type=None, taper=(0, 0, 0)):  # noqa: A002 - reference name
        a, b = _TYPES.get(type, (a, b))
        self.a, self.b = a, b
        am, ad = _axes(model, 3), _axes(data, 3)
        super().__init__(model, data, [interp_filter(1, ad[k].n, ad[k].d, am[k].n, am[k].d, a, b, taper[k])
                                       for k in range(3)])
