"""The reference's unit operators (propagator/src/python/py_wem_ops.cpp:25-206)
on the MI355X library: RefSampler, Phshift, SplitStep, SSF / PSPI (one depth
step), Down / Up, Scatter, LinDown / LinUp, Reflect, dReflect, IC and
Injection, with the same constructor arguments, setFreq / setDepth / setSlow /
setLocation / setStep / setShot state and
forward(model, data, add) / adjoint(model, data, add) argument order as the
C++ classes.

Slowness and density are api.Vector objects on a Hypercube([nx, nz]) (the
propagation grid; dx, dz from the axes); wavefields are complex64 CUDA tensors
of one row [nx] / one wavefield [nz][nx], or a batch with a leading row axis.
setFreq takes the frequency in Hz, as the reference's setFreq(freq[i]).
"""

from __future__ import annotations

import ctypes as C
from typing import Mapping, Optional, Sequence

import numpy as np
import torch

from ._lib import check, lib
from .api import Vector, _host_c64
from .core import Model, _dev_c64, _stream_ptr
from .core import refsampler as _refsampler


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _model(slow: Vector, par: Mapping, den: Optional[Vector] = None) -> Model:
    ax_x, ax_z = slow.getHyper().getAxis(1), slow.getHyper().getAxis(2)
    return Model(_host_c64(slow), ax_x.d, ax_z.d, tap=int(par.get("tap", 0)), pad=int(par.get("pad", 0)),
                 nref=int(par.get("nref", 1)), prop=str(par.get("prop", "ssf")), eps=float(par.get("eps", 0.04)),
                 ntaylor=int(par.get("ntaylor", 1)), den=None if den is None else _host_c64(den))


def _rows(t: torch.Tensor, row_shape) -> int:
    _dev_c64(t, "wavefield")
    if tuple(t.shape[-len(row_shape):]) != tuple(row_shape):
        raise ValueError(f"wavefield shape {tuple(t.shape)} does not end in {tuple(row_shape)}")
    return int(np.prod(t.shape[:-len(row_shape)], dtype=np.int64)) if t.dim() > len(row_shape) else 1


class RefSampler:
    """Reference slownesses and their x positions per depth (RefSampler.cpp:10-91)."""

    def __init__(self, slow: Vector, nref: int):
        self.nref = int(nref)
        self.sref, self.refid = _refsampler(_host_c64(slow), self.nref)

    def getRefSlow(self, iz: int, iref: Optional[int] = None):  # noqa: N802 - reference name
        """Reference slowness iref of depth iz (RefSampler.h:19); without
        iref, the list of all nref of them."""
        if iref is None:
            return [complex(v) for v in self.sref[iz]]
        return complex(self.sref[iz][int(iref)])

    def getRefLoc(self, iz: int, iref: int):  # noqa: N802
        return list(np.nonzero(self.refid[iz] == iref)[0])


class _Stateful:
    freq = 0.0
    depth = 0

    def setFreq(self, f: float):  # noqa: N802
        self.freq = float(f)

    def setDepth(self, iz: int):  # noqa: N802
        self.depth = int(iz)


class Phshift:
    """The phase shift of one reference slowness in the wavenumber domain
    (Phshift.h:15-72, Phshift_cmplx.cpp:8-113, phase_shift.ispc:2-32) on rows
    [..., nk], nk and the spacing from the first axis of hyper:
    forward data (+)= model exp(-i kz dz), adjoint model (+)= data times the
    conjugate factor, kz the damped vertical wavenumber of the slowness set by
    setSlow at the frequency (Hz) set by setFreq.  nref and
    setRef select the reference for the reflection helper of the C++ class,
    which the Python bindings do not expose; they are kept as state."""

    def __init__(self, dz: float, hyper, nref: int, eps: float = 0.04):
        ax = hyper.getAxis(1)
        self.nk, self.dx = int(ax.n), float(ax.d)
        if self.nk < 2:
            raise ValueError("Phshift: the wavenumber axis needs at least 2 samples")
        self.dz, self.nref, self.eps = float(dz), int(nref), float(eps)
        self.freq, self.slow, self.iref = 0.0, complex(0.0), 0

    def setFreq(self, f: float):  # noqa: N802
        self.freq = float(f)

    def setSlow(self, s: complex):  # noqa: N802
        self.slow = complex(s)

    def setRef(self, iref: int):  # noqa: N802
        self.iref = int(iref)

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.nk,))
        if tuple(dst.shape) != tuple(src.shape):
            raise ValueError("Phshift: model and data shapes differ")
        _dev_c64(dst, "wavefield")
        f = np.full(n, self.freq, np.float32)
        sr = np.full(n, self.slow, np.complex64)
        check(lib.wemfwi_phshift(self.nk, self.dx, self.dz, self.eps, int(adjoint), n,
                                 f.ctypes.data_as(C.POINTER(C.c_float)), C.c_void_p(sr.ctypes.data), _ptr(src),
                                 _ptr(dst), int(bool(add)), _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class SplitStep(_Stateful):
    """The split-step correction of one reference slowness (SplitStep.h:8-32,
    SplitStep.cpp:5-29) on rows [..., nx] of the slowness grid: at the
    positions set by setLocation, forward data[x] = model[x] exp(-i w dz
    (S[iz][x] - sref)/sqrt(sref)), adjoint model[x] = data[x] exp(+i w dz
    conj(...)); the other samples are zero (add = False) or left as they
    were (add = True), as in the reference."""

    def __init__(self, slow: Vector):
        ax_x, ax_z = slow.getHyper().getAxis(1), slow.getHyper().getAxis(2)
        self.m = Model(_host_c64(slow), ax_x.d, ax_z.d)
        self.loc = np.zeros(0, np.int32)
        self.slow = complex(1.0)

    def setLocation(self, loc: Sequence[int]):  # noqa: N802
        self.loc = np.ascontiguousarray(np.asarray(loc, dtype=np.int32).reshape(-1))

    def setSlow(self, s: complex):  # noqa: N802
        self.slow = complex(s)

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.m.nx,))
        if tuple(dst.shape) != tuple(src.shape):
            raise ValueError("SplitStep: model and data shapes differ")
        _dev_c64(dst, "wavefield")
        f = np.full(n, self.freq, np.float32)
        d = np.full(n, self.depth, np.int32)
        sr = np.full(n, self.slow, np.complex64)
        check(lib.wemfwi_splitstep(self.m._h, int(adjoint), n, f.ctypes.data_as(C.POINTER(C.c_float)),
                                   d.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(sr.ctypes.data), len(self.loc),
                                   self.loc.ctypes.data_as(C.POINTER(C.c_int)), _ptr(src), _ptr(dst),
                                   int(bool(add)), _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class SSF(_Stateful):
    """One depth step (SSF.cpp:8-60; PSPI.cpp:8-53 when par['prop'] == 'pspi')
    on rows [..., nx] at the current frequency and depth."""

    def __init__(self, slow: Vector, par: Mapping, ref: Optional[RefSampler] = None):
        self.m = _model(slow, par)

    def _apply(self, adjoint, model, data, add):
        n = _rows(model, (self.m.nx,))
        f = np.full(n, self.freq, np.float32)
        d = np.full(n, self.depth, np.int32)
        check(lib.wemfwi_onestep(self.m._h, int(adjoint), n, f.ctypes.data_as(C.POINTER(C.c_float)),
                                 d.ctypes.data_as(C.POINTER(C.c_int)), _ptr(model if not adjoint else data),
                                 _ptr(data if not adjoint else model), int(bool(add)), _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


PSPI = SSF


class _OneWay(_Stateful):
    up = False

    def __init__(self, slow: Vector, par: Mapping, ref: Optional[RefSampler] = None):
        self.m = _model(slow, par)

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.m.nz, self.m.nx))
        self.m.oneway(self.up, adjoint, [self.freq] * n, src, dst, add=add)

    def forward(self, model, data, add=False):
        """In place (model is data) runs the recursion of OneWay.cpp:32-109; with
        separate vectors one bidiagonal step per depth, as the reference."""
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class Down(_OneWay):
    up = False


class Up(_OneWay):
    up = True


class Scatter(_Stateful):
    """Scatter (Scatter.cpp:6-111) on rows [..., nx]; adjoint_mode 'compat'
    keeps the reference's operator order, 'exact' is the true adjoint."""

    def __init__(self, slow: Vector, ntaylor: int, par: Mapping, adjoint_mode: str = "compat"):
        p = dict(par)
        p["ntaylor"] = int(ntaylor)
        self.m = _model(slow, p)
        self.mode = 1 if adjoint_mode == "exact" else 0

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.m.nx,))
        f = np.full(n, self.freq, np.float32)
        d = np.full(n, self.depth, np.int32)
        check(lib.wemfwi_scatter(self.m._h, int(adjoint), self.mode, n, f.ctypes.data_as(C.POINTER(C.c_float)),
                                 d.ctypes.data_as(C.POINTER(C.c_int)), _ptr(src), _ptr(dst), int(bool(add)),
                                 _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class _Lin(_Stateful):
    up = False

    def __init__(self, slow: Vector, par: Mapping, bg: torch.Tensor, oneway=None, adjoint_mode: str = "compat"):
        self.m = _model(slow, par)
        self.bg = _dev_c64(bg, "bg")
        self.mode = 1 if adjoint_mode == "exact" else 0

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.m.nz, self.m.nx))
        if self.bg.shape != src.shape:
            raise ValueError("LinDown/LinUp: background and wavefield shapes differ")
        f = np.full(n, self.freq, np.float32)
        check(lib.wemfwi_lin_oneway(self.m._h, int(self.up), int(adjoint), self.mode, n,
                                    f.ctypes.data_as(C.POINTER(C.c_float)), _ptr(self.bg), _ptr(src), _ptr(dst),
                                    int(bool(add)), _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class LinDown(_Lin):
    """LinDown (LinOneWay.cpp:33-87) around the background wavefield bg."""
    up = False


class LinUp(_Lin):
    """LinUp (LinOneWay.cpp:89-142) around the background wavefield bg."""
    up = True


class Reflect:
    """Reflection coefficients of [slow, den] (Reflect.h:12-48)."""

    def __init__(self, model: Sequence[Vector]):
        self.m = Model(_host_c64(model[0]), 1.0, 1.0, den=_host_c64(model[1]))

    def _apply(self, adjoint, model, data, add):
        src, dst = (model, data) if not adjoint else (data, model)
        n = _rows(src, (self.m.nz, self.m.nx))
        check(lib.wemfwi_reflect(self.m._h, int(adjoint), n, _ptr(src), _ptr(dst), int(bool(add)),
                                 _stream_ptr(None)))

    def forward(self, model, data, add=False):
        self._apply(False, model, data, add)

    def adjoint(self, model, data, add=False):
        self._apply(True, model, data, add)


class dReflect:  # noqa: N801 - reference name
    """Linearised reflection coefficients of [slow, den] (dReflect.h:11-84):
    model = [dslow, dden], data = dr, all [nz][nx]."""

    def __init__(self, model: Sequence[Vector]):
        self.m = Model(_host_c64(model[0]), 1.0, 1.0, den=_host_c64(model[1]))

    def forward(self, model, data, add=False):
        dS, dR = model
        check(lib.wemfwi_dreflect(self.m._h, 0, _ptr(_dev_c64(dS, "dslow")), _ptr(_dev_c64(dR, "dden")),
                                  _ptr(_dev_c64(data, "dr")), int(bool(add)), _stream_ptr(None)))

    def adjoint(self, model, data, add=False):
        dS, dR = model
        check(lib.wemfwi_dreflect(self.m._h, 1, _ptr(_dev_c64(dS, "dslow")), _ptr(_dev_c64(dR, "dden")),
                                  _ptr(_dev_c64(data, "dr")), int(bool(add)), _stream_ptr(None)))


class IC:
    """Zero-lag imaging condition with background bg (ic.cpp:5-45): forward
    d = bg * m, adjoint m = conj(bg) * d; on a whole [nz][nx] field, or on one
    row (setDepth) when the model is a row."""

    def __init__(self, bg: torch.Tensor):
        self.bg = _dev_c64(bg, "bg")
        self.depth = 0

    def setDepth(self, iz: int):  # noqa: N802
        self.depth = int(iz)

    def _bg_for(self, t):
        if t.dim() == 1 and self.bg.dim() == 2:
            return self.bg[self.depth]
        return self.bg

    def forward(self, model, data, add=False):
        bg = self._bg_for(model)
        check(lib.wemfwi_ic(0, model.numel(), _ptr(bg), _ptr(_dev_c64(model, "model")), _ptr(_dev_c64(data, "data")),
                            int(bool(add)), _stream_ptr(None)))

    def adjoint(self, model, data, add=False):
        bg = self._bg_for(model)
        check(lib.wemfwi_ic(1, model.numel(), _ptr(bg), _ptr(_dev_c64(model, "model")), _ptr(_dev_c64(data, "data")),
                            int(bool(add)), _stream_ptr(None)))


class Injection:
    """Devices at (z[i], x[i]) with cos^2 spread ng and x offset tap
    (Injection.h:19-65, Injection.cpp:8-80).  setStep selects the frequency
    bin, setShot the shot.  model is a wavelet [nt] (every device injects
    wavelet[step]), a data cube [ns][nr][nt] (device i injects
    data[shot][i][step]) or a [ndev][nt] array (no spread, no offset: the
    two-dimensional form); data is a wavefield [nz][nx]."""

    def __init__(self, z: Sequence[int], x: Sequence[int], ng: int = 0, tap: int = 0):
        self.iz = np.ascontiguousarray(np.asarray(z, dtype=np.int32))
        self.ix = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
        self.ng, self.tap = int(ng), int(tap)
        self.step = 0
        self.shot = 0

    def setStep(self, i: int):  # noqa: N802
        self.step = int(i)

    def setShot(self, s: int):  # noqa: N802
        self.shot = int(s)

    def _call(self, adjoint, vals, wfld, add, ng, tap):
        nz, nx = wfld.shape[-2:]
        check(lib.wemfwi_injection(nz, nx, len(self.ix), self.iz.ctypes.data_as(C.POINTER(C.c_int)),
                                   self.ix.ctypes.data_as(C.POINTER(C.c_int)), ng, tap, int(adjoint), 1,
                                   _ptr(vals), _ptr(wfld), int(bool(add)), _stream_ptr(None)))

    def _vals(self, model):
        n = len(self.ix)
        if model.dim() == 1:
            return model[self.step].expand(n).contiguous(), (self.ng, self.tap)
        if model.dim() == 3:
            return model[self.shot, :, self.step].contiguous(), (self.ng, self.tap)
        return model[:, self.step].contiguous(), (0, 0)

    def forward(self, model, data, add=False):
        _dev_c64(data, "wavefield")
        vals, (ng, tap) = self._vals(_dev_c64(model, "model"))
        self._call(False, vals, data, add, ng, tap)

    def adjoint(self, model, data, add=False):
        _dev_c64(data, "wavefield")
        _dev_c64(model, "model")
        if not add:
            model.zero_()          # the reference scales the whole model by 0 (Injection.cpp:22,47,71)
        n = len(self.ix)
        ng, tap = (self.ng, self.tap) if model.dim() != 2 else (0, 0)
        vals = torch.zeros(n, dtype=torch.complex64, device=data.device)
        self._call(True, vals, data, False, ng, tap)
        if model.dim() == 1:
            model[self.step] += vals.sum()
        elif model.dim() == 3:
            model[self.shot, :, self.step] += vals
        else:
            model[:, self.step] += vals
