"""The reference's Python operator API on the MI355X library.

Mirrors propagator/src/python/WEM.py (classes WEM, Born, BornRefl) and
propagator/src/python/newExtWEM.py (extWEM, extBorn): the same constructor
arguments (model, data, wave, par), the same forward(add, model, data) /
adjoint(add, model, data) argument meaning, set_background, Born's
mode='real', and the reference's paramObj keys.  SEP vectors are replaced by a
minimal Hypercube / Vector / SuperVector whose storage is a device-resident
torch tensor (complexHyper storage, complex/include/complexHyper.h:14-107).

Axis convention (SEP): axis 1 is the fastest, so a model on
Hypercube(ns=[nx, nz]) is a tensor of shape (nz, nx); data on
Hypercube(ns=[nt, nr, ns]) has shape (ns, nr, nt).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Mapping, Optional, Sequence

import numpy as np
import torch

from ._lib import check, lib
from .core import Model, _stream_ptr
from . import ops
from .params import CMapPlan, cmap_from_par, device_indices, prop_param


# --------------------------------------------------------------------- vectors

@dataclass
class Axis:
    n: int
    o: float = 0.0
    d: float = 1.0


class Hypercube:
    """Regular axes, fastest first (SEP hypercube)."""

    def __init__(self, ns: Sequence[int], os: Optional[Sequence[float]] = None,
                 ds: Optional[Sequence[float]] = None):
        os = list(os) if os is not None else [0.0] * len(ns)
        ds = list(ds) if ds is not None else [1.0] * len(ns)
        self.axes = [Axis(int(n), float(o), float(d)) for n, o, d in zip(ns, os, ds)]

    def getAxis(self, i: int) -> Axis:  # noqa: N802 - reference name
        return self.axes[i - 1]

    def getNdim(self) -> int:  # noqa: N802
        return len(self.axes)

    @property
    def shape(self):
        return tuple(a.n for a in reversed(self.axes))


class Vector:
    """A complex (or float) array on a hypercube, stored on the GPU."""

    def __init__(self, hyper: Hypercube, storage: str = "dataComplex", device="cuda",
                 arr: Optional[torch.Tensor] = None):
        self.hyper = hyper
        dtype = torch.complex64 if storage == "dataComplex" else torch.float32
        if arr is None:
            arr = torch.zeros(hyper.shape, dtype=dtype, device=device)
        if tuple(arr.shape) != hyper.shape:
            raise ValueError(f"array shape {tuple(arr.shape)} does not match hypercube {hyper.shape}")
        self.arr = arr

    def getHyper(self) -> Hypercube:  # noqa: N802
        return self.hyper

    def getNdArray(self) -> torch.Tensor:  # noqa: N802
        return self.arr

    def clone(self) -> "Vector":
        return Vector(self.hyper, arr=self.arr.clone())

    def zero(self) -> "Vector":
        self.arr.zero_()
        return self

    def set(self, v) -> "Vector":
        self.arr.fill_(v)
        return self

    def scale(self, s) -> "Vector":
        self.arr.mul_(s)
        return self

    def rand(self, generator: Optional[torch.Generator] = None) -> "Vector":
        """Uniform in [-1/2, 1/2) for real and imaginary parts (complexHyper::random)."""
        if self.arr.is_complex():
            re = torch.rand(self.arr.shape, generator=generator, device=self.arr.device) - 0.5
            im = torch.rand(self.arr.shape, generator=generator, device=self.arr.device) - 0.5
            self.arr.copy_(torch.complex(re, im))
        else:
            self.arr.copy_(torch.rand(self.arr.shape, generator=generator, device=self.arr.device) - 0.5)
        return self

    def dot(self, other: "Vector") -> float:
        """Re sum a * conj(b) in double precision (complexHyper::dot, complexHyper.cc:73-84)."""
        return cdot(self.arr, other.arr)

    def norm(self) -> float:
        return float(np.sqrt(self.dot(self)))


class SuperVector:
    """A list of vectors treated as one (pyVector.superVector)."""

    def __init__(self, *vecs):
        if len(vecs) == 1 and isinstance(vecs[0], (list, tuple)):
            vecs = tuple(vecs[0])
        self.vecs: List[Vector] = list(vecs)

    def __getitem__(self, i):
        return self.vecs[i]

    def __len__(self):
        return len(self.vecs)

    def __iter__(self):
        return iter(self.vecs)

    def clone(self) -> "SuperVector":
        return SuperVector(*[v.clone() for v in self.vecs])

    def zero(self):
        for v in self.vecs:
            v.zero()
        return self

    def rand(self, generator=None):
        for v in self.vecs:
            v.rand(generator)
        return self

    def scale(self, s):
        for v in self.vecs:
            v.scale(s)
        return self

    def dot(self, other: "SuperVector") -> float:
        return sum(a.dot(b) for a, b in zip(self.vecs, other.vecs))


def Trace(values, d: float, o: float = 0.0) -> Vector:  # noqa: N802 - float1DReg stand-in
    """A float 1-D vector (the wavelet, float1DReg) with sampling d and origin o."""
    v = torch.as_tensor(np.asarray(values, dtype=np.float32))
    return Vector(Hypercube([v.numel()], [o], [d]), storage="dataFloat", device="cpu", arr=v)


def cdot(a: torch.Tensor, b: torch.Tensor) -> float:
    """Re sum a * conj(b), accumulated in double precision on the device."""
    if a.is_complex():
        return float((a.to(torch.complex128) * b.to(torch.complex128).conj()).real.sum().item())
    return float((a.double() * b.double()).sum().item())


# ----------------------------------------------------------------- operators

class Operator:
    """forward/adjoint pair with the reference's dot-product test
    (operator/src/Operator.h:23-48)."""

    def __init__(self, domain, rng):
        self.domain = domain
        self.range = rng

    def forward(self, add, model, data):  # pragma: no cover - abstract
        raise NotImplementedError

    def adjoint(self, add, model, data):  # pragma: no cover - abstract
        raise NotImplementedError

    def dotTest(self, verbose: bool = False, seed: int = 0):  # noqa: N802
        """<L m, d> against <m, L^H d> for random m, d (add = False and True);
        returns the two relative mismatches."""
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        m = self.domain.clone().rand(g)
        d = self.range.clone().rand(g)
        m1 = self.domain.clone().zero()
        d1 = self.range.clone().zero()
        self.forward(False, m, d1)
        self.adjoint(False, m1, d)
        a, b = d1.dot(d), m.dot(m1)
        err0 = abs(a - b) / max(abs(a), abs(b), 1e-300)
        self.forward(True, m, d1)
        self.adjoint(True, m1, d)
        a2, b2 = d1.dot(d), m.dot(m1)
        err1 = abs(a2 - b2) / max(abs(a2), abs(b2), 1e-300)
        if verbose:
            print(f"dot test: <Lm,d> = {a:.8e}  <m,L*d> = {b:.8e}  rel {err0:.2e}; add: rel {err1:.2e}")
        return err0, err1


def _host_c64(v) -> np.ndarray:
    a = v.arr if isinstance(v, Vector) else v
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex64))


def _edge_pad(a: np.ndarray, pad: int) -> np.ndarray:
    """Pad.h:77-96 with extend=true along x."""
    if pad == 0:
        return a
    return np.ascontiguousarray(np.pad(a, [(0, 0)] * (a.ndim - 1) + [(pad, pad)], mode="edge"))


def _zero_pad_dev(t: torch.Tensor, pad: int) -> torch.Tensor:
    """Pad.h:77-96 with extend=false along x (device copy)."""
    if pad == 0:
        return t.contiguous()
    out = torch.zeros(t.shape[:-1] + (t.shape[-1] + 2 * pad,), dtype=t.dtype, device=t.device)
    out[..., pad:pad + t.shape[-1]] = t
    return out


class _Acquisition:
    """What PropParam/Acquisition derive from (wave, model hypercube, par)."""

    def __init__(self, wave: Vector, hyper: Hypercube, par: Mapping):
        self.par = dict(par)
        ax_x, ax_z = hyper.getAxis(1), hyper.getAxis(2)
        self.nx, self.nz = ax_x.n, ax_z.n
        self.dx, self.dz = ax_x.d, ax_z.d
        wa = wave.getHyper().getAxis(1)
        self.nt = wa.n
        self.tpad = int(self.par.get("tpad", 0))
        vals = wave.arr.detach().cpu().numpy() if isinstance(wave.arr, torch.Tensor) else wave.arr
        self.freq, self.index, self.wave_f = prop_param(
            vals, wa.d, ot=wa.o, tpad=self.tpad, fmin=float(self.par.get("fmin", 1.0)),
            fmax=float(self.par.get("fmax", 0.0)), nfreq=int(self.par.get("nfreq", 0)))
        self.ox, self.oz = ax_x.o, ax_z.o
        self.src_iz, self.src_ix = device_indices(self.par, "s", ax_x.o, ax_z.o, ax_x.d, ax_z.d)
        self.rec_iz, self.rec_ix = device_indices(self.par, "r", ax_x.o, ax_z.o, ax_x.d, ax_z.d)
        self.pad = int(self.par.get("pad", 0))
        self.tap = int(self.par.get("tap", 0))
        self.nref = int(self.par.get("nref", 1))
        self.prop = str(self.par.get("prop", "ssf"))
        self.eps = float(self.par.get("eps", 0.04))
        self.ntaylor = int(self.par.get("ntaylor", 1))
        self.ngs = int(self.par.get("ngs", 0))
        self.ngr = int(self.par.get("ngr", 0))
        self.ng = int(self.par.get("ng", 0))
        self._surveys = {}
        self._cmap: Optional[CMapPlan] = None
        self._cmap_done = False
        self._plans = {}

    # ---------------------------------------------------- conformal maps

    @property
    def cmap(self) -> Optional[CMapPlan]:
        """Per-shot maps of PropParam::prepareConformalMaps (PropParam.h:89-131)
        when par cmap is 1 or 2; used by WEM and BornFull (wem_prop.cpp:20-45,
        born_prop.cpp:28-70), not by the extended operators or BornRefl, whose
        propagation loops take the model grid as it is."""
        if not self._cmap_done:
            self._cmap = cmap_from_par(self.par, self.nz, self.nx, self.oz, self.ox, self.dz, self.dx,
                                       self.src_iz, self.src_ix, self.rec_iz, self.rec_ix)
            self._cmap_done = True
        return self._cmap

    def shot_models(self, slow, den=None) -> List[Model]:
        """One background model per shot on its mapped grid, built as
        wem_prop.cpp:24-40 does: map, edge-pad, Reflect from the padded map,
        then the slowness scaled by the map's Jacobian factors, in order, twice,
        over the first nz*nx samples of the padded array."""
        cm = self.cmap
        s = _host_c64(slow).ravel()
        d = np.ones_like(s) if den is None else _host_c64(den).ravel()
        n = self.nz * self.nx
        out = []
        try:
            for k in range(cm.ns):
                f = cm.fwd[k]
                g = np.where(f >= 0, s[np.maximum(f, 0)], 0).astype(np.complex64).reshape(self.nz, self.nx)
                gd = np.where(f >= 0, d[np.maximum(f, 0)], 0).astype(np.complex64).reshape(self.nz, self.nx)
                sp = _edge_pad(g, self.pad)
                dp = _edge_pad(gd, self.pad)
                refl = sp.copy()
                flat = sp.reshape(-1)
                for _ in range(2):
                    for j in (cm.jac[k, 0], cm.jac[k, 1]):
                        flat.real[:n] *= j
                        flat.imag[:n] *= j
                out.append(Model(sp, float(cm.axes[k, 1]), float(cm.axes[k, 3]), tap=self.tap, pad=self.pad,
                                 nref=self.nref, prop=self.prop, eps=self.eps, ntaylor=self.ntaylor, den=dp,
                                 slow_refl=refl))
        except Exception:
            for m in out:
                m.close()
            raise
        return out

    def map_plan(self, which: str) -> ops.GatherPlan:
        """Device gather of map_csr(which) (cached)."""
        if which not in self._plans:
            ptr, src = map_csr(self.cmap, self.nz, self.nx, self.pad, which)
            self._plans[which] = ops.GatherPlan(ptr, src)
        return self._plans[which]

    @property
    def nw(self) -> int:
        return len(self.freq)

    def survey(self, nbins: int, refl: bool = False, mapped: bool = False) -> ops.Survey:
        key = (nbins, refl, mapped)
        if key not in self._surveys:
            if int(self.index.max()) >= nbins:
                raise ValueError(f"data has {nbins} frequency bins; the survey needs bin {int(self.index.max())}")
            if refl:   # BornRefl: source spread ng without offset, point receivers (ext_born_refl.cpp:25-26)
                sv = ops.Survey(self.src_iz, self.src_ix, self.rec_iz, self.rec_ix, self.freq, self.index,
                                self.wave_f[self.index], nbins, ngs=self.ng, ngr=0, offset=0)
            else:      # wem_prop.cpp:43-44, born_prop.cpp:63-64
                sv = ops.Survey(self.src_iz, self.src_ix, self.rec_iz, self.rec_ix, self.freq, self.index,
                                self.wave_f[self.index], nbins, ngs=self.ngs, ngr=self.ngr, offset=self.pad)
                if mapped:   # the mapped devices of every shot (wem_prop.cpp:21-22, 43-44)
                    cm = self.cmap
                    for k in range(cm.ns):
                        sv.set_shot_devices(k, cm.src[k][0], cm.src[k][1], cm.rec_iz[k], cm.rec_ix[k])
            self._surveys[key] = sv
        return self._surveys[key]

    def model(self, slow, den=None, padded: bool = True) -> Model:
        s = _host_c64(slow)
        d = None if den is None else _host_c64(den)
        if padded:   # Pad(modspace, pad, pad, true) of the background (PropParam.h:29-30)
            s = _edge_pad(s, self.pad)
            d = None if d is None else _edge_pad(d, self.pad)
        return Model(s, self.dx, self.dz, tap=self.tap, pad=self.pad, nref=self.nref, prop=self.prop,
                     eps=self.eps, ntaylor=self.ntaylor, den=d)

    # ---------------------------------------------------- the data boundary

    def _is_time(self, data: Vector) -> bool:
        return not data.arr.is_complex()

    def freq_data(self, data: Vector) -> torch.Tensor:
        """Complex frequency data to propagate into: the vector itself for complex
        data, a zeroed [ns][nr][nt+tpad] buffer for float traces."""
        if not self._is_time(data):
            return data.arr
        ns, nr, nt = data.arr.shape
        return torch.zeros((ns, nr, nt + self.tpad), dtype=torch.complex64, device=data.arr.device)

    def to_time(self, f: torch.Tensor, data: Vector, add: bool, stream=None):
        """FFT1::adjoint + Pad::adjoint of the frequency data into float traces
        (WEMnl.cpp:61-63)."""
        ns, nr, nt = data.arr.shape
        b = np.ascontiguousarray(self.index, dtype=np.int32)
        check(lib.wemfwi_freq_to_trace(ns * nr, nt, self.tpad, C.c_void_p(f.data_ptr()), f.shape[-1], len(b),
                                       b.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(data.arr.data_ptr()),
                                       int(bool(add)), _stream_ptr(stream)))

    def to_freq(self, data: Vector, stream=None) -> torch.Tensor:
        """Pad::forward + FFT1::forward of float traces (BornFull.cpp:80-88); complex
        data pass through."""
        if not self._is_time(data):
            return data.arr
        ns, nr, nt = data.arr.shape
        f = torch.empty((ns, nr, nt + self.tpad), dtype=torch.complex64, device=data.arr.device)
        b = np.ascontiguousarray(self.index, dtype=np.int32)
        check(lib.wemfwi_trace_to_freq(ns * nr, nt, self.tpad, C.c_void_p(data.arr.contiguous().data_ptr()),
                                       len(b), b.ctypes.data_as(C.POINTER(C.c_int)), f.shape[-1],
                                       C.c_void_p(f.data_ptr()), 0, _stream_ptr(stream)))
        return f


def map_csr(cm: CMapPlan, nz: int, nx: int, pad: int, which: str):
    """CSR lists (ptr, src) of the gathers between the model grid [nz][nx]
    and the per-shot padded grids [ns][nz][nx + 2 pad]:
      'fwd'   a perturbation mapped onto every shot's grid, zero-padded
              (born_prop.cpp:50-58): out = ns*nz*(nx+2 pad) samples;
      'inv'   the reference's inverse map of the per-shot results summed over
              shots in shot order (born_prop.cpp:217-231): out = nz*nx;
      'exact' the transpose of 'fwd' (the adjoint of the mapping).
    A sample with an empty list is zero."""
    nxp = nx + 2 * pad
    n = nz * nx
    i = np.arange(n)
    pos = (i // nx) * nxp + pad + (i % nx)      # model-grid sample -> padded slice position
    if which == "fwd":
        idx = np.full((cm.ns, nz * nxp), -1, np.int64)
        idx[:, pos] = cm.fwd
        idx = idx.ravel()
        ok = idx >= 0
        return np.concatenate([[0], np.cumsum(ok)]), idx[ok]
    if which == "inv":
        v = cm.inv.astype(np.int64)                                          # [ns][n]
        src = np.where(v >= 0, np.arange(cm.ns)[:, None] * nz * nxp + pos[np.maximum(v, 0)], -1).T
        cnt = (src >= 0).sum(axis=1)                                         # [n], shots in order
        return np.concatenate([[0], np.cumsum(cnt)]), src[src >= 0]
    if which == "exact":
        f = cm.fwd.astype(np.int64)
        k, ii = np.nonzero(f >= 0)
        tgt = f[k, ii]
        order = np.lexsort((ii, k, tgt))
        cnt = np.bincount(tgt, minlength=n)
        return np.concatenate([[0], np.cumsum(cnt)]), (k * nz * nxp + pos[ii])[order]
    raise ValueError(which)


def _model_vecs(model) -> List[Vector]:
    if isinstance(model, SuperVector):
        return model.vecs
    if isinstance(model, (list, tuple)):
        return list(model)
    return [model]


def _par_dict(par) -> dict:
    """A paramObj-like argument as a dict (a Mapping, or an object holding
    one in .pars as genericIO.pythonParams does)."""
    return dict(getattr(par, "pars", par))


class _Picklable:
    """Pickling of an operator (WEM.py:32-35, 45-47, 77-80): the native
    objects (the acquisition's surveys and device plans, the background
    models) are dropped and rebuilt on unpickling from the wave, the
    parameters and a host copy of the background, so an unpickled operator
    computes what the pickled one did."""

    _native = ("acq", "_bg")

    def __getstate__(self):
        state = self.__dict__.copy()
        for k in self._native:
            state.pop(k, None)
        return state

    def __setstate__(self, state):
        self.__dict__.update(state)
        self.acq = _Acquisition(self.wave, self._hyper, self.par)
        self._bg = None
        if getattr(self, "_bg_host", None) is not None:
            self._load_bg(self._bg_host)

    def get_nfreq(self) -> int:
        """Number of propagated frequencies (PropParam.h:42, pyWEM.cpp:36-39)."""
        return self.acq.nw


class Propagator(_Picklable, Operator):
    """Base of the shot x frequency operators (WEM.py:8-41)."""

    def __init__(self, model, data, wave, par):
        if not isinstance(model, SuperVector):
            raise TypeError("Model vector should be a superVector")
        super().__init__(model, data)
        self.wave = wave
        self.par = _par_dict(par)
        self._hyper = _model_vecs(model)[0].getHyper()
        self.acq = _Acquisition(wave, self._hyper, self.par)

    @classmethod
    def from_subspace(cls, subdomain, subrange, wave, par, *args, **kw):
        """The operator of a subset of the shots (WEM.py:24-30): par's ns and
        osx are taken from the third (shot) axis of subrange."""
        pars = _par_dict(par)
        ax = subrange.getHyper().getAxis(3)
        pars["ns"] = ax.n
        pars["osx"] = ax.o
        return cls(subdomain, subrange, wave, pars, *args, **kw)


class WEM(Propagator):
    """Non-linear modeling d = WEM([S, rho]) (WEMnl.cpp:39-82, wem_prop.cpp:9-78).
    Only forward exists, as in the reference."""

    def forward(self, add, model, data):
        vecs = _model_vecs(model)
        den = vecs[1] if len(vecs) > 1 else None
        if self.acq.cmap is not None:   # one mapped grid per shot
            ms = self.acq.shot_models(vecs[0], den)
            try:
                f = self.acq.freq_data(data)
                sv = self.acq.survey(f.shape[-1], mapped=True)
                if self.acq._is_time(data):
                    ops.shotwem_forward(ms, sv, f, add=False)
                    self.acq.to_time(f, data, add)
                else:
                    ops.shotwem_forward(ms, sv, f, add=bool(add))
            finally:
                for m in ms:
                    m.close()
            return
        m = self.acq.model(vecs[0], den)
        try:
            f = self.acq.freq_data(data)
            sv = self.acq.survey(f.shape[-1])
            if self.acq._is_time(data):
                ops.wem_forward(m, sv, f, add=False)
                self.acq.to_time(f, data, add)
            else:
                ops.wem_forward(m, sv, f, add=bool(add))
        finally:
            m.close()

    def adjoint(self, add, model, data):
        raise NotImplementedError("WEM is non-linear: use Born for the adjoint")


def _with_checkpoint(sv: ops.Survey, every: int) -> ops.Survey:
    if sv.checkpoint != every:
        sv.set_checkpoint(every)
    return sv


class Born(Propagator):
    """Linearised modeling around a background (BornFull.cpp:13-117), model
    [dS, drho] on the unpadded grid.  mode='real' zeroes imaginary parts of the
    model before forward and after adjoint (WEM.py:52-67).  adjoint_mode
    'compat' keeps Scatter::adjoint's operator order (Scatter.cpp:52-66);
    'exact' is the exact adjoint.  checkpoint=k > 0 makes the adjoint keep
    wavefield rows every k depth steps and recompute the rest (bitwise the
    same result, far less device memory per row)."""

    def __init__(self, model, data, wave, par, mode=None, adjoint_mode: str = "compat", checkpoint: int = 0):
        super().__init__(model, data, wave, par)
        self.mode = mode
        if adjoint_mode not in ("compat", "exact"):
            raise ValueError("adjoint_mode must be 'compat' or 'exact'")
        self.adjoint_mode = adjoint_mode
        self.checkpoint = int(checkpoint)
        self._bg = None
        self.set_background(model)

    def set_background(self, slow):
        vecs = _model_vecs(slow)
        self._load_bg([_host_c64(v) for v in vecs[:2]])

    def _load_bg(self, host):
        den = host[1] if len(host) > 1 else None
        self._close_bg()
        if self.acq.cmap is not None:
            self._bg = self.acq.shot_models(host[0], den)
        else:
            self._bg = self.acq.model(host[0], den)
        self._bg_host = host

    def _close_bg(self):
        for m in (self._bg if isinstance(self._bg, list) else [self._bg]):
            if m is not None:
                m.close()
        self._bg = None

    def _mapped_forward(self, add, vecs, data):
        acq = self.acq
        ns, nz, nxp = acq.cmap.ns, acq.nz, acq.nx + 2 * acq.pad
        plan = acq.map_plan("fwd")
        dS = plan(vecs[0].arr.contiguous(), torch.empty((ns, nz, nxp), dtype=torch.complex64, device=vecs[0].arr.device))
        dR = torch.zeros_like(dS)
        if len(vecs) > 1:
            plan(vecs[1].arr.contiguous(), dR)
        f = acq.freq_data(data)
        sv = acq.survey(f.shape[-1], mapped=True)
        if acq._is_time(data):
            ops.shotborn_forward(self._bg, sv, dS, dR, f, add=False)
            acq.to_time(f, data, add)
        else:
            ops.shotborn_forward(self._bg, sv, dS, dR, f, add=bool(add))

    def _mapped_adjoint(self, add, vecs, data):
        acq = self.acq
        f = acq.to_freq(data)
        sv = _with_checkpoint(acq.survey(f.shape[-1], mapped=True), self.checkpoint)
        ns, nz, nxp = acq.cmap.ns, acq.nz, acq.nx + 2 * acq.pad
        gS = torch.zeros((ns, nz, nxp), dtype=torch.complex64, device=f.device)
        gR = torch.zeros_like(gS)
        exact = self.adjoint_mode == "exact"
        ops.shotborn_adjoint(self._bg, sv, gS, gR, f, add=False, exact=exact)
        plan = acq.map_plan("exact" if exact else "inv")
        for v, g in zip(vecs, (gS, gR)):
            if v.arr.is_contiguous():
                plan(g, v.arr, add=bool(add))
            else:
                t = plan(g, torch.empty(v.arr.shape, dtype=torch.complex64, device=g.device))
                v.arr.add_(t) if add else v.arr.copy_(t)

    def forward(self, add, model, data):
        vecs = _model_vecs(model)
        if self.mode == "real":
            for v in vecs:
                v.arr.imag.zero_()
        if self.acq.cmap is not None:
            self._mapped_forward(add, vecs, data)
            return
        p = self.acq.pad
        dS = _zero_pad_dev(vecs[0].arr, p)
        dR = _zero_pad_dev(vecs[1].arr, p) if len(vecs) > 1 else torch.zeros_like(dS)
        f = self.acq.freq_data(data)
        sv = self.acq.survey(f.shape[-1])
        if self.acq._is_time(data):
            ops.born_forward(self._bg, sv, dS, dR, f, add=False)
            self.acq.to_time(f, data, add)
        else:
            ops.born_forward(self._bg, sv, dS, dR, f, add=bool(add))

    def adjoint(self, add, model, data):
        vecs = _model_vecs(model)
        if self.acq.cmap is not None:
            self._mapped_adjoint(add, vecs, data)
            if self.mode == "real":
                for v in vecs:
                    v.arr.imag.zero_()
            return
        f = self.acq.to_freq(data)
        sv = _with_checkpoint(self.acq.survey(f.shape[-1]), self.checkpoint)
        p = self.acq.pad
        nz, nxp = self._bg.nz, self._bg.nx
        gS = torch.zeros((nz, nxp), dtype=torch.complex64, device=f.device)
        gR = torch.zeros_like(gS)
        ops.born_adjoint(self._bg, sv, gS, gR, f, add=False, exact=self.adjoint_mode == "exact")
        for v, g in zip(vecs, (gS, gR)):   # Pad::adjoint = truncation
            g = g[:, p:nxp - p]
            if add:
                v.arr.add_(g)
            else:
                v.arr.copy_(g)
        if self.mode == "real":
            for v in vecs:
                v.arr.imag.zero_()


class BornRefl(_Picklable, Operator):
    """Frequency-extended reflectivity modeling / migration
    (BornRefl.h:9-58, ext_born_refl.cpp:8-110): model is an image
    [nw][nz][nx] on the unpadded grid, data float traces (the reference's
    interface) or complex frequency data.  add_prec / add_pad register
    operators that set_background applies to the slowness first
    (WEM.py:101-115)."""

    def __init__(self, model, slowness, data, wave, par):
        super().__init__(model, data)
        self.wave = wave
        self.par = _par_dict(par)
        self._hyper = slowness.getHyper()
        self.acq = _Acquisition(wave, self._hyper, self.par)
        self._bg = None
        self.prec = None
        self.pad = None
        self.set_background(slowness)

    def add_prec(self, op):
        self.prec = op
        self.big_model = op.range.clone()

    def add_pad(self, op):
        self.pad = op
        self.pad_model = op.range.clone()

    def set_background(self, slow):
        if self.prec is not None:
            self.prec.forward(False, slow, self.big_model)
            slow = self.big_model
        if self.pad is not None:
            self.pad.forward(False, slow, self.pad_model)
            slow = self.pad_model
        self._load_bg([_host_c64(slow)])

    def _load_bg(self, host):
        if self._bg is not None:
            self._bg.close()
            self._bg = None
        self._bg = self.acq.model(host[0], None, padded=False)
        self._bg_host = host

    def forward(self, add, model, data):
        f = self.acq.freq_data(data)
        sv = self.acq.survey(f.shape[-1], refl=True)
        if self.acq._is_time(data):
            ops.bornrefl_forward(self._bg, sv, model.arr, f, add=False)
            self.acq.to_time(f, data, add)
        else:
            ops.bornrefl_forward(self._bg, sv, model.arr, f, add=bool(add))

    def adjoint(self, add, model, data):
        f = self.acq.to_freq(data)
        sv = self.acq.survey(f.shape[-1], refl=True)
        ops.bornrefl_adjoint(self._bg, sv, model.arr, f, add=bool(add))


class extWEM(Propagator):  # noqa: N801 - reference name
    """WEM with one [S, rho] slice per frequency (extWEM.cpp:54-70); model is a
    SuperVector of two [nw][nz][nx] vectors."""

    def forward(self, add, model, data):
        vecs = _model_vecs(model)
        S = _host_c64(vecs[0])
        R = _host_c64(vecs[1]) if len(vecs) > 1 else None
        models = [self.acq.model(S[i], None if R is None else R[i]) for i in range(S.shape[0])]
        try:
            f = self.acq.freq_data(data)
            sv = self.acq.survey(f.shape[-1])
            if self.acq._is_time(data):
                ops.extwem_forward(models, sv, f, add=False)
                self.acq.to_time(f, data, add)
            else:
                ops.extwem_forward(models, sv, f, add=bool(add))
        finally:
            for m in models:
                m.close()


class extBorn(Propagator):  # noqa: N801 - reference name
    """BornFull with one background and one perturbation slice per frequency
    (extBornFull.cpp:68-121).  checkpoint as for Born."""

    def __init__(self, model, data, wave, par, adjoint_mode: str = "compat", checkpoint: int = 0):
        super().__init__(model, data, wave, par)
        self.adjoint_mode = adjoint_mode
        self.checkpoint = int(checkpoint)
        self._bg: List[Model] = []
        self.set_background(model)

    def set_background(self, slow):
        vecs = _model_vecs(slow)
        self._load_bg([_host_c64(v) for v in vecs[:2]])

    def _load_bg(self, host):
        S = host[0]
        R = host[1] if len(host) > 1 else None
        for m in self._bg or []:
            m.close()
        self._bg = [self.acq.model(S[i], None if R is None else R[i]) for i in range(S.shape[0])]
        self._bg_host = host

    def forward(self, add, model, data):
        vecs = _model_vecs(model)
        p = self.acq.pad
        dS = _zero_pad_dev(vecs[0].arr, p)
        dR = _zero_pad_dev(vecs[1].arr, p) if len(vecs) > 1 else torch.zeros_like(dS)
        f = self.acq.freq_data(data)
        sv = self.acq.survey(f.shape[-1])
        if self.acq._is_time(data):
            ops.extborn_forward(self._bg, sv, dS, dR, f, add=False)
            self.acq.to_time(f, data, add)
        else:
            ops.extborn_forward(self._bg, sv, dS, dR, f, add=bool(add))

    def adjoint(self, add, model, data):
        vecs = _model_vecs(model)
        f = self.acq.to_freq(data)
        sv = _with_checkpoint(self.acq.survey(f.shape[-1]), self.checkpoint)
        p = self.acq.pad
        nw = len(self._bg)
        nz, nxp = self._bg[0].nz, self._bg[0].nx
        gS = torch.zeros((nw, nz, nxp), dtype=torch.complex64, device=f.device)
        gR = torch.zeros_like(gS)
        ops.extborn_adjoint(self._bg, sv, gS, gR, f, add=False, exact=self.adjoint_mode == "exact")
        for v, g in zip(vecs, (gS, gR)):
            g = g[..., p:nxp - p]
            if add:
                v.arr.add_(g)
            else:
                v.arr.copy_(g)
