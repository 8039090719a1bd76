"""Acquisition parameters of the reference's operators, computed by the
library's host code: the frequency grid + wavelet spectrum of
PropParam::calcPropParam (propagator/src/PropParam.h:45-85), the regular and
irregular device geometry of Device (propagator/src/Device.h:39-85) and the
per-shot conformal maps of PropParam::prepareConformalMaps
(PropParam.h:89-131).

Parameter names are the reference's paramObj keys (fmin, fmax, nfreq, tpad,
ns/osx/osz/dsx/dsz, nr/orx/orz/drx/drz, coord, cmap, eps_log, cmap_angles).
"""

from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from typing import List, Mapping, Optional, Sequence, Tuple

import numpy as np

from ._lib import check, lib


def prop_param(wave, dt: float, ot: float = 0.0, tpad: int = 0, fmin: float = 1.0,
               fmax: float = 0.0, nfreq: int = 0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(freq [nw] float32 Hz, index [nw] int32 data bins, wave_f [nt+tpad] complex64)
    exactly as PropParam::calcPropParam derives them from a float wavelet."""
    w = np.ascontiguousarray(np.asarray(wave, dtype=np.float32))
    nt = int(w.size)
    nw = C.c_int()
    check(lib.wemfwi_prop_param(nt, float(ot), float(dt), None, int(tpad), float(fmin), float(fmax),
                                int(nfreq), 0, C.byref(nw), None, None, None))
    n = nw.value
    freq = np.zeros(n, dtype=np.float32)
    index = np.zeros(n, dtype=np.int32)
    wave_f = np.zeros(nt + int(tpad), dtype=np.complex64)
    check(lib.wemfwi_prop_param(nt, float(ot), float(dt), w.ctypes.data_as(C.POINTER(C.c_float)),
                                int(tpad), float(fmin), float(fmax), int(nfreq), n, C.byref(nw),
                                freq.ctypes.data_as(C.POINTER(C.c_float)),
                                index.ctypes.data_as(C.POINTER(C.c_int)),
                                wave_f.ctypes.data_as(C.c_void_p)))
    return freq, index, wave_f


def device_indices(par: Mapping, name: str, ox: float, oz: float, dx: float,
                   dz: float) -> Tuple[np.ndarray, np.ndarray]:
    """(iz, ix) of the devices named 's' (sources) or 'r' (receivers) from the
    parameters n<name>, o<name>x, o<name>z, d<name>x, d<name>z on a model grid
    with origin (ox, oz) and spacing (dx, dz)."""
    if name not in ("s", "r"):
        raise ValueError("Device must be specified to be r or s!")
    coord = str(par.get("coord", "") or "")
    if coord:   # Device::createIrregGeom (Device.h:65-85)
        with open(coord) as f:
            js = json.load(f)
        xs, zs = js.get(name + "x", []), js.get(name + "z", [])
        if len(xs) != len(zs):
            raise ValueError("z- and x-coordinates not equal size!")
        x = np.ascontiguousarray(np.asarray(xs, dtype=np.float32))
        z = np.ascontiguousarray(np.asarray(zs, dtype=np.float32))
        iz = np.zeros(x.size, dtype=np.int32)
        ix = np.zeros(x.size, dtype=np.int32)
        check(lib.wemfwi_device_coords(int(x.size), x.ctypes.data_as(C.POINTER(C.c_float)),
                                       z.ctypes.data_as(C.POINTER(C.c_float)), float(ox), float(oz),
                                       float(dx), float(dz), iz.ctypes.data_as(C.POINTER(C.c_int)),
                                       ix.ctypes.data_as(C.POINTER(C.c_int))))
        return iz, ix
    n = int(par.get("n" + name, 0))
    iz = np.zeros(max(n, 0), dtype=np.int32)
    ix = np.zeros(max(n, 0), dtype=np.int32)
    check(lib.wemfwi_device_indices(n, float(par.get(f"o{name}x", 0.0)), float(par.get(f"o{name}z", 0.0)),
                                    float(par.get(f"d{name}x", 0.0)), float(par.get(f"d{name}z", 0.0)),
                                    float(ox), float(oz), float(dx), float(dz),
                                    iz.ctypes.data_as(C.POINTER(C.c_int)),
                                    ix.ctypes.data_as(C.POINTER(C.c_int))))
    return iz, ix


@dataclass
class CMapPlan:
    """Per-shot conformal-map tables (see cmap_plan)."""

    kind: int
    axes: np.ndarray          # [ns][4] o_x, d_x, o_z, d_z of each shot's propagation grid
    fwd: np.ndarray           # [ns][nz*nx] model -> grid gather (-1: zero)
    inv: np.ndarray           # [ns][nz*nx] grid -> model gather of the Born adjoint (-1: zero)
    jac: np.ndarray           # [ns][2][nz*nx] Jacobian factors, applied in order, twice
    src: List[Tuple[np.ndarray, np.ndarray]]   # mapped source points (iz, ix) per shot
    rec_iz: np.ndarray        # [ns][nr]
    rec_ix: np.ndarray        # [ns][nr]

    @property
    def ns(self) -> int:
        return self.axes.shape[0]


def cmap_plan(kind: int, nz: int, nx: int, oz: float, ox: float, dz: float, dx: float,
              src_iz: Sequence[int], src_ix: Sequence[int], rec_iz: Sequence[int], rec_ix: Sequence[int],
              eps_log: float = 0.01, angles: Optional[Sequence[float]] = None) -> CMapPlan:
    """Tables of PropParam::prepareConformalMaps for every shot: kind 1 =
    ShiftAndScale + ComplexLog(eps_log), kind 2 = Rotation(angles[s]), 0 =
    identity."""
    siz = np.ascontiguousarray(np.asarray(src_iz, dtype=np.int32))
    six = np.ascontiguousarray(np.asarray(src_ix, dtype=np.int32))
    riz = np.ascontiguousarray(np.asarray(rec_iz, dtype=np.int32))
    rix = np.ascontiguousarray(np.asarray(rec_ix, dtype=np.int32))
    ns, nr, n = siz.size, riz.size, nz * nx
    ang = None if angles is None else np.ascontiguousarray(np.asarray(angles, dtype=np.float32))
    if kind == 2 and (ang is None or ang.size < ns):
        raise ValueError("cmap 2 needs one angle per shot")
    axes = np.zeros((ns, 4), np.float32)
    fwd = np.zeros((ns, n), np.int32)
    inv = np.zeros((ns, n), np.int32)
    jac = np.zeros((ns, 2, n), np.float32)
    nsrc = np.zeros(ns, np.int32)
    oiz = np.zeros((ns, nx), np.int32)
    oix = np.zeros((ns, nx), np.int32)
    orz = np.zeros((ns, nr), np.int32)
    orx = np.zeros((ns, nr), np.int32)
    ip = C.POINTER(C.c_int)
    fp = C.POINTER(C.c_float)
    check(lib.wemfwi_cmap_plan(int(kind), int(nz), int(nx), float(oz), float(ox), float(dz), float(dx), ns,
                               siz.ctypes.data_as(ip), six.ctypes.data_as(ip), nr, riz.ctypes.data_as(ip),
                               rix.ctypes.data_as(ip), float(eps_log),
                               None if ang is None else ang.ctypes.data_as(fp), axes.ctypes.data_as(fp),
                               fwd.ctypes.data_as(ip), inv.ctypes.data_as(ip), jac.ctypes.data_as(fp),
                               nsrc.ctypes.data_as(ip), oiz.ctypes.data_as(ip), oix.ctypes.data_as(ip),
                               orz.ctypes.data_as(ip), orx.ctypes.data_as(ip)))
    src = [(oiz[s, :nsrc[s]].copy(), oix[s, :nsrc[s]].copy()) for s in range(ns)]
    return CMapPlan(int(kind), axes, fwd, inv, jac, src, orz, orx)


def cmap_from_par(par: Mapping, nz: int, nx: int, oz: float, ox: float, dz: float, dx: float,
                  src_iz, src_ix, rec_iz, rec_ix) -> Optional[CMapPlan]:
    """The maps the parameters ask for (cmap=1: eps_log, default 0.01;
    cmap=2: the "a" array of the JSON file cmap_angles), or None for cmap=0."""
    kind = int(par.get("cmap", 0))
    if kind == 0:
        return None
    angles = None
    if kind == 2:
        with open(str(par["cmap_angles"])) as f:
            angles = [float(a) for a in json.load(f)["a"]]
    return cmap_plan(kind, nz, nx, oz, ox, dz, dx, src_iz, src_ix, rec_iz, rec_ix,
                     eps_log=float(par.get("eps_log", 0.01)), angles=angles)
