"""Oracle of the acquisition-geometry tables: irregular devices and the
per-shot conformal maps.

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module; the
product computes the same tables in its own host code (``wemfwi_cmap_plan``,
``wemfwi_device_coords``) and never calls this file.

NumPy restatement, in float32 / complex64 arithmetic (the reference's types),
of (paths relative to the reference repository):

  ======================  ====================================================
  function here           reference it follows
  ======================  ====================================================
  device_coords           propagator/src/Device.h:65-85
  grid_coords             conformal/src/ConformalMap.h:98-109
  shift_scale             conformal/src/ShiftAndScale.h:15-55
  complex_log             conformal/src/ComplexLog.h:14-66
  rotation                conformal/src/Rotation.h:14-99
  forward_gather          conformal/src/ConformalMap.cpp:53-72
  log_forward_gather      conformal/src/ComplexLog.cpp:77-109
  inverse_gather          conformal/src/ConformalMap.cpp:74-87
  log_inverse_gather      conformal/src/ComplexLog.cpp:111-130
  map_point               conformal/src/ConformalMap.h:34-66,
                          ComplexLog.cpp:132-190
  cmap_tables             propagator/src/PropParam.h:89-131 with
                          conformal/src/MapChain.h:17-60
  ======================  ====================================================

Deviation kept on purpose: samples a map leaves unwritten are zero.  The
reference writes into per-task buffers reused from shot to shot (the grid of
wem_prop.cpp:15-17 for Rotation, the ``mm`` of born_prop.cpp:124 for the
inverse), so there they keep a previous shot's values, depending on the TBB
partition.

numpy's complex64 log, exp and division may differ from the C library's in
the last bit; tables are compared with the product except where a mapped
coordinate lies within a tiny distance of a cell boundary (see ``ties``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Sequence, Tuple

import numpy as np

F32 = np.float32
C64 = np.complex64
PI = F32(4) * np.arctan(F32(1))


def device_coords(x: Sequence[float], z: Sequence[float], ox: float, oz: float, dxv: float, dzv: float):
    """Device.h:72-82: ix = (int)((x - ox)/dxv + .5f) in float32."""
    x = np.asarray(x, dtype=F32)
    z = np.asarray(z, dtype=F32)
    fx = (x - F32(ox)) / F32(dxv) + F32(0.5)
    fz = (z - F32(oz)) / F32(dzv) + F32(0.5)
    ix = np.trunc(fx).astype(np.int64)
    iz = np.trunc(fz).astype(np.int64)
    if np.any(ix < 0) or np.any(iz < 0):
        raise ValueError("device out of bounds")
    return iz, ix


@dataclass
class Grid:
    n1: int
    n2: int
    o1: np.float32
    d1: np.float32
    o2: np.float32
    d2: np.float32

    def coords(self) -> np.ndarray:
        """z + i x of every sample, row-major (axis 1 fastest)."""
        i2, i1 = np.meshgrid(np.arange(self.n2), np.arange(self.n1), indexing="ij")
        z = F32(self.o2) + i2.astype(F32) * F32(self.d2)
        x = F32(self.o1) + i1.astype(F32) * F32(self.d1)
        return (z + 1j * x).astype(C64).ravel()

    def frac(self, c: np.ndarray):
        """Unrounded cell coordinates (z, x) of coordinates c."""
        fz = (c.real.astype(F32) - F32(self.o2)) / F32(self.d2)
        fx = (c.imag.astype(F32) - F32(self.o1)) / F32(self.d1)
        return fz, fx

    def cells(self, c: np.ndarray):
        fz, fx = self.frac(c)
        iz = np.trunc(fz).astype(np.int64)
        ix = np.trunc(fx).astype(np.int64)
        ok = (iz >= 0) & (iz < self.n2) & (ix >= 0) & (ix < self.n1)
        return iz, ix, ok


@dataclass
class Map:
    inp: Grid
    out: Grid
    zmap: np.ndarray
    umap: np.ndarray
    jac: np.ndarray


def _cdiv(x, y):
    """Complex division in the C library's order (Smith's method with true
    divisions, as libgcc's __divsc3 evaluates it); numpy's complex64 division
    multiplies by a reciprocal instead, which rounds differently."""
    x = np.asarray(x, C64)
    y = C64(y)
    c, d = F32(y.real), F32(y.imag)
    a, b = x.real.astype(F32), x.imag.astype(F32)
    if abs(c) >= abs(d):
        r = d / c
        den = c + d * r
        re, im = (a + b * r) / den, (b - a * r) / den
    else:
        r = c / d
        den = c * r + d
        re, im = (a * r + b) / den, (b * r - a) / den
    return (re.astype(F32) + 1j * im.astype(F32)).astype(C64)


def shift_scale(g: Grid, a: complex, b: complex) -> Map:
    """f(z) = a (z - b)."""
    a = C64(a)
    b = C64(b)
    lo = C64(complex(g.o2, g.o1))
    hi = C64(complex(F32(g.o2) + F32(g.n2 - 1) * F32(g.d2), F32(g.o1) + F32(g.n1 - 1) * F32(g.d1)))
    lo = C64(a * C64(lo - b))
    hi = C64(a * C64(hi - b))
    span = C64(hi - lo)
    out = Grid(g.n1, g.n2, F32(lo.imag), F32(span.imag) / F32(g.n1 - 1), F32(lo.real), F32(span.real) / F32(g.n2 - 1))
    zmap = (_cdiv(out.coords(), a) + b).astype(C64)
    umap = (a * (g.coords() - b)).astype(C64)
    jac = np.full(zmap.size, F32(abs(_cdiv(C64(1), a)[()])), F32)
    return Map(g, out, zmap, umap, jac)


def complex_log(g: Grid, eps: float) -> Map:
    """u = log z; |z| < e^eps clamps to the circle; angles in [0, 2 pi)."""
    e = F32(np.log(F32(eps)))
    out = Grid(g.n1, g.n2, F32(0), F32(2) * PI / F32(g.n1 - 1), e, -e / F32(g.n2 - 1))
    u = out.coords()
    zmap = np.exp(u).astype(C64)
    z = g.coords()
    rmin = F32(np.exp(e))
    small = np.abs(z) < rmin
    v = np.where(small, (e + 1j * np.angle(z).astype(F32)).astype(C64), np.log(np.where(small, 1, z)).astype(C64))
    v = np.where(v.imag < 0, v + C64(2j * PI), v).astype(C64)
    jac = np.exp(u.real.astype(F32)).astype(F32)
    return Map(g, out, zmap, v, jac)


def rotation(g: Grid, rad: float) -> Map:
    """Rotation by -rad with Rotation.h's clamping of outside preimages."""
    rad = F32(rad)
    rot = C64(np.exp(C64(-1j) * rad))
    irot = C64(np.exp(C64(1j) * rad))
    xe = F32(g.o1) + F32(g.n1 - 1) * F32(g.d1)
    ze = F32(g.o2) + F32(g.n2 - 1) * F32(g.d2)
    v = np.array([complex(g.o2, g.o1), complex(g.o2, xe), complex(ze, g.o1), complex(ze, xe)], C64)
    u = (v * rot).astype(C64)
    o1 = min(F32(0), *u.imag.astype(F32))
    o2 = min(F32(0), *u.real.astype(F32))
    m1_ = max(F32(0), *u.imag.astype(F32))
    m2_ = max(F32(0), *u.real.astype(F32))
    out = Grid(g.n1, g.n2, F32(o1), (F32(m1_) - F32(o1)) / F32(g.n1 - 1), F32(o2), (F32(m2_) - F32(o2)) / F32(g.n2 - 1))
    m1 = xe
    m2 = ze
    t = F32(np.tan(rad))
    zm = np.empty(g.n1 * g.n2, C64)
    oc = out.coords()
    for i in range(oc.size):
        z = C64(oc[i] * irot)
        zr, zi = F32(z.real), F32(z.imag)
        i1 = int(np.trunc((zi - F32(g.o1)) / F32(g.d1)))
        i2 = int(np.trunc((zr - F32(g.o2)) / F32(g.d2)))
        if i1 < 0:
            x = -(zi - F32(g.o1)) / t
            x = min(F32(zr + x), m2)
            zr, zi = F32(x), F32(g.o1)
        if i2 < 0:
            x = -(zr - F32(g.o2)) * t
            x = min(F32(zi + x), m1)
            zr, zi = F32(g.o2), F32(x)
        if i1 >= g.n1:
            x = (zi - m1) / t
            x = max(F32(zr - x), F32(g.o1))
            zr, zi = F32(x), m1
        if i2 >= g.n2:
            x = (zr - m2) * t
            x = min(F32(zi - x), F32(g.o1))
            zr, zi = m2, F32(x)
        zm[i] = C64(complex(zr, zi))
    umap = (g.coords() * rot).astype(C64)
    return Map(g, out, zm, umap, np.ones(zm.size, F32))


def forward_gather(m: Map) -> np.ndarray:
    iz, ix, ok = m.inp.cells(m.zmap)
    return np.where(ok, iz * m.inp.n1 + ix, -1)


def log_forward_gather(m: Map) -> np.ndarray:
    """A sample without an in-grid preimage takes the last one of its column."""
    iz, ix, ok = m.inp.cells(m.zmap)
    idx = np.where(ok, iz * m.inp.n1 + ix, -1).reshape(m.inp.n2, m.inp.n1)
    col = np.full(m.inp.n1, -1)
    out = np.empty_like(idx)
    for r in range(m.inp.n2):
        col = np.where(idx[r] >= 0, idx[r], col)
        out[r] = col
    return out.ravel()


def inverse_gather(m: Map) -> np.ndarray:
    iz, ix, ok = m.out.cells(m.umap)
    return np.where(ok, iz * m.out.n1 + ix, -1)


def log_inverse_gather(m: Map, eps: float) -> np.ndarray:
    amp = np.abs(m.inp.coords()).astype(F32)
    rmin = F32(np.exp(F32(np.log(F32(eps)))))
    sel = (amp > rmin) & (amp < F32(0.99))
    iz, ix, ok = m.out.cells(m.umap)
    if np.any(sel & ~ok):
        raise ValueError("inverse map leaves the grid")
    return np.where(sel, iz * m.out.n1 + ix, -1)


def map_point(m: Map, iz: int, ix: int) -> Tuple[int, int]:
    c = m.umap[ix + iz * m.inp.n1]
    fz, fx = m.out.frac(np.asarray([c]))
    return int(np.trunc(fz[0])), int(np.trunc(fx[0]))


@dataclass
class Tables:
    axes: np.ndarray                       # [ns][4] o_x, d_x, o_z, d_z
    fwd: np.ndarray                        # [ns][n]
    inv: np.ndarray                        # [ns][n]
    jac: np.ndarray                        # [ns][2][n]
    src: List[Tuple[np.ndarray, np.ndarray]]
    rec_iz: np.ndarray                     # [ns][nr]
    rec_ix: np.ndarray
    maps: list                             # the Map objects of every shot


def cmap_tables(kind: int, nz: int, nx: int, oz: float, ox: float, dz: float, dx: float, src_iz, src_ix,
                rec_iz, rec_ix, eps_log: float = 0.01, angles=None) -> Tables:
    g = Grid(nx, nz, F32(ox), F32(dx), F32(oz), F32(dz))
    ns, nr, n = len(src_iz), len(rec_iz), nz * nx
    axes = np.zeros((ns, 4), F32)
    fwd = np.zeros((ns, n), np.int64)
    inv = np.zeros((ns, n), np.int64)
    jac = np.ones((ns, 2, n), F32)
    src, maps = [], []
    riz = np.zeros((ns, nr), np.int64)
    rix = np.zeros((ns, nr), np.int64)
    a = C64(0)
    for s in range(ns):
        if kind == 0:
            fwd[s] = np.arange(n)
            inv[s] = np.arange(n)
            out = g
            src.append((np.array([src_iz[s]]), np.array([src_ix[s]])))
            riz[s], rix[s] = rec_iz, rec_ix
            maps.append(None)
        elif kind == 1:
            b = C64(complex(F32(g.o2) + F32(src_iz[s]) * F32(g.d2), F32(g.o1) + F32(src_ix[s]) * F32(g.d1)))
            for i2 in (0, nz - 1):
                for i1 in (0, nx - 1):
                    corner = C64(complex(F32(g.o2) + F32(i2) * F32(g.d2), F32(g.o1) + F32(i1) * F32(g.d1)))
                    a = C64(max(F32(a.real), F32(abs(C64(corner - b)))))
            a = C64(_cdiv(C64(1), a)[()])
            sas = shift_scale(g, a, b)
            lg = complex_log(sas.out, eps_log)
            out = lg.out
            fs, fl = forward_gather(sas), log_forward_gather(lg)
            fwd[s] = np.where(fl >= 0, fs[np.maximum(fl, 0)], -1)
            is_, il = inverse_gather(sas), log_inverse_gather(lg, eps_log)
            inv[s] = np.where(is_ >= 0, il[np.maximum(is_, 0)], -1)
            jac[s, 0], jac[s, 1] = sas.jac, lg.jac
            src.append((np.zeros(nx, np.int64), np.arange(nx)))
            for i in range(nr):
                z1, x1 = map_point(sas, rec_iz[i], rec_ix[i])
                z2, x2 = map_point(lg, z1, x1)
                riz[s, i], rix[s, i] = max(0, z2), x2
            maps.append((sas, lg))
        else:
            rt = rotation(g, angles[s])
            out = rt.out
            fwd[s] = forward_gather(rt)
            inv[s] = inverse_gather(rt)
            src.append(tuple(np.array([v]) for v in map_point(rt, src_iz[s], src_ix[s])))
            for i in range(nr):
                riz[s, i], rix[s, i] = map_point(rt, rec_iz[i], rec_ix[i])
            maps.append((rt,))
        axes[s] = (out.o1, out.d1, out.o2, out.d2)
    return Tables(axes, fwd, inv, jac, src, riz, rix, maps)


def ties(grid: Grid, c: np.ndarray, tol: float = 1e-4) -> np.ndarray:
    """Samples whose mapped cell coordinate is within tol of an integer: where
    a last-bit difference of a libm function may change the truncation."""
    fz, fx = grid.frac(np.asarray(c))
    dz = np.abs(fz - np.round(fz))
    dx = np.abs(fx - np.round(fx))
    return (dz < tol) | (dx < tol)
