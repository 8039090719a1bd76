"""Oracle of the model interpolation operators (FWI preconditioning).

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module; the
product builds and applies the filters in the library (wemfwi.interp).

FP64 NumPy restatement of (paths relative to the reference repository):
  lanczos_filter   operator/src/LanczosInterpolation3D.h:29-55 (2-D: taper 0,
                   LanczosInterpolation2D.h:95-125)
  spline_filter    operator/src/Spline3D.h:22-52 (2-D: taper 0, Spline2D.h:97-127)
  apply / adjoint  operator/src/Interpolation3D.h:21-95 (a dense separable
                   product), the Spline2D adjoint's edge copy (Spline2D.h:84-92)
"""

from __future__ import annotations

import math

import numpy as np


def lanczos_filter(nd, dd, nm, dm, a_cells, taper=0.0):
    """Window membership and the x == 0 test in float32, as the reference's
    float coordinates decide them (a sample exactly at the window edge is in
    or out by their rounding); kernel values in double."""
    f32 = np.float32
    dd, dm = f32(dd), f32(dm)
    a = f32(f32(a_cells) * dm)
    nf = int(f32(f32(2) * a / dd + f32(1)))
    na = (nf - 1) // 2
    half = f32(f32(na) * dd)
    new_d = f32(f32(1 + taper) * dm)

    def L(x):
        x = float(x)
        return float(new_d) * float(a) * math.sin(math.pi * x / float(new_d)) * math.sin(math.pi * x / float(a)) / (
            math.pi ** 2 * x * x)

    F = np.zeros((nd, nm))
    for i in range(nd):
        for j in range(nm):
            x = f32(f32(i) * dd - f32(j) * dm)
            if abs(x) <= half:
                F[i, j] = 1.0 if x == 0 else L(x)
    for i in range(na):
        for j in range(nm):
            x = f32(f32(i) * dd - f32(j) * dm)
            xx = f32(f32(i) * dd + f32(j) * dm)
            if abs(xx) < half and x != xx:
                F[i, j] += L(xx)
                F[nd - i - 1, nm - j - 1] += L(xx)
    return F


def _cubic(y, a, b):
    if y < 1:
        return ((-6 * a - 9 * b + 12) * y ** 3 + (6 * a + 12 * b - 18) * y * y - 2 * b + 6) / 6
    if y < 2:
        return ((-6 * a - b) * y ** 3 + (30 * a + 6 * b) * y * y + (-48 * a - 12 * b) * y + 24 * a + 8 * b) / 6
    return 0.0


def spline_filter(nd, dd, nm, dm, a, b, taper=0.0):
    """Distances and the branch tests in float32 (the reference's types); the
    cubic in double."""
    f32 = np.float32
    dd, dm = f32(dd), f32(dm)
    new_d = f32(f32(1 + taper) * dm)
    F = np.zeros((nd, nm))
    for i in range(nd):
        for j in range(nm):
            F[i, j] = _cubic(float(f32(abs(f32(f32(i) * dd - f32(j) * dm)) / new_d)), a, b)
    for i in range(nd):
        for j in range(1, nm - 1):
            f = _cubic(float(f32(abs(f32(f32(i) * dd + f32(j) * dm)) / new_d)), a, b)
            F[i, j] += f
            F[nd - i - 1, nm - j - 1] += f
    return F


def apply(filters, m):
    """data = F_k ... F_1 m, filters[k] acting on SEP axis k+1 (numpy axis -(k+1))."""
    x = np.asarray(m, complex)
    for k, F in enumerate(filters):
        x = np.moveaxis(np.tensordot(F, np.moveaxis(x, -1 - k, 0), axes=(1, 0)), 0, -1 - k)
    return x


def adjoint(filters, d, spline2d=False):
    x = np.asarray(d, complex)
    for k in reversed(range(len(filters))):
        x = np.moveaxis(np.tensordot(filters[k].T, np.moveaxis(x, -1 - k, 0), axes=(1, 0)), 0, -1 - k)
    if spline2d:
        x = x.copy()
        x[:, 0] = x[:, 1]
        x[:, -1] = x[:, -2]
        x[0, :] = x[1, :]
        x[-1, :] = x[-2, :]
    return x
