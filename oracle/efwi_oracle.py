"""Oracle of the frequency-extended FWI objective (scripts/efwi.py).

TEST INFRASTRUCTURE ONLY.  Only ``tests/`` may import this module; the
product runs these operators as library kernels (wemfwi.efwi).

FP64 NumPy restatement of (paths relative to the reference repository):

  =====================  =====================================================
  function here          reference it follows
  =====================  =====================================================
  h, dh                  operator/src/python/Operator.py:619-623
  hyperbolic, softclip   Operator.py:625-671
  softmax, softmin,      Operator.py:674-768
  dsoftmax, dsoftmin
  derivative_filter      Operator.py:241-268
  derivative_fwd / _adj  Operator.py:272-292 (scipy.ndimage correlate1d /
                         convolve1d along the frequency axis, as there)
  objective              scripts/efwi.py:145-230 with python-solver's
                         ProblemL2NonLinearReg (J = 1/2 |f(m) - d|^2 +
                         1/2 |eps g(m)|^2; branch rustam-dev of
                         ettore88/python-solver, not in the reference tree)
  =====================  =====================================================
"""

from __future__ import annotations

import math
from typing import Optional

import numpy as np
from scipy.ndimage import convolve1d, correlate1d

from . import wem_oracle as O


def h(x, l, tau):
    return l * x - np.sqrt(l * l * x * x + tau ** 2)


def dh(x, l, tau):
    return l - l * l * x / np.sqrt(l * l * x * x + tau ** 2)


def hyperbolic(S, R, l=1, tau=.001):
    S = np.asarray(S, complex)
    return S.real + 1j * h(S.imag, l / 2, tau), np.asarray(R, complex)


def softclip(dS, dR, bgS, l=1, tau=.001):
    dS = np.asarray(dS, complex)
    return dS.real + 1j * dS.imag * dh(np.asarray(bgS).imag, l / 2, tau), np.asarray(dR, complex)


def softmax(S, R, xmax, l=1, tau=.001):
    S = np.asarray(S, complex)
    return h(S.real - xmax, l / 2, tau) + xmax + 1j * S.imag, np.asarray(R, complex)


def softmin(S, R, xmin, l=1, tau=.001):
    S = np.asarray(S, complex)
    return -h(-(S.real - xmin), l / 2, tau) + xmin + 1j * S.imag, np.asarray(R, complex)


def dsoftmax(dS, dR, bgS, xmax, l=1, tau=.001):
    dS = np.asarray(dS, complex)
    return dS.real * dh(np.asarray(bgS).real - xmax, l / 2, tau) + 1j * dS.imag, np.asarray(dR, complex)


def dsoftmin(dS, dR, bgS, xmin, l=1, tau=.001):
    dS = np.asarray(dS, complex)
    return dS.real * dh(-(np.asarray(bgS).real - xmin), l / 2, tau) + 1j * dS.imag, np.asarray(dR, complex)


def derivative_filter(which=1, order=4, dt=None, tmax=None):
    if which == 1:
        table = {1: [-1, 1], 2: [-1 / 2, 0, 1 / 2], 4: [1 / 12, -2 / 3, 0, 2 / 3, -1 / 12],
                 8: [1 / 280, -4 / 105, 1 / 5, -4 / 5, 0, 4 / 5, -1 / 5, 4 / 105, -1 / 280]}
        if order in table:
            return np.array(table[order], float)
        if order == "exact":
            t = np.arange(-tmax, tmax, dt)
            w = np.flip(2 * np.pi * np.linspace(-1 / (2 * dt), 1 / (2 * dt), t.size))
            with np.errstate(divide="ignore", invalid="ignore"):
                f = np.where(w == 0, 0, (np.cos(w * tmax / 2) / w) * tmax - 2 * (np.sin(w * tmax / 2) / w ** 2))
            return f[::2]
    if which == 2:
        return np.array({2: [1, -2, 1], 4: [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12]}[order], float)
    raise ValueError(order)


def derivative_fwd(m, filt, alpha=0.0, beta=1.0, mode="reflect"):
    m = np.asarray(m, complex)
    mean = np.sum(m.imag, axis=0) / m.shape[0]
    out_re = -beta * correlate1d(m.imag, filt, axis=0, mode=mode)
    out_im = beta * correlate1d(m.real, filt, axis=0, mode=mode) + alpha * mean
    return out_re + 1j * out_im


def derivative_adj(d, filt, alpha=0.0, beta=1.0, mode="reflect"):
    d = np.asarray(d, complex)
    mean = np.sum(d.imag, axis=0) / d.shape[0]
    out_re = beta * convolve1d(d.imag, filt, axis=0, mode=mode)
    out_im = -beta * convolve1d(d.real, filt, axis=0, mode=mode) + alpha * mean
    return out_re + 1j * out_im


def objective(sv: O.Survey, S3, R3, obs, eps, filt, alpha=0.0, beta=1.0, mode="reflect", tau=.001,
              vmin: Optional[float] = None, vmax: Optional[float] = None, mask=None):
    """(J, gS, gR) of the efwi.py problem with the exact adjoint of extBorn."""
    S3 = np.asarray(S3, complex)
    R3 = np.asarray(R3, complex)
    q = (S3, R3)
    if vmin is not None and vmax is not None:
        xmin, xmax = 1 / vmax ** 2, 1 / vmin ** 2
        a = softmax(S3, R3, xmax, tau=0.01)
        q = softmin(*a, xmin, tau=0.001)
    p = hyperbolic(*q, tau=tau)
    res = O.extwem_forward(sv, p[0], p[1]) - obs
    gS, gR = O.extborn_adjoint(sv, p[0], p[1], res, mode="exact")
    if mask is not None:
        gS, gR = gS * mask[0], gR * mask[1]
    gS, gR = softclip(gS, gR, q[0], tau=tau)
    if vmin is not None and vmax is not None:
        gS, gR = dsoftmin(gS, gR, a[0], xmin, tau=0.001)
        gS, gR = dsoftmax(gS, gR, S3, xmax, tau=0.01)
    dm = [derivative_fwd(v, filt, alpha, beta, mode) for v in (S3, R3)]
    dd = [derivative_adj(v, filt, alpha, beta, mode) for v in dm]
    J = 0.5 * np.vdot(res, res).real + 0.5 * eps ** 2 * sum(np.vdot(v, v).real for v in dm)
    return J, gS + eps ** 2 * dd[0], gR + eps ** 2 * dd[1]


def lag_weight_transform(w, T, n=20001):
    """i * int_{-T/2}^{T/2} t e^{i w t} dt by the composite Simpson rule (for
    pinning the 'exact' filter)."""
    t = np.linspace(-T / 2, T / 2, n)
    y = t * np.exp(1j * w * t)
    hstep = t[1] - t[0]
    s = y[0] + y[-1] + 4 * y[1:-1:2].sum() + 2 * y[2:-1:2].sum()
    return 1j * s * hstep / 3
