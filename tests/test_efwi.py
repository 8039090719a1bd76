"""CPU checks of the frequency-extended FWI operators (scripts/efwi.py,
Operator.py:228-292,619-775): the host filter matrix against scipy.ndimage
applied to basis vectors (the reference's own call), and the oracle pinned by
calculus -- h' is the derivative of h, the 'exact' DSO taps are
i int t e^{iwt} dt over the lag window (quadrature), the Derivative adjoint is
exact in mode 'constant', the alpha term is the mean of Im over frequency."""

import ctypes as C
import math

import numpy as np
import pytest
from scipy.ndimage import convolve1d, correlate1d

from oracle import efwi_oracle as E

MODES = ["reflect", "constant", "nearest", "mirror", "wrap"]


@pytest.mark.parametrize("n", [1, 5, 12, 50])
def test_filter_matrix_matches_scipy(n):
    from wemfwi._lib import check, lib
    rng = np.random.default_rng(n)
    for L in (1, 2, 5, 8, 31, 150):
        w = rng.standard_normal(L)
        for mi, mode in enumerate(MODES):
            for conv in (0, 1):
                M = np.zeros((n, n))
                check(lib.wemfwi_filter_matrix(n, L, w.ctypes.data_as(C.POINTER(C.c_double)), mi, conv,
                                               M.ctypes.data_as(C.POINTER(C.c_double))))
                f = convolve1d if conv else correlate1d
                ref = np.stack([f(np.eye(n)[j], w, mode=mode) for j in range(n)], axis=1)
                assert np.abs(M - ref).max() < 1e-12, (n, L, mode, conv)


def test_product_filter_taps_equal_oracle():
    from wemfwi.efwi import derivative_filter
    for order in (1, 2, 4, 8):
        assert np.array_equal(derivative_filter(1, order), E.derivative_filter(1, order))
    a = derivative_filter(1, "exact", dt=0.004, tmax=3.0)
    b = E.derivative_filter(1, "exact", dt=0.004, tmax=3.0)
    assert np.allclose(a, b, rtol=1e-13, atol=1e-13)


def test_oracle_dh_is_the_derivative_of_h():
    x = np.linspace(-3, 3, 61)
    for l, tau in ((0.5, 0.1), (0.5, 1e-3), (1.0, 0.7)):
        e = 1e-6
        fd = (E.h(x + e, l, tau) - E.h(x - e, l, tau)) / (2 * e)
        assert np.allclose(E.dh(x, l, tau), fd, atol=1e-6)
    # l = 1/2 (the operators' l=1): h(x) -> x for x -> -inf, -> 0 for x -> +inf
    assert abs(E.h(-1e6, 0.5, 1e-3) - (-1e6)) < 1e-3 and abs(E.h(1e6, 0.5, 1e-3)) < 1e-3
    S = np.array([0.1 + 0j, 5.0, -5.0])
    up, _ = E.softmax(S, S, 1.0, tau=1e-6)
    lo, _ = E.softmin(S, S, -1.0, tau=1e-6)
    assert np.all(up.real <= 1.0 + 1e-9) and abs(up[0].real - 0.1) < 1e-6
    assert np.all(lo.real >= -1.0 - 1e-9) and abs(lo[0].real - 0.1) < 1e-6


def test_oracle_exact_filter_is_the_lag_weight_transform():
    dt, tmax = 0.004, 1.2
    f = E.derivative_filter(1, "exact", dt=dt, tmax=tmax)
    t = np.arange(-tmax, tmax, dt)
    w = np.flip(2 * np.pi * np.linspace(-1 / (2 * dt), 1 / (2 * dt), t.size))[::2]
    for k in (1, 7, 50, len(f) // 2 + 3, len(f) - 2):
        if w[k] == 0:
            continue
        q = E.lag_weight_transform(w[k], tmax)
        assert abs(q.imag) < 1e-9 * max(1.0, abs(q.real))
        assert abs(f[k] - q.real) < 1e-6 * max(1.0, abs(q.real))


def test_oracle_derivative_adjoint_and_alpha_term():
    rng = np.random.default_rng(1)
    nw, nz, nx = 11, 3, 4
    filt = E.derivative_filter(1, "exact", dt=0.004, tmax=0.5)
    m = rng.standard_normal((nw, nz, nx)) + 1j * rng.standard_normal((nw, nz, nx))
    d = rng.standard_normal((nw, nz, nx)) + 1j * rng.standard_normal((nw, nz, nx))
    a = np.vdot(d, E.derivative_fwd(m, filt, 0.3, 1.7, "constant")).real
    b = np.vdot(E.derivative_adj(d, filt, 0.3, 1.7, "constant"), m).real
    assert abs(a - b) < 1e-10 * abs(a)
    # beta = 0: only the mean of Im over frequency, in Im
    c = np.zeros((nw, nz, nx), complex) + 1j * np.arange(nw)[:, None, None]
    out = E.derivative_fwd(c, filt, 2.0, 0.0, "reflect")
    assert np.allclose(out, 1j * 2.0 * (nw - 1) / 2)
