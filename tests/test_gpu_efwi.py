"""GPU parity of the frequency-extended FWI operators and objective
(scripts/efwi.py:145-230, Operator.py:228-292,619-775) against the FP64
oracle (oracle/efwi_oracle.py, with scipy.ndimage's filters as the
reference calls them): relative L2 <= 1e-5; the Derivative adjoint passes a
dot test in mode 'constant'.  Units are km and s (slowness squared in
s^2/km^2), the scale efwi.py's soft-clip widths are written for."""

import numpy as np
import pytest

from oracle import efwi_oracle as E
from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NX, NZ, DX, DZ = 100, 16, 0.01, 0.01     # km; nx + 2 pad = 128
NT, DT = 64, 0.008
PAR = {"ns": 2, "osx": 0.2, "dsx": 0.5, "osz": 0.0, "nr": 12, "orx": 0.01, "drx": 0.08, "orz": 0.0,
       "eps": 0.02, "prop": "ssf", "nref": 2, "tap": 0, "pad": 14, "fmin": 5.0, "fmax": 8.0, "ngs": 1,
       "ngr": 0, "ntaylor": 1, "tpad": 0}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _wave():
    t = np.arange(NT) * DT - 0.1
    a = (np.pi * 10.0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _nw():
    freq, _ = O.freq_grid(NT, DT, fmin=PAR["fmin"], fmax=PAR["fmax"])
    return len(freq)


def _vec3(arr):
    import torch
    from wemfwi.api import Hypercube, Vector
    nw = arr.shape[0]
    return Vector(Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0]),
                  arr=torch.from_numpy(np.ascontiguousarray(arr.astype(np.complex64))).cuda())


def _models3(seed=3, atten=True):
    rng = np.random.default_rng(seed)
    nw = _nw()
    z = np.linspace(0, 1, NZ)[None, :, None]
    v = (1.5 + 1.2 * z) * (1 + 0.03 * rng.standard_normal((nw, NZ, NX)))
    S = (1 / v ** 2).astype(np.complex64)
    if atten:
        S = (S - 1j * 0.01 * np.abs(rng.standard_normal(S.shape))).astype(np.complex64)
    R = (1 + 0.2 * rng.random((nw, NZ, NX)) + 0.01j * rng.standard_normal((nw, NZ, NX))).astype(np.complex64)
    return S, R


def test_soft_clips_match_oracle():
    from wemfwi.api import SuperVector
    from wemfwi.efwi import HyperbolicPenalty, Softclip, Softmax, Softmin, dSoftmax, dSoftmin
    S, R = _models3(1)
    rng = np.random.default_rng(2)
    S = (S + 1j * 0.02 * rng.standard_normal(S.shape)).astype(np.complex64)
    m = SuperVector(_vec3(S), _vec3(R))
    bgS = (S * (1 + 0.1 * rng.standard_normal(S.shape)) + 1j * 0.02 * rng.standard_normal(S.shape)).astype(np.complex64)
    bg = SuperVector(_vec3(bgS), _vec3(R))
    out = m.clone().zero()
    xmin, xmax = float(np.percentile(S.real, 20)), float(np.percentile(S.real, 80))
    tau = 1e-3
    cases = [
        (HyperbolicPenalty(m, out, l=1, tau=tau), E.hyperbolic(S, R, tau=tau)),
        (Softmax(m, out, xmax, tau=tau), E.softmax(S, R, xmax, tau=tau)),
        (Softmin(m, out, xmin, tau=tau), E.softmin(S, R, xmin, tau=tau)),
    ]
    for jac, want in ((Softclip(bg, out, tau=tau), E.softclip(S, R, bgS, tau=tau)),
                      (dSoftmax(bg, out, xmax, tau=tau), E.dsoftmax(S, R, bgS, xmax, tau=tau)),
                      (dSoftmin(bg, out, xmin, tau=tau), E.dsoftmin(S, R, bgS, xmin, tau=tau))):
        cases.append((jac, want))
    for op, (wS, wR) in cases:
        op.forward(False, m, out)
        assert rel(out[0].arr.cpu().numpy(), wS) < TOL, type(op).__name__
        assert rel(out[1].arr.cpu().numpy(), wR) < TOL
        before = out[0].arr.clone()
        op.forward(True, m, out)                       # add = True accumulates
        assert rel(out[0].arr.cpu().numpy(), 2 * before.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("mode", ["reflect", "constant", "nearest", "mirror", "wrap"])
def test_derivative_matches_oracle(mode):
    from wemfwi.api import SuperVector
    from wemfwi.efwi import Derivative
    S, R = _models3(4)
    rng = np.random.default_rng(5)
    D = (rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)).astype(np.complex64)
    m = SuperVector(_vec3(S), _vec3(R))
    d = SuperVector(_vec3(D), _vec3(D[::-1].copy()))
    for order, kw in ((4, {}), ("exact", {"dt": 0.004, "tmax": 0.6})):
        op = Derivative(m, d, which=1, order=order, f0=5.0, alpha=0.7, beta=1.3, mode=mode,
                        tmax=kw.get("tmax", 1.0), dt=kw.get("dt"))
        filt = E.derivative_filter(1, order, dt=kw.get("dt"), tmax=kw.get("tmax"))
        out = d.clone().zero()
        op.forward(False, m, out)
        for k, src in enumerate((S, R)):
            assert rel(out[k].arr.cpu().numpy(), E.derivative_fwd(src.astype(complex), filt, 0.7, 1.3, mode)) < TOL
        back = m.clone().zero()
        op.adjoint(False, back, d)
        for k, src in enumerate((D, D[::-1])):
            assert rel(back[k].arr.cpu().numpy(), E.derivative_adj(src.astype(complex), filt, 0.7, 1.3, mode)) < TOL
        if mode == "constant":
            e0, e1 = op.dotTest()
            assert e0 < 1e-5 and e1 < 1e-5


@pytest.mark.parametrize("clip,masked", [(False, False), (True, True)])
def test_extended_objective_and_gradient_match_oracle(clip, masked):
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector, extWEM
    from wemfwi.efwi import ExtFWIObjective
    S0, R0 = _models3(6, atten=False)
    St, Rt = _models3(7, atten=False)
    nw = S0.shape[0]
    wave = Trace(_wave(), DT)
    h3 = Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0])
    obs = Vector(Hypercube([NT, PAR["nr"], PAR["ns"]]))
    mt = SuperVector(_vec3(St), _vec3(Rt))
    extWEM(mt, obs, wave, PAR).forward(False, mt, obs)
    rng = np.random.default_rng(8)
    mask = None
    if masked:
        mk = (rng.random((nw, NZ, NX)) > 0.2).astype(np.complex64)
        mask = [torch.from_numpy(mk).cuda(), torch.from_numpy(mk).cuda()]
    vmin, vmax = (1.55, 2.5) if clip else (None, None)
    eps, alpha, beta, tau = 0.05, 0.2, 1.0, 1e-3
    obj = ExtFWIObjective(obs, wave, PAR, h3, epsilon=eps, alpha=alpha, beta=beta, der_mode="reflect", tau=tau,
                          vmin=vmin, vmax=vmax, mask=mask)
    m = SuperVector(_vec3(S0), _vec3(R0))
    J, g = obj.value_and_gradient(m)
    assert abs(obj.value(m) - J) < 1e-6 * J
    freq, index = O.freq_grid(NT, DT, fmin=PAR["fmin"], fmax=PAR["fmax"])
    acq = obj._wem.acq
    sv = O.Survey(nx=NX, nz=NZ, dx=DX, dz=DZ, freq=freq, index=index, wave_f=O.wavelet_spectrum(_wave()),
                  src_iz=list(acq.src_iz), src_ix=list(acq.src_ix), rec_iz=list(acq.rec_iz), rec_ix=list(acq.rec_ix),
                  nt_f=NT, tap=0, pad=PAR["pad"], nref=PAR["nref"], eps=PAR["eps"], prop="ssf", ntaylor=1,
                  ngs=PAR["ngs"], ngr=PAR["ngr"])
    wa = wave.getHyper().getAxis(1)
    filt = E.derivative_filter(1, "exact", dt=wa.d, tmax=((wa.n - 1) * wa.d + wa.o) / 2)
    mk = None if mask is None else [t.cpu().numpy() for t in mask]
    wJ, wS, wR = E.objective(sv, S0.astype(complex), R0.astype(complex), obs.arr.cpu().numpy().astype(complex),
                             eps, filt, alpha, beta, "reflect", tau, vmin, vmax, mk)
    assert abs(J - wJ) < 1e-5 * abs(wJ)
    assert rel(g[0].arr.cpu().numpy(), wS) < TOL
    assert rel(g[1].arr.cpu().numpy(), wR) < TOL
