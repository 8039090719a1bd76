"""GPU parity of the reference-style Python API (WEM, Born, BornRefl,
extWEM, extBorn, the time <-> frequency data boundary and the FWI objective)
against the FP64 oracle, built from a parameter set shaped like the
reference's own test (propagator/test/test_wem_ops.py:15-47).  Tolerance: 1e-5
relative L2."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# nx + 2 pad = 256 (WEM, Born) and nx = 128 (BornRefl on the unpadded grid)
NX, NZ, DX, DZ = 128, 40, 10.0, 10.0
PAR = {
    "ns": 2, "osx": 300.0, "dsx": 500.0, "osz": 0.0,
    "nr": 25, "orx": 0.0, "drx": 50.0, "orz": 20.0,
    "eps": 0.01, "prop": "ssf", "nref": 2, "tap": 0, "pad": 64,
    "fmin": 4.0, "fmax": 8.0, "ngs": 2, "ngr": 1, "ng": 2, "ntaylor": 1, "tpad": 20,
}
NT, DT = 100, 0.008


def _wave():
    t = np.arange(NT) * DT - 0.15
    a = (np.pi * 8.0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _models(seed=4):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.03 * rng.standard_normal((NZ, NX)))
    slow = (1 / v ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((NZ, NX))).astype(np.complex64)
    return slow, den


def _oracle_survey(par=PAR):
    n = NT + par["tpad"]
    freq, index = O.freq_grid(NT, DT, tpad=par["tpad"], fmin=par["fmin"], fmax=par["fmax"])
    wave_f = O.wavelet_spectrum(_wave(), par["tpad"])
    siz, six = O.device_indices(par["ns"], par["osx"], par["osz"], par["dsx"], 0.0, 0.0, 0.0, DX, DZ)
    riz, rix = O.device_indices(par["nr"], par["orx"], par["orz"], par["drx"], 0.0, 0.0, 0.0, DX, DZ)
    return O.Survey(nx=NX, nz=NZ, dx=DX, dz=DZ, freq=freq, index=index, wave_f=wave_f,
                    src_iz=siz, src_ix=six, rec_iz=riz, rec_ix=rix, nt_f=n, tap=par["tap"],
                    pad=par["pad"], nref=par["nref"], eps=par["eps"], prop=par["prop"],
                    ntaylor=par["ntaylor"], ngs=par["ngs"], ngr=par["ngr"], ng=par["ng"])


def _to_time(df, nt=NT):
    """FFT1::adjoint + crop: Re(unitary inverse FFT) of full-length spectra."""
    n = df.shape[-1]
    return (np.fft.ifft(df, axis=-1) * np.sqrt(n)).real[..., :nt]


def _api(slow, den):
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DZ])
    S = Vector(h, arr=torch.from_numpy(slow).cuda())
    R = Vector(h, arr=torch.from_numpy(den).cuda())
    wave = Trace(_wave(), DT)
    return torch, h, SuperVector(S, R), wave


def _data(torch, complex_data=True, par=PAR):
    from wemfwi.api import Hypercube, Vector
    n = NT + par["tpad"]
    if complex_data:
        return Vector(Hypercube([n, par["nr"], par["ns"]]))
    return Vector(Hypercube([NT, par["nr"], par["ns"]]), storage="dataFloat")


def test_wem_complex_and_float_data_match_oracle():
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.api import WEM
    sv = _oracle_survey()
    want = O.wem_forward(sv, slow, den)
    d = _data(torch, True)
    op = WEM(model, d, wave, PAR)
    op.forward(False, model, d)
    got = d.arr.cpu().numpy()
    assert rel(got[:, :, sv.index], want[:, :, sv.index]) < TOL
    dt = _data(torch, False)
    op.forward(False, model, dt)
    assert rel(dt.arr.cpu().numpy(), _to_time(want)) < TOL
    # add=True accumulates
    op.forward(True, model, dt)
    assert rel(dt.arr.cpu().numpy(), 2 * _to_time(want)) < TOL


@pytest.mark.parametrize("complex_data", [True, False])
def test_born_forward_adjoint_match_oracle(complex_data):
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.api import Born, SuperVector, Vector
    sv = _oracle_survey()
    rng = np.random.default_rng(12)
    dS = (rng.standard_normal((NZ, NX)) * 1e-8).astype(np.complex64)
    dR = (rng.standard_normal((NZ, NX)) * 0.05).astype(np.complex64)
    pert = SuperVector(Vector(h, arr=torch.from_numpy(dS).cuda()), Vector(h, arr=torch.from_numpy(dR).cuda()))
    d = _data(torch, complex_data)
    op = Born(model, d, wave, PAR)
    op.forward(False, pert, d)
    want = O.born_forward(sv, slow, den, dS, dR)
    if complex_data:
        assert rel(d.arr.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL
        res = np.zeros(want.shape, complex)
        res[:, :, sv.index] = rng.standard_normal((PAR["ns"], PAR["nr"], len(sv.index))) * (1 + 1j)
        d.arr.copy_(torch.from_numpy(res.astype(np.complex64)))
        res_f = res
    else:
        assert rel(d.arr.cpu().numpy(), _to_time(want)) < TOL
        tr = rng.standard_normal((PAR["ns"], PAR["nr"], NT)).astype(np.float32)
        d.arr.copy_(torch.from_numpy(tr))
        n = NT + PAR["tpad"]
        res_f = np.fft.fft(np.concatenate([tr, np.zeros(tr.shape[:-1] + (PAR["tpad"],))], -1), axis=-1) / np.sqrt(n)
    g = pert.clone().zero()
    op.adjoint(False, g, d)
    wS, wR = O.born_adjoint(sv, slow, den, res_f, mode="compat")
    assert rel(g[0].arr.cpu().numpy(), wS) < TOL
    assert rel(g[1].arr.cpu().numpy(), wR) < TOL


def test_born_dot_test_exact_mode_and_real_mode():
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.api import Born
    d = _data(torch, True)
    op = Born(model, d, wave, PAR, adjoint_mode="exact")
    # bins outside the survey are neither written by forward nor read by adjoint
    e0, e1 = op.dotTest()
    assert e0 < 1e-5 and e1 < 1e-5
    # mode='real' zeroes the imaginary part of the model (WEM.py:59-67)
    op_r = Born(model, d, wave, PAR, mode="real")
    g = model.clone().rand()
    op_r.forward(False, g, d)
    assert float(g[0].arr.imag.abs().max()) == 0.0
    op_r.adjoint(False, g, d)
    assert float(g[1].arr.imag.abs().max()) == 0.0


def test_bornrefl_float_data_matches_oracle():
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.api import BornRefl, Hypercube, Vector
    sv = _oracle_survey()
    nw = len(sv.freq)
    rng = np.random.default_rng(13)
    img = (rng.standard_normal((nw, NZ, NX)) + 1j * rng.standard_normal((nw, NZ, NX))).astype(np.complex64)
    image = Vector(Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0]), arr=torch.from_numpy(img).cuda())
    d = _data(torch, False)
    op = BornRefl(image, model[0], d, wave, PAR)
    op.forward(False, image, d)
    want = O.bornrefl_forward(sv, slow, img.astype(complex))
    assert rel(d.arr.cpu().numpy(), _to_time(want)) < TOL
    tr = rng.standard_normal(d.arr.shape).astype(np.float32)
    d.arr.copy_(torch.from_numpy(tr))
    g = image.clone().zero()
    op.adjoint(False, g, d)
    n = NT + PAR["tpad"]
    res_f = np.fft.fft(np.concatenate([tr, np.zeros(tr.shape[:-1] + (PAR["tpad"],))], -1), axis=-1) / np.sqrt(n)
    assert rel(g.arr.cpu().numpy(), O.bornrefl_adjoint(sv, slow, res_f)) < TOL


def test_extwem_equals_wem_through_api():
    # propagator/test/test_wem_ops.py:268-282
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.api import Hypercube, SuperVector, Vector, WEM, extWEM
    sv = _oracle_survey()
    nw = len(sv.freq)
    h3 = Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0])
    m3 = SuperVector(Vector(h3, arr=model[0].arr.expand(nw, NZ, NX).contiguous()),
                     Vector(h3, arr=model[1].arr.expand(nw, NZ, NX).contiguous()))
    d1, d2 = _data(torch, True), _data(torch, True)
    WEM(model, d1, wave, PAR).forward(False, model, d1)
    extWEM(m3, d2, wave, PAR).forward(False, m3, d2)
    assert torch.equal(d1.arr, d2.arr)


def test_trace_transforms_match_numpy_and_are_adjoint():
    import ctypes as C
    import torch
    from wemfwi._lib import check, lib
    rng = np.random.default_rng(14)
    ntr, nt, tpad, nbins = 7, 131, 29, 200
    n = nt + tpad
    bins = np.array([3, 4, 9, 17, 40, 80, 159], dtype=np.int32)
    x = rng.standard_normal((ntr, nt)).astype(np.float32)
    xd = torch.from_numpy(x).cuda()
    f = torch.zeros((ntr, nbins), dtype=torch.complex64, device="cuda")
    check(lib.wemfwi_trace_to_freq(ntr, nt, tpad, C.c_void_p(xd.data_ptr()), len(bins),
                                   bins.ctypes.data_as(C.POINTER(C.c_int)), nbins, C.c_void_p(f.data_ptr()), 0,
                                   None))
    full = np.fft.fft(np.concatenate([x, np.zeros((ntr, tpad))], -1), axis=-1) / np.sqrt(n)
    got = f.cpu().numpy()
    assert rel(got[:, bins], full[:, bins]) < 1e-6
    other = np.setdiff1d(np.arange(nbins), bins)
    assert np.all(got[:, other] == 0)
    F = np.zeros((ntr, nbins), complex)
    F[:, bins] = rng.standard_normal((ntr, len(bins))) + 1j * rng.standard_normal((ntr, len(bins)))
    Fd = torch.from_numpy(F.astype(np.complex64)).cuda()
    y = torch.zeros((ntr, nt), dtype=torch.float32, device="cuda")
    check(lib.wemfwi_freq_to_trace(ntr, nt, tpad, C.c_void_p(Fd.data_ptr()), nbins, len(bins),
                                   bins.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(y.data_ptr()), 0, None))
    Fn = np.zeros((ntr, n), complex)
    Fn[:, bins] = F[:, bins]
    assert rel(y.cpu().numpy(), _to_time(Fn, nt)) < 1e-6
    lhs = O.cdot(got.astype(complex), F)
    rhs = float(np.sum(x.astype(float) * y.cpu().numpy().astype(float)))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


@pytest.mark.parametrize("complex_data", [True, False])
def test_fwi_misfit_and_autograd_gradient_match_oracle(complex_data):
    slow, den = _models()
    torch, h, model, wave = _api(slow, den)
    from wemfwi.fwi import WEMMisfit
    sv = _oracle_survey()
    # observed data from a model 20% slower plus an anomaly: a residual of the
    # size of the data itself
    s_true = (1.2 * slow * (1 + 0.05 * np.exp(-((np.arange(NX) - 64) / 10.0) ** 2))[None, :]).astype(np.complex64)
    obs_f = O.wem_forward(sv, s_true, den)
    obs = _data(torch, complex_data)
    if complex_data:
        obs.arr.copy_(torch.from_numpy(obs_f.astype(np.complex64)))
    else:
        obs.arr.copy_(torch.from_numpy(_to_time(obs_f).astype(np.float32)))
    J = WEMMisfit(obs, wave, PAR, h, mode="real")
    S = torch.from_numpy(slow).cuda().requires_grad_(True)
    R = torch.from_numpy(den).cuda().requires_grad_(True)
    val = J(S, R)
    val.backward()
    pred = O.wem_forward(sv, slow, den)
    if complex_data:
        res_f = pred - np.asarray(obs.arr.cpu().numpy(), complex)
        res_f[:, :, np.setdiff1d(np.arange(res_f.shape[-1]), sv.index)] = 0
        j_want = 0.5 * O.cdot(res_f, res_f)
    else:
        rt = _to_time(pred) - obs.arr.cpu().numpy().astype(float)
        j_want = 0.5 * float(np.sum(rt * rt))
        n = NT + PAR["tpad"]
        res_f = np.fft.fft(np.concatenate([rt, np.zeros(rt.shape[:-1] + (PAR["tpad"],))], -1), axis=-1) / np.sqrt(n)
    assert abs(float(val) - j_want) < 1e-5 * j_want
    wS, wR = O.born_adjoint(sv, slow, den, res_f, mode="exact")
    assert rel(S.grad.cpu().numpy(), wS.real) < TOL
    assert rel(R.grad.cpu().numpy(), wR.real) < TOL
    assert float(S.grad.imag.abs().max()) == 0.0
