"""Wavefield checkpointing of the Born adjoints: the gradient computed from
rows kept every k depth steps must be bitwise the one computed with whole
wavefields stored (the same kernels, the same order of operations), for
every segment layout (k = 1, a last segment of one step, k = nz - 2, k not
dividing nz - 1), both adjoint modes, sources and receivers at depths inside
segments and on their boundaries, per-frequency models (extBorn, whose
compat mode zeroes the last depth row of all but the last shot) and per-shot
models with their own device lists.  One case is also checked against the
FP64 oracle at the parity tolerance."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5  # relative L2 against the FP64 oracle


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(nx=100, nz=40, pad=14, nw=4, ns=3, nref=2, prop="pspi", seed=21, nr=14):
    rng = np.random.default_rng(seed)
    nt = 64
    freq = (2.5 + np.arange(nw) * 3.5).astype(np.float32)
    index = np.arange(nw) + 4
    wave_f = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    # sources and receivers on segment boundaries (depth 5 for k = 5, 10 for
    # k = 5 / 10) and inside segments
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freq, index=index, wave_f=wave_f,
                  src_iz=[0, 5, 13][:ns] + [0] * max(0, ns - 3),
                  src_ix=list(np.linspace(10, nx - 11, ns).astype(int)),
                  rec_iz=[0] * (nr // 2) + [10] * (nr // 4) + [3] * (nr - nr // 2 - nr // 4),
                  rec_ix=list(np.linspace(0, nx - 1, nr).astype(int)),
                  nt_f=nt, tap=0, pad=pad, nref=nref, eps=0.04, prop=prop, ntaylor=1,
                  ngs=2, ngr=1, ng=2)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    slow = (1 / v ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, nw)) + 1j * rng.standard_normal((sv.ns, sv.nr, nw))
    return sv, slow, den, data


def _model(sv, slow, den):
    from wemfwi import Model
    sp, dp = O.edge_pad(slow, sv.pad), O.edge_pad(den, sv.pad)
    return Model(sp.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
                 eps=sv.eps, ntaylor=sv.ntaylor, den=dp.astype(np.complex64))


def _survey(sv):
    from wemfwi.ops import Survey
    return Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
                  ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)


def _c(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).cuda()


def _grad(fn, survey, shape, every):
    import torch
    survey.set_checkpoint(every)
    gS = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
    gR = torch.full(shape, float("nan"), dtype=torch.complex64, device="cuda")
    fn(gS, gR)
    return gS, gR


@pytest.mark.parametrize("exact", [False, True])
def test_born_adjoint_checkpoint_bitwise(exact):
    import torch
    from wemfwi.ops import born_adjoint
    sv, slow, den, data = _case()
    m, s, d = _model(sv, slow, den), _survey(sv), _c(data)
    shape = (sv.nz, sv.nx + 2 * sv.pad)
    fn = lambda gS, gR: born_adjoint(m, s, gS, gR, d, add=False, exact=exact)  # noqa: E731
    aS, aR = _grad(fn, s, shape, 0)
    assert torch.isfinite(aS).all() and torch.isfinite(aR).all()
    for k in (1, 2, 3, 5, 7, 10, 13, 25, sv.nz - 2, sv.nz - 1, sv.nz + 5):
        gS, gR = _grad(fn, s, shape, k)
        print(f"born_adjoint exact={exact} k={k}", flush=True)
        assert torch.equal(gS, aS), k
        assert torch.equal(gR, aR), k
    # and the checkpointed gradient against the oracle
    gS, gR = _grad(fn, s, shape, 6)
    wS, wR = O.born_adjoint(sv, slow, den, data, mode="exact" if exact else "compat")
    sl = slice(sv.pad, sv.pad + sv.nx)
    assert rel(gS.cpu().numpy()[:, sl], wS) < TOL
    assert rel(gR.cpu().numpy()[:, sl], wR) < TOL


def test_born_adjoint_checkpoint_ssf_add():
    import torch
    from wemfwi.ops import born_adjoint
    sv, slow, den, data = _case(nref=1, prop="ssf", nz=33, ns=2, seed=5)
    m, s, d = _model(sv, slow, den), _survey(sv), _c(data)
    shape = (sv.nz, sv.nx + 2 * sv.pad)
    base = torch.randn(shape, dtype=torch.complex64, device="cuda")
    out = []
    for k in (0, 4):
        s.set_checkpoint(k)
        gS, gR = base.clone(), base.clone()
        born_adjoint(m, s, gS, gR, d, add=True, exact=True)
        out.append((gS, gR))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


@pytest.mark.parametrize("exact", [False, True])
def test_extborn_adjoint_checkpoint_bitwise(exact):
    import torch
    from wemfwi.ops import extborn_adjoint
    sv, slow, den, data = _case(nz=30, ns=3, nw=3, seed=8)
    nw = len(sv.freq)
    ms = [_model(sv, (slow * (1 + 0.01 * i)).astype(np.complex64), den) for i in range(nw)]
    s, d = _survey(sv), _c(data)
    shape = (nw, sv.nz, sv.nx + 2 * sv.pad)
    fn = lambda gS, gR: extborn_adjoint(ms, s, gS, gR, d, add=False, exact=exact)  # noqa: E731
    aS, aR = _grad(fn, s, shape, 0)
    for k in (1, 4, 9, 28):
        gS, gR = _grad(fn, s, shape, k)
        print(f"extborn_adjoint exact={exact} k={k}", flush=True)
        assert torch.equal(gS, aS), k
        assert torch.equal(gR, aR), k


def test_shotborn_adjoint_checkpoint_bitwise():
    import torch
    from wemfwi.ops import shotborn_adjoint
    sv, slow, den, data = _case(nz=36, ns=3, nw=3, seed=13)
    ms = [_model(sv, (slow * (1 + 0.02 * i)).astype(np.complex64), den) for i in range(sv.ns)]
    s, d = _survey(sv), _c(data)
    rng = np.random.default_rng(2)
    # every shot its own source points (several depths) and receivers
    for sh in range(sv.ns):
        siz = np.array([sh, 4 + 3 * sh, 9 + sh])
        six = rng.integers(5, sv.nx - 5, size=3)
        riz = rng.integers(0, 12, size=sv.nr)
        rix = rng.integers(0, sv.nx, size=sv.nr)
        s.set_shot_devices(sh, siz, six, riz, rix)
    shape = (sv.ns, sv.nz, sv.nx + 2 * sv.pad)
    fn = lambda gS, gR: shotborn_adjoint(ms, s, gS, gR, d, add=False, exact=True)  # noqa: E731
    aS, aR = _grad(fn, s, shape, 0)
    for k in (1, 3, 4, 11, 34):
        gS, gR = _grad(fn, s, shape, k)
        print(f"shotborn_adjoint k={k}", flush=True)
        assert torch.equal(gS, aS), k
        assert torch.equal(gR, aR), k


def test_misfit_gradient_checkpoint_bitwise():
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector, WEM
    from wemfwi.fwi import WEMMisfit
    NX, NZ, DX, NT, DT = 100, 27, 10.0, 64, 0.008
    par = {"ns": 2, "osx": 200.0, "dsx": 300.0, "osz": 0.0, "nr": 16, "orx": 10.0, "drx": 45.0, "orz": 0.0,
           "eps": 0.02, "prop": "pspi", "nref": 3, "tap": 0, "pad": 14, "fmin": 5.0, "fmax": 11.0, "ngs": 1,
           "ngr": 0, "ntaylor": 1, "tpad": 0}
    t = np.arange(NT) * DT - 0.1
    a = (np.pi * 10.0 * t) ** 2
    wave = Trace(((1 - 2 * a) * np.exp(-a)).astype(np.float32), DT)
    z = np.arange(NZ)[:, None] * DX
    X, Z = np.meshgrid(np.arange(NX) * DX, z[:, 0])
    v0 = 1500.0 + 1.5 * z + np.zeros((NZ, NX))
    vt = v0 + 150.0 * np.exp(-((X - 400) ** 2 + (Z - 120) ** 2) / 50.0 ** 2)
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DX])
    den = torch.ones((NZ, NX), dtype=torch.complex64, device="cuda")
    S0 = torch.from_numpy((1 / v0 ** 2).astype(np.complex64)).cuda()
    St = torch.from_numpy((1 / vt ** 2).astype(np.complex64)).cuda()
    obs = Vector(Hypercube([NT, par["nr"], par["ns"]]))
    mt = SuperVector(Vector(h, arr=St), Vector(h, arr=den))
    WEM(mt, obs, wave, par).forward(False, mt, obs)
    grads = []
    for k in (0, 5):
        misfit = WEMMisfit(obs, wave, par, h, mode="real", checkpoint=k)
        slow = S0.clone().requires_grad_(False)
        res = misfit.residual(slow, den)
        grads.append(misfit.gradient(slow, den, res))
    assert torch.equal(grads[0][0], grads[1][0])
    assert torch.equal(grads[0][1], grads[1][1])
    assert float(grads[0][0].abs().max()) > 0


def test_checkpoint_interval_validated():
    from wemfwi.ops import born_adjoint  # noqa: F401 - loads the library
    sv, slow, den, data = _case(nz=12, ns=1, nw=2)
    s = _survey(sv)
    with pytest.raises(Exception):
        s.set_checkpoint(-1)
    s.set_checkpoint(3)
    assert s.checkpoint == 3
