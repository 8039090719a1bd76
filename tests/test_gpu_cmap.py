"""GPU parity of WEM and BornFull on conformal-map grids (par cmap = 1, 2;
PropParam.h:89-131, wem_prop.cpp:20-72, born_prop.cpp:28-231) through the
reference-style API, against the FP64 oracle on the same per-shot tables
(the tables themselves are pinned on the CPU in test_geometry.py).
Tolerance: 1e-5 relative L2; the exact adjoint passes a dot test to 1e-5."""

import json

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NX, NZ, DX, DZ = 100, 24, 10.0, 10.0     # nx + 2 pad = 128
NT, DT = 64, 0.008
BASE = {
    "ns": 2, "osx": 200.0, "dsx": 500.0, "osz": 20.0,
    "nr": 20, "orx": 15.0, "drx": 50.0, "orz": 0.0,
    "eps": 0.02, "prop": "ssf", "nref": 2, "tap": 0, "pad": 14,
    "fmin": 5.0, "fmax": 8.0, "ngs": 1, "ngr": 1, "ntaylor": 1, "tpad": 0,
}


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _wave():
    t = np.arange(NT) * DT - 0.1
    a = (np.pi * 10.0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _models(seed=5):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 1000 * z) * (1 + 0.03 * rng.standard_normal((NZ, NX)))
    return (1 / v ** 2).astype(np.complex64), (1 + 0.2 * rng.random((NZ, NX))).astype(np.complex64)


@pytest.fixture(params=[1, 2], ids=["log", "rotation"])
def par(request, tmp_path):
    p = dict(BASE)
    p["cmap"] = request.param
    if request.param == 1:
        p["eps_log"] = 0.02
    else:
        f = tmp_path / "angles.json"
        f.write_text(json.dumps({"a": [0.2, -0.15]}))
        p["cmap_angles"] = str(f)
    return p


def _setup(par, slow, den):
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DZ])
    m = SuperVector(Vector(h, arr=torch.from_numpy(slow).cuda()), Vector(h, arr=torch.from_numpy(den).cuda()))
    d = Vector(Hypercube([NT + par["tpad"], par["nr"], par["ns"]]))
    return torch, h, m, d, Trace(_wave(), DT)


def _oracle(acq, par):
    """Oracle survey with the product's per-shot devices, and the tables."""
    cm = acq.cmap
    freq, index = O.freq_grid(NT, DT, tpad=par["tpad"], fmin=par["fmin"], fmax=par["fmax"])
    sv = O.Survey(nx=NX, nz=NZ, dx=DX, dz=DZ, freq=freq, index=index, wave_f=O.wavelet_spectrum(_wave(), par["tpad"]),
                  src_iz=list(acq.src_iz), src_ix=list(acq.src_ix), rec_iz=list(acq.rec_iz),
                  rec_ix=list(acq.rec_ix), nt_f=NT + par["tpad"], tap=par["tap"], pad=par["pad"],
                  nref=par["nref"], eps=par["eps"], prop=par["prop"], ntaylor=par["ntaylor"], ngs=par["ngs"],
                  ngr=par["ngr"],
                  shot_src=[(list(cm.src[k][0]), list(cm.src[k][1])) for k in range(cm.ns)],
                  shot_rec=[(list(cm.rec_iz[k]), list(cm.rec_ix[k])) for k in range(cm.ns)])
    return sv, cm


def test_wem_on_mapped_grids_matches_oracle(par):
    from wemfwi.api import WEM
    slow, den = _models()
    torch, h, m, d, wave = _setup(par, slow, den)
    op = WEM(m, d, wave, par)
    op.forward(False, m, d)
    sv, cm = _oracle(op.acq, par)
    assert cm.ns == 2 and len(set(np.round(cm.axes[:, 1], 6))) >= 1
    want = O.wem_forward_mapped(sv, slow, den, cm.fwd, cm.jac, cm.axes)
    got = d.arr.cpu().numpy()
    assert np.abs(want).max() > 0
    assert rel(got, want) < TOL


def test_born_on_mapped_grids_matches_oracle(par):
    from wemfwi.api import Born, Hypercube, SuperVector, Vector
    slow, den = _models()
    torch, h, m, d, wave = _setup(par, slow, den)
    rng = np.random.default_rng(9)
    dS = (rng.standard_normal((NZ, NX)) * 3e-9).astype(np.complex64)
    dR = (rng.standard_normal((NZ, NX)) * 2e-2).astype(np.complex64)
    dm = SuperVector(Vector(h, arr=torch.from_numpy(dS).cuda()), Vector(h, arr=torch.from_numpy(dR).cuda()))
    for mode in ("compat", "exact"):
        op = Born(m, d, wave, par, adjoint_mode=mode)
        sv, cm = _oracle(op.acq, par)
        if mode == "compat":
            op.forward(False, dm, d)
            want = O.born_forward_mapped(sv, slow, den, dS, dR, cm.fwd, cm.jac, cm.axes)
            assert rel(d.arr.cpu().numpy(), want) < TOL
        data = (rng.standard_normal(d.arr.shape) + 1j * rng.standard_normal(d.arr.shape)).astype(np.complex64)
        dd = Vector(d.getHyper(), arr=torch.from_numpy(data).cuda())
        g = SuperVector(Vector(h), Vector(h))
        op.adjoint(False, g, dd)
        wS, wR = O.born_adjoint_mapped(sv, slow, den, data.astype(complex), cm.fwd, cm.inv, cm.jac, cm.axes,
                                       mode=mode, map_adjoint="inverse" if mode == "compat" else "transpose")
        assert rel(g[0].arr.cpu().numpy(), wS) < TOL
        assert rel(g[1].arr.cpu().numpy(), wR) < TOL


def test_born_on_mapped_grids_exact_adjoint_dot_test(par):
    from wemfwi.api import Born, Hypercube, SuperVector, Vector, cdot
    slow, den = _models(7)
    torch, h, m, d, wave = _setup(par, slow, den)
    op = Born(m, d, wave, par, adjoint_mode="exact")
    rng = np.random.default_rng(11)
    x = SuperVector(Vector(h, arr=torch.from_numpy((rng.standard_normal((NZ, NX)) * 1e-8).astype(np.complex64)).cuda()),
                    Vector(h, arr=torch.from_numpy((rng.standard_normal((NZ, NX)) * 1e-2).astype(np.complex64)).cuda()))
    y = Vector(d.getHyper(), arr=torch.from_numpy(
        (rng.standard_normal(d.arr.shape) + 1j * rng.standard_normal(d.arr.shape)).astype(np.complex64)).cuda())
    Bx = Vector(d.getHyper())
    op.forward(False, x, Bx)
    g = SuperVector(Vector(h), Vector(h))
    op.adjoint(False, g, y)
    lhs = cdot(Bx.arr, y.arr)
    rhs = cdot(x[0].arr, g[0].arr) + cdot(x[1].arr, g[1].arr)
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)
