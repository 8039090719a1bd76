"""Operators that carry a background across pickling and shot subsets
(WEM.py:24-35, 52-80, 101-115): an unpickled Born / BornRefl / extBorn keeps
the background it had (also one set after construction) and computes bitwise
what the original did; Born.from_subspace models the subset's shots as the
full operator does; BornRefl's add_pad operator is applied to the slowness
before it becomes the background."""

import pickle

import numpy as np
import pytest

from wemfwi.api import Operator

pytestmark = pytest.mark.gpu

NX, NZ, DX, DZ, NT, DT = 128, 24, 10.0, 10.0, 80, 0.008
PAR = {"ns": 3, "osx": 200.0, "dsx": 400.0, "osz": 0.0, "nr": 20, "orx": 0.0, "drx": 60.0, "orz": 10.0,
       "eps": 0.04, "prop": "pspi", "nref": 2, "tap": 0, "pad": 64, "fmin": 4.0, "fmax": 9.0, "tpad": 16,
       "ngs": 1, "ngr": 1, "ng": 1, "ntaylor": 1}


def _slow(seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.03 * rng.standard_normal((NZ, NX)))
    return (1 / v ** 2).astype(np.complex64)


def _setup(ns=3, osx=200.0):
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DZ])
    model = SuperVector(Vector(h, arr=torch.from_numpy(_slow(1)).cuda()),
                        Vector(h, arr=torch.ones((NZ, NX), dtype=torch.complex64, device="cuda")))
    t = np.arange(NT) * DT - 0.12
    a = (np.pi * 8 * t) ** 2
    wave = Trace(((1 - 2 * a) * np.exp(-a)).astype(np.float32), DT)
    n = NT + PAR["tpad"]
    data = Vector(Hypercube([n, PAR["nr"], ns], [0.0, 0.0, osx], [1.0, PAR["drx"], PAR["dsx"]]))
    return torch, h, model, wave, data


def _pert(torch, h, seed=5):
    from wemfwi.api import SuperVector, Vector
    rng = np.random.default_rng(seed)
    dS = (rng.standard_normal((NZ, NX)) * 1e-8).astype(np.complex64)
    dR = (rng.standard_normal((NZ, NX)) * 0.05).astype(np.complex64)
    return SuperVector(Vector(h, arr=torch.from_numpy(dS).cuda()), Vector(h, arr=torch.from_numpy(dR).cuda()))


class Scale(Operator):
    """Multiplies the slowness by 1.02: a stand-in for the padding operator of
    BornRefl.add_pad (module level, so that pickle finds it)."""

    def forward(self, add, m, d):
        d.arr.copy_(m.arr * 1.02)


def test_born_pickle_keeps_the_current_background():
    from wemfwi.api import Born, SuperVector, Vector
    torch, h, model, wave, data = _setup()
    op = Born(model, data, wave, PAR, checkpoint=5)
    other = SuperVector(Vector(h, arr=torch.from_numpy(_slow(2)).cuda()), model[1].clone())
    op.set_background(other)
    pert = _pert(torch, h)
    d1 = data.clone()
    op.forward(False, pert, d1)
    op2 = pickle.loads(pickle.dumps(op))
    d2 = data.clone()
    op2.forward(False, pert, d2)
    print("Born forward after unpickling", flush=True)
    assert torch.equal(d1.arr, d2.arr)
    g1, g2 = pert.clone().zero(), pert.clone().zero()
    op.adjoint(False, g1, d1)
    op2.adjoint(False, g2, d1)
    assert torch.equal(g1[0].arr, g2[0].arr) and torch.equal(g1[1].arr, g2[1].arr)
    assert op2.checkpoint == 5 and op2.get_nfreq() == op.get_nfreq()


def test_born_from_subspace_models_the_subset_of_shots():
    from wemfwi.api import Born
    torch, h, model, wave, data = _setup()
    pert = _pert(torch, h)
    full = Born(model, data, wave, PAR)
    d = data.clone()
    full.forward(False, pert, d)
    _, _, _, _, sub = _setup(ns=2, osx=600.0)          # shots 1 and 2
    part = Born.from_subspace(model, sub, wave, PAR)
    part.forward(False, pert, sub)
    print("Born on a subset of the shots", flush=True)
    want = d.arr[1:3]
    err = float((sub.arr - want).abs().norm() / want.abs().norm())
    assert err < 1e-6, err


def test_bornrefl_pickle_and_add_pad():
    from wemfwi.api import BornRefl, Hypercube, Vector
    torch, h, model, wave, data = _setup()
    slow = model[0]
    op0 = BornRefl(Vector(Hypercube([NX, NZ, 1])), slow, data, wave, PAR)
    nw = op0.get_nfreq()
    img_h = Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0])
    rng = np.random.default_rng(8)
    img = Vector(img_h, arr=torch.from_numpy(
        (rng.standard_normal((nw, NZ, NX)) * (1 + 1j)).astype(np.complex64)).cuda())
    op = BornRefl(img, slow, data, wave, PAR)
    op.add_pad(Scale(slow, slow.clone()))
    op.set_background(slow)                        # background = 1.02 slow
    ref = BornRefl(img, Vector(h, arr=slow.arr * 1.02), data, wave, PAR)
    d1, d2 = data.clone(), data.clone()
    op.forward(False, img, d1)
    ref.forward(False, img, d2)
    print("BornRefl with a pad operator", flush=True)
    assert torch.equal(d1.arr, d2.arr)
    op2 = pickle.loads(pickle.dumps(op))
    d3 = data.clone()
    op2.forward(False, img, d3)
    assert torch.equal(d1.arr, d3.arr)
    i1, i2 = img.clone().zero(), img.clone().zero()
    op.adjoint(False, i1, d1)
    op2.adjoint(False, i2, d1)
    assert torch.equal(i1.arr, i2.arr)


def test_extborn_pickle():
    from wemfwi.api import Hypercube, SuperVector, Vector, WEM, extBorn
    torch, h, model, wave, data = _setup()
    nw = WEM(model, data, wave, PAR).get_nfreq()
    h3 = Hypercube([NX, NZ, nw], [0.0, 0.0, 0.0], [DX, DZ, 1.0])
    S = torch.from_numpy(np.stack([_slow(10 + i) for i in range(nw)])).cuda()
    bg = SuperVector(Vector(h3, arr=S), Vector(h3, arr=torch.ones_like(S)))
    rng = np.random.default_rng(9)
    pert = SuperVector(Vector(h3, arr=torch.from_numpy((rng.standard_normal((nw, NZ, NX)) * 1e-8)
                                                       .astype(np.complex64)).cuda()),
                       Vector(h3, arr=torch.zeros_like(S)))
    op = extBorn(bg, data, wave, PAR)
    op2 = pickle.loads(pickle.dumps(op))
    d1, d2 = data.clone(), data.clone()
    op.forward(False, pert, d1)
    op2.forward(False, pert, d2)
    print("extBorn after unpickling", flush=True)
    assert torch.equal(d1.arr, d2.arr)
