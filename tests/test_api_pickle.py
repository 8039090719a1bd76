"""Host side of the operator wrappers' from_subspace and pickling
(WEM.py:24-35, 45-47): the shot subset's geometry, and an unpickled
operator's acquisition (frequencies, devices, parameters), built without a
GPU from the WEM wrapper, which holds no background."""

import pickle

import numpy as np

NX, NZ, DX, DZ, NT, DT = 128, 20, 10.0, 10.0, 64, 0.008
PAR = {"ns": 6, "osx": 100.0, "dsx": 100.0, "osz": 0.0, "nr": 10, "orx": 0.0, "drx": 100.0, "orz": 0.0,
       "eps": 0.04, "prop": "ssf", "nref": 1, "tap": 0, "pad": 0, "fmin": 4.0, "fmax": 10.0, "tpad": 0}


class _Params:
    """genericIO.pythonParams stand-in: the dict lives in .pars."""

    def __init__(self, pars):
        self.pars = dict(pars)


def _op(ns=6, osx=100.0, par=PAR):
    import torch
    from wemfwi.api import WEM, Hypercube, SuperVector, Trace, Vector
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DZ])
    model = SuperVector(Vector(h, device="cpu"), Vector(h, device="cpu"))
    data = Vector(Hypercube([NT, PAR["nr"], ns], [0.0, 0.0, osx], [DT, PAR["drx"], PAR["dsx"]]), device="cpu")
    t = np.arange(NT) * DT - 0.1
    wave = Trace(((1 - 2 * (np.pi * 8 * t) ** 2) * np.exp(-(np.pi * 8 * t) ** 2)).astype(np.float32), DT)
    return WEM, model, data, wave, torch


def test_from_subspace_takes_the_shot_axis_of_the_subrange():
    WEM, model, data, wave, _ = _op()
    full = WEM(model, data, wave, PAR)
    _, _, sub, _, _ = _op(ns=2, osx=300.0)
    part = WEM.from_subspace(model, sub, wave, _Params(PAR))
    assert part.par["ns"] == 2 and part.par["osx"] == 300.0
    assert PAR["ns"] == 6                                    # the caller's parameters are left alone
    np.testing.assert_array_equal(part.acq.src_ix, full.acq.src_ix[2:4])
    np.testing.assert_array_equal(part.acq.rec_ix, full.acq.rec_ix)
    assert part.range is sub and part.domain is model


def test_unpickled_operator_rebuilds_its_acquisition():
    WEM, model, data, wave, _ = _op()
    op = WEM(model, data, wave, _Params(PAR))
    op2 = pickle.loads(pickle.dumps(op))
    assert op2 is not op and op2.par == op.par
    assert op2.get_nfreq() == op.get_nfreq() == len(op.acq.freq)
    np.testing.assert_array_equal(op2.acq.freq, op.acq.freq)
    np.testing.assert_array_equal(op2.acq.index, op.acq.index)
    np.testing.assert_array_equal(op2.acq.wave_f, op.acq.wave_f)
    np.testing.assert_array_equal(op2.acq.src_ix, op.acq.src_ix)
    state = op.__getstate__()
    assert "acq" not in state and "_bg" not in state
