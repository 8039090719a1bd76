"""Survey device depths are checked on the host against the model before any
work is queued: a source or receiver depth outside [0, nz) returns
WEMFWI_ERR_ARG from every shot x frequency operator (the drivers clear and
inject only the rows of these depths, so an unchecked index would address
memory outside the wavefield buffers), and a survey whose injection plans
were built for a deeper model is checked again, not reused, with a
shallower model of the same width."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(nz, nx, seed=3):
    from wemfwi import Model
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    slow = (1 / ((1500 + 1000 * z) * (1 + 0.03 * rng.standard_normal((nz, nx)))) ** 2).astype(np.complex64)
    return Model(slow, 10.0, 10.0, nref=2, prop="pspi", eps=0.04)


def _survey(src_iz, rec_iz, nx, nw=2):
    from wemfwi.ops import Survey
    freq = (3.0 + 2.0 * np.arange(nw)).astype(np.float32)
    rix = list(range(0, nx, 4))
    riz = [rec_iz] * len(rix)
    return Survey(src_iz, [nx // 3, 2 * nx // 3][:len(src_iz)], riz, rix, freq, np.arange(nw),
                  np.ones(nw, np.complex64), nw)


def test_out_of_range_depths_are_argument_errors():
    import torch
    from wemfwi._lib import ERR_ARG, WemfwiError
    from wemfwi.ops import (born_adjoint, born_forward, bornrefl_adjoint, bornrefl_forward, shotwem_forward,
                            wem_forward)
    nz, nx = 12, 64
    m = _model(nz, nx)
    dev = "cuda"
    gS = torch.zeros((nz, nx), dtype=torch.complex64, device=dev)
    gR = torch.zeros_like(gS)
    img = torch.zeros((2, nz, nx), dtype=torch.complex64, device=dev)
    for src_iz, rec_iz in (([0, nz], 0), ([0, -1], 0), ([0, 1], nz), ([0, 1], -3), ([0, 1], nz + 40)):
        s = _survey(src_iz, rec_iz, nx)
        d = torch.zeros(s.data_shape, dtype=torch.complex64, device=dev)
        calls = [lambda: wem_forward(m, s, d),
                 lambda: born_forward(m, s, gS, gR, d),
                 lambda: born_adjoint(m, s, gS, gR, d),
                 lambda: bornrefl_forward(m, s, img, d),
                 lambda: bornrefl_adjoint(m, s, img, d),
                 lambda: shotwem_forward([m, m], s, d)]
        for k, call in enumerate(calls):
            print(f"src_iz={src_iz} rec_iz={rec_iz} call {k}", flush=True)
            with pytest.raises(WemfwiError) as e:
                call()
            assert e.value.code == ERR_ARG, e.value
        s.set_checkpoint(4)
        with pytest.raises(WemfwiError) as e:
            born_adjoint(m, s, gS, gR, d)
        assert e.value.code == ERR_ARG
        s.close()
    # per-shot device lists are checked the same way
    s = _survey([0, 1], 0, nx)
    s.set_shot_devices(1, [nz + 2], [5], [0] * s.nr, list(range(0, nx, 4)))
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device=dev)
    with pytest.raises(WemfwiError) as e:
        shotwem_forward([m, m], s, d)
    assert e.value.code == ERR_ARG
    s.close()
    m.close()


def test_survey_reused_with_a_shallower_model():
    import torch
    from wemfwi._lib import ERR_ARG, WemfwiError
    from wemfwi.ops import wem_forward
    nx = 64
    deep, shallow = _model(20, nx), _model(8, nx, seed=4)
    s = _survey([0, 2], 12, nx)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(deep, s, d)               # builds and caches the receiver plans (depth 12 of 20)
    first = d.clone()
    assert torch.isfinite(first.abs()).all() and float(first.abs().max()) > 0
    with pytest.raises(WemfwiError) as e:
        wem_forward(shallow, s, d)        # depth 12 is outside the 8-deep model
    assert e.value.code == ERR_ARG
    d.zero_()
    wem_forward(deep, s, d)               # the survey still works with the deep model
    assert torch.equal(d, first)
    s.close()
    deep.close()
    shallow.close()
