"""The graph-replayed FWI gradient (wemfwi_gradient_create / _launch) against
the eager calls it records: res = WEM(m) - obs and (gS, gR) = Born^H res
(exact adjoint) bitwise equal on creation and on four replays, with eager
torch work, other library calls and cleared outputs between the replays;
J equal to the FP64 misfit of the eager residual.  Store-all, checkpointed
every 4 depths, and chunked (rows cut into chunks of 5) surveys."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(ckpt):
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey, wem_forward
    nz, nx, ns, nw = 40, 128, 3, 5
    rng = np.random.default_rng(8)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.03 * rng.standard_normal((nz, nx)))
    v0 = (1500 + 1500 * z) * np.ones((1, nx))
    den = np.ones((nz, nx), np.complex64)
    m_true = Model((1 / v ** 2).astype(np.complex64), 10.0, 10.0, nref=2, prop="pspi", eps=0.04, den=den)
    m0 = Model((1 / v0 ** 2).astype(np.complex64), 10.0, 10.0, nref=2, prop="pspi", eps=0.04, den=den)
    freqs = np.linspace(4.0, 20.0, nw).astype(np.float32)
    wav = (rng.standard_normal(nw) + 1j * rng.standard_normal(nw)).astype(np.complex64)
    sv = Survey([0] * ns, [20, 64, 100], [0] * nx, list(range(nx)), freqs, np.arange(nw), wav, nw)
    if ckpt:
        sv.set_checkpoint(ckpt)
    obs = torch.zeros(sv.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m_true, sv, obs)
    return m_true, m0, sv, obs


def _eager(m0, sv, obs):
    import torch
    from wemfwi.ops import born_adjoint, wem_forward
    pred = torch.zeros_like(obs)
    wem_forward(m0, sv, pred)
    res = pred - obs
    gS = torch.zeros((m0.nz, m0.nx), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m0, sv, gS, gR, res, exact=True)
    J = 0.5 * float((res.to(torch.complex128).abs() ** 2).sum())
    return res, gS, gR, J


@pytest.mark.parametrize("ckpt,chunk", [(0, None), (4, None), (0, 5), (4, 5)])
def test_gradient_plan_replays_the_eager_gradient(ckpt, chunk):
    import torch
    from wemfwi.ops import GradientPlan, wem_forward
    old = os.environ.get("WEMFWI_CHUNK_ROWS")
    if chunk:
        os.environ["WEMFWI_CHUNK_ROWS"] = str(chunk)
    try:
        m_true, m0, sv, obs = _setup(ckpt)
        res, gS, gR, J = _eager(m0, sv, obs)
        assert float(gS.abs().max()) > 0
        print(f"checkpoint {ckpt} chunk {chunk}: J = {J:.6e}", flush=True)
        plan = GradientPlan(m0, sv, obs)
        assert torch.equal(plan.res, res)
        assert torch.equal(plan.gS, gS) and torch.equal(plan.gR, gR)
        assert abs(plan.misfit - J) <= 1e-12 * J
        scratch = torch.zeros_like(obs)
        for k in range(4):
            plan.gS.zero_()
            plan.gR.fill_(1.0)
            plan.res.zero_()
            # eager work of the program and of the library between replays
            wem_forward(m_true, sv, scratch)
            _ = (scratch * 2).sum()
            torch.cuda.synchronize()
            j = plan.run()
            print(f"  replay {k}: J = {j:.6e}", flush=True)
            assert torch.equal(plan.res, res)
            assert torch.equal(plan.gS, gS) and torch.equal(plan.gR, gR)
            assert abs(j - J) <= 1e-12 * J
        plan.close()
        for m in (m_true, m0):
            m.close()
        sv.close()
    finally:
        if old is None:
            os.environ.pop("WEMFWI_CHUNK_ROWS", None)
        else:
            os.environ["WEMFWI_CHUNK_ROWS"] = old
