"""Depths where RefSampler's stale buffer leaves no live reference
(RefSampler.cpp:63-81; for a slowness decreasing with depth this starts at
depth nx - 1): Selector assigns nothing there (Selector.cpp:7-9), so the step
output is zero.  The sweep and single-step kernels must give exactly that;
GPU against the FP64 oracle, relative L2 <= 1e-5, through the C ABI."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

NZ, NX, F = 160, 128, 3.0


def _slow():
    z = np.arange(NZ)[:, None] * 10.0
    v = 1500.0 + 2000.0 * z / ((NZ - 1) * 10.0) + np.zeros((NZ, NX))
    return (1.0 / v ** 2).astype(np.complex64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("up,adjoint", [(False, False), (True, False), (False, True), (True, True)])
def test_sweeps_through_depths_without_reference(up, adjoint):
    import torch
    from wemfwi import Model
    slow = _slow()
    ref = O.refsampler(slow, 1)
    assert (np.asarray(ref.refid) < 0).all(axis=1).sum() > 20
    p = O.StepParams(nx=NX, dx=10.0, dz=10.0, nref=1, prop="ssf", eps=0.04)
    rng = np.random.default_rng(2)
    src = (rng.standard_normal((NZ, NX)) + 1j * rng.standard_normal((NZ, NX))).astype(np.complex64)
    m = Model(slow, 10.0, 10.0, nref=1, prop="ssf", eps=0.04)
    a = torch.from_numpy(src).cuda()[None]
    out = torch.zeros_like(a)
    m.oneway_recursive(up, adjoint, [F], a, out)
    fn = {(False, False): O.down_fwd, (True, False): O.up_fwd, (False, True): O.down_adj,
          (True, True): O.up_adj}[(up, adjoint)]
    want = fn(src.astype(complex), 2 * np.pi * np.float32(F), slow, ref, p)
    got = out[0].cpu().numpy()
    assert np.isfinite(got).all()
    assert rel(got, want) < 1e-5
    m.close()


def test_single_step_at_a_depth_without_reference_is_zero():
    import torch
    from wemfwi import Model
    slow = _slow()
    m = Model(slow, 10.0, 10.0, nref=1, prop="ssf", eps=0.04)
    x = torch.ones((2, NX), dtype=torch.complex64, device="cuda")
    y = torch.full_like(x, 7.0)
    m.onestep(False, [F, F], [NX + 5, 3], x, y)
    assert torch.count_nonzero(y[0]) == 0 and torch.count_nonzero(y[1]) > 0
    m.onestep(True, [F, F], [NX + 5, 3], x, y)
    assert torch.count_nonzero(y[0]) == 0 and torch.count_nonzero(y[1]) > 0
    m.close()
