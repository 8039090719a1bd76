"""GPU parity of the Phshift and SplitStep unit operators (py_wem_ops.cpp:25-47,
95-113) against the FP64 oracle, relative L2 <= 1e-5, through the C ABI:
the phase shift on the wavenumber grid of Phshift_cmplx.cpp:12-24 for even
and odd nk, lossy complex slowness and evanescent wavenumbers; the split-step
correction at the selected positions only; the add semantics of each
(accumulate for Phshift, assign at the positions for SplitStep); dot tests of
both adjoints; argument checks."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _c(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex64))).cuda()


def _rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


@pytest.mark.parametrize("nk,dx", [(128, 10.0), (101, 12.5)])
@pytest.mark.parametrize("s_ref", [1 / 2000.0 ** 2, complex(1 / 1800.0 ** 2, 2e-9)])
def test_phshift_matches_oracle_add_and_dot(nk, dx, s_ref):
    import torch
    from wemfwi.api import Hypercube
    from wemfwi.unit import Phshift
    dz, eps = 8.0, 0.04
    op = Phshift(dz, Hypercube([nk], [0.0], [dx]), 2, eps)
    op.setRef(1)
    k = O.kgrid(nk, dx)
    # 5 Hz leaves most of the k grid evanescent at these spacings, 40 Hz little
    for f in (5.0, 40.0):
        op.setFreq(f)
        op.setSlow(s_ref)
        w = 2 * np.pi * float(np.float32(f))
        P = O.phase_factor(w, dz, complex(np.complex64(s_ref)), eps, k)
        x = _rnd((3, nk), 1)
        d = _rnd((3, nk), 2)
        print(f"phshift nk={nk} f={f} s_ref={s_ref}", flush=True)
        y = _c(np.full((3, nk), 5.0))
        op.forward(_c(x), y)
        assert rel(y.cpu().numpy(), x * P) < TOL
        op.forward(_c(x), y, add=True)      # ps_fwd accumulates (phase_shift.ispc:13-14)
        assert rel(y.cpu().numpy(), 2 * x * P) < TOL
        g = _c(np.full((3, nk), -2.0))
        op.adjoint(g, _c(d))
        assert rel(g.cpu().numpy(), d * np.conj(P)) < TOL
        gx = _c(x)
        op.adjoint(gx, _c(d), add=True)
        assert rel(gx.cpu().numpy(), x + d * np.conj(P)) < TOL
        op.forward(_c(x), y)
        lhs = O.cdot(y.cpu().numpy().astype(complex), d)
        rhs = O.cdot(x, g.cpu().numpy().astype(complex))
        assert abs(lhs - rhs) < 1e-5 * abs(lhs)
        # one row equals the batch's row
        y1 = _c(np.zeros(nk))
        op.forward(_c(x[1]), y1)
        assert torch.equal(y1, y[1])


def _slow(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    s = 1 / v ** 2
    return (s + 1j * 1e-9 * rng.random((nz, nx))).astype(np.complex64)


def _vec(a, dx, dz):
    import torch
    from wemfwi.api import Hypercube, Vector
    return Vector(Hypercube([a.shape[-1], a.shape[-2]], [0.0, 0.0], [dx, dz]),
                  arr=torch.from_numpy(np.ascontiguousarray(a)).cuda())


def test_splitstep_matches_oracle_add_and_dot():
    import torch
    from wemfwi.unit import SplitStep
    nz, nx, dx, dz = 12, 97, 10.0, 7.5
    s = _slow(nz, nx, 3)
    op = SplitStep(_vec(s, dx, dz))
    loc = [0, 5, 6, 7, 40, 96, 41]
    sref = complex(np.complex64(s[4, 40] * 1.03))
    op.setLocation(loc)
    op.setSlow(sref)
    op.setDepth(4)
    x = _rnd((2, nx), 4)
    d = _rnd((2, nx), 5)
    for f in (3.0, 25.0):
        op.setFreq(f)
        w = 2 * np.pi * float(np.float32(f))
        M = np.zeros(nx, complex)
        M[loc] = np.exp(O._ss_phase(w, dz, s[4], sref))[loc]
        sel = np.zeros(nx, bool)
        sel[loc] = True
        print(f"splitstep f={f}", flush=True)
        y = _c(np.full((2, nx), 3.0))
        op.forward(_c(x), y)                                  # add=False zeroes the rest (SplitStep.cpp:6)
        assert rel(y.cpu().numpy(), x * M) < TOL
        y2 = _c(np.full((2, nx), 3.0))
        op.forward(_c(x), y2, add=True)                       # assigns at loc, keeps the rest
        want = np.where(sel, x * M, 3.0)
        assert rel(y2.cpu().numpy(), want) < TOL
        g = _c(np.full((2, nx), -1.0))
        op.adjoint(g, _c(d))
        assert rel(g.cpu().numpy(), d * np.conj(M)) < TOL
        lhs = O.cdot(y.cpu().numpy().astype(complex), d)
        rhs = O.cdot(x, g.cpu().numpy().astype(complex))
        assert abs(lhs - rhs) < 1e-5 * abs(lhs)
    # no positions: add=False leaves zeros, add=True leaves the data as it was
    op.setLocation([])
    y = _c(np.full(nx, 2.0))
    op.forward(_c(x[0]), y, add=True)
    assert torch.equal(y, _c(np.full(nx, 2.0)))
    op.forward(_c(x[0]), y)
    assert float(y.abs().max()) == 0.0


def test_splitstep_and_phshift_reject_bad_arguments():
    import torch
    from wemfwi._lib import WemfwiError
    from wemfwi.api import Hypercube
    from wemfwi.unit import Phshift, SplitStep
    s = _slow(6, 40, 1)
    op = SplitStep(_vec(s, 10.0, 10.0))
    x = torch.zeros(40, dtype=torch.complex64, device="cuda")
    op.setFreq(5.0)
    op.setDepth(6)                      # past the last depth
    op.setLocation([1])
    with pytest.raises(WemfwiError):
        op.forward(x, torch.zeros_like(x))
    op.setDepth(0)
    op.setLocation([40])                # past the row
    with pytest.raises(WemfwiError):
        op.forward(x, torch.zeros_like(x))
    with pytest.raises(ValueError):
        op.forward(torch.zeros(41, dtype=torch.complex64, device="cuda"), x)
    with pytest.raises(ValueError):
        Phshift(10.0, Hypercube([1], [0.0], [10.0]), 1)
    ps = Phshift(10.0, Hypercube([64], [0.0], [10.0]), 1)
    with pytest.raises(ValueError):
        ps.forward(torch.zeros(64, dtype=torch.complex64, device="cuda"),
                   torch.zeros(65, dtype=torch.complex64, device="cuda"))
