"""GPU checks of the FWI driver pieces: the interpolation operators against
the oracle (forward and adjoint, relative L2 <= 1e-5; dot tests where the
reference's adjoint is exact), and a short preconditioned FWI run (L-BFGS and
NLCG on the library's objective, a Lanczos coarse model fitted by CG) whose
objective decreases at every iteration and whose preconditioned gradient is
P^H of the full gradient."""

import numpy as np
import pytest

from oracle import interp_oracle as I

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _vec(shape, d, seed=None):
    import torch
    from wemfwi.api import Hypercube, Vector
    ns = list(reversed(shape))
    h = Hypercube(ns, [0.0] * len(ns), list(reversed(d)))
    if seed is None:
        return Vector(h)
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    return Vector(h, arr=torch.from_numpy(a).cuda())


@pytest.mark.parametrize("which", ["lanczos2", "spline2", "lanczos3", "spline3"])
def test_interpolation_matches_oracle(which):
    from wemfwi import interp
    if which.endswith("2"):
        ms, md, dm, dd = (7, 9), (31, 41), (5.0, 5.0), (1.0, 1.0)
    else:
        ms, md, dm, dd = (3, 6, 8), (9, 26, 36), (2.0, 5.0, 5.0), (0.5, 1.0, 1.0)
    m, d = _vec(ms, dm, 1), _vec(md, dd, 2)
    # filters per SEP axis k+1 (numpy axis -(k+1))
    axes_m = [(ms[-1 - k], dm[-1 - k]) for k in range(len(ms))]
    axes_d = [(md[-1 - k], dd[-1 - k]) for k in range(len(ms))]
    if which == "lanczos2":
        op = interp.LanczosInterpolation2D(m, d, a=[3, 2])
        fs = [I.lanczos_filter(axes_d[k][0], axes_d[k][1], axes_m[k][0], axes_m[k][1], [3, 2][k]) for k in range(2)]
    elif which == "spline2":
        op = interp.Spline2D(m, d, type="CR-spline")
        fs = [I.spline_filter(axes_d[k][0], axes_d[k][1], axes_m[k][0], axes_m[k][1], 0.5, 0.0) for k in range(2)]
    elif which == "lanczos3":
        op = interp.LanczosInterpolation3D(m, d, a=[2, 2, 1], taper=[0.1, 0.0, 0.2])
        fs = [I.lanczos_filter(axes_d[k][0], axes_d[k][1], axes_m[k][0], axes_m[k][1], [2, 2, 1][k],
                               [0.1, 0.0, 0.2][k]) for k in range(3)]
    else:
        op = interp.Spline3D(m, d, type="B-spline", taper=[0.0, 0.1, 0.0])
        fs = [I.spline_filter(axes_d[k][0], axes_d[k][1], axes_m[k][0], axes_m[k][1], 0.0, 1.0, [0.0, 0.1, 0.0][k])
              for k in range(3)]
    out = _vec(md, dd)
    op.forward(False, m, out)
    assert rel(out.arr.cpu().numpy(), I.apply(fs, m.arr.cpu().numpy())) < TOL
    back = _vec(ms, dm)
    op.adjoint(False, back, d)
    assert rel(back.arr.cpu().numpy(), I.adjoint(fs, d.arr.cpu().numpy(), spline2d=which == "spline2")) < TOL
    if which != "spline2":
        e0, e1 = op.dotTest()
        assert e0 < 1e-5 and e1 < 1e-5


def _fwi_setup():
    import torch
    from wemfwi.api import Hypercube, SuperVector, Trace, Vector, WEM
    from wemfwi.fwi import WEMMisfit
    NX, NZ, DX = 100, 16, 10.0
    NT, DT = 64, 0.008
    par = {"ns": 2, "osx": 200.0, "dsx": 500.0, "osz": 0.0, "nr": 20, "orx": 10.0, "drx": 50.0, "orz": 0.0,
           "eps": 0.02, "prop": "ssf", "nref": 2, "tap": 0, "pad": 14, "fmin": 5.0, "fmax": 9.0, "ngs": 1,
           "ngr": 0, "ntaylor": 1, "tpad": 0}
    t = np.arange(NT) * DT - 0.1
    a = (np.pi * 10.0 * t) ** 2
    wave = Trace(((1 - 2 * a) * np.exp(-a)).astype(np.float32), DT)
    z = np.arange(NZ)[:, None] * DX
    X, Z = np.meshgrid(np.arange(NX) * DX, z[:, 0])
    v0 = 1500.0 + 1.5 * z + np.zeros((NZ, NX))
    vt = v0 + 150.0 * np.exp(-((X - 500) ** 2 + (Z - 90) ** 2) / 60.0 ** 2)
    h = Hypercube([NX, NZ], [0.0, 0.0], [DX, DX])
    den = torch.ones((NZ, NX), dtype=torch.complex64, device="cuda")
    S0 = torch.from_numpy((1 / v0 ** 2).astype(np.complex64)).cuda()
    St = torch.from_numpy((1 / vt ** 2).astype(np.complex64)).cuda()
    obs = Vector(Hypercube([NT, par["nr"], par["ns"]]))
    mt = SuperVector(Vector(h, arr=St), Vector(h, arr=den))
    WEM(mt, obs, wave, par).forward(False, mt, obs)
    misfit = WEMMisfit(obs, wave, par, h, mode="real")
    return misfit, S0, den, h, (NZ, NX, DX)


@pytest.mark.parametrize("method", ["lbfgs", "nlcg"])
def test_preconditioned_fwi_decreases_the_misfit(method):
    import torch
    from wemfwi import solver
    from wemfwi.api import Hypercube, Vector
    from wemfwi.fwi import fit_preconditioned, fwi_problem
    from wemfwi.interp import LanczosInterpolation2D
    misfit, S0, den, h, (NZ, NX, DX) = _fwi_setup()
    cs = (6, 11)
    hc = Hypercube([cs[1], cs[0]], [0.0, 0.0], [(NX - 1) * DX / (cs[1] - 1), (NZ - 1) * DX / (cs[0] - 1)])
    P = LanczosInterpolation2D(Vector(hc), Vector(h), a=[2, 2])
    m_pre = fit_preconditioned(P, S0, cs, 20)
    full = torch.zeros_like(S0)
    P.forward(False, m_pre, full)
    assert rel(full.cpu().numpy(), S0.cpu().numpy()) < 3e-2        # a coarse Lanczos fit of the start
    prob = fwi_problem(misfit, den, precond=P, coarse_shape=cs, vmin=1400.0, vmax=1800.0)
    f0, g0 = prob.fg([m_pre])
    # preconditioned gradient = P^H (full gradient) on the fine grid
    gS, _ = misfit.gradient(full, den, misfit.residual(full, den))
    fs = [I.lanczos_filter(NX, DX, cs[1], (NX - 1) * DX / (cs[1] - 1), 2),
          I.lanczos_filter(NZ, DX, cs[0], (NZ - 1) * DX / (cs[0] - 1), 2)]
    assert rel(g0[0].cpu().numpy(), I.adjoint(fs, gS.cpu().numpy())) < 1e-4
    s = solver.LBFGS(3) if method == "lbfgs" else solver.NLCG(3, beta_type="HZ")
    x, hist = s.run(prob, [m_pre])
    assert len(hist.objective) >= 3
    assert all(b < a for a, b in zip(hist.objective, hist.objective[1:]))
    assert hist.objective[-1] < 0.95 * hist.objective[0]
    lo, hi = 1 / 1800.0 ** 2, 1 / 1400.0 ** 2
    r = x[0].real
    assert float(r.min()) >= lo - 1e-12 and float(r.max()) <= hi + 1e-12
