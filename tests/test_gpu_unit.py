"""GPU parity of the unit operators (wemfwi.unit, the reference's
py_wem_ops.cpp classes) against the FP64 oracle, relative L2 <= 1e-5, through
the C ABI; dot tests of the exact adjoints; the add=True semantics the
reference gives each operator."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NX, NZ, TAP, PAD, DX, DZ = 100, 16, 14, 10, 10.0, 8.0     # FFT length 128
PAR = {"tap": TAP, "pad": PAD, "nref": 2, "prop": "ssf", "eps": 0.04, "ntaylor": 1}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _slow(seed=0):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((NZ, NX)))
    return (1 / v ** 2).astype(np.complex64)


def _vec(a):
    import torch
    from wemfwi.api import Hypercube, Vector
    return Vector(Hypercube([a.shape[-1], a.shape[-2]], [0.0, 0.0], [DX, DZ]),
                  arr=torch.from_numpy(np.ascontiguousarray(a)).cuda())


def _c(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex64))).cuda()


def _rnd(shape, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


def _p(ntaylor=1):
    return O.StepParams(nx=NX, dx=DX, dz=DZ, tap=TAP, pad=PAD, nref=PAR["nref"], eps=PAR["eps"], prop="ssf",
                        ntaylor=ntaylor)


@pytest.mark.parametrize("ntaylor", [0, 1, 2])
def test_scatter_matches_oracle_and_exact_dot(ntaylor):
    import torch
    from wemfwi.unit import Scatter
    s = _slow(1)
    x = _rnd((NX,), 2)
    d = _rnd((NX,), 3)
    p = _p(ntaylor)
    for mode in ("compat", "exact"):
        op = Scatter(_vec(s), ntaylor, PAR, adjoint_mode=mode)
        op.setFreq(7.5)
        op.setDepth(5)
        w = 2 * np.pi * np.float32(7.5)
        y = torch.zeros(NX, dtype=torch.complex64, device="cuda")
        op.forward(_c(x), y)
        assert rel(y.cpu().numpy(), O.scatter_fwd(x.astype(complex), 5, w, s, p)) < TOL
        g = torch.zeros_like(y)
        op.adjoint(g, _c(d))
        assert rel(g.cpu().numpy(), O.scatter_adj(d.astype(complex), 5, w, s, p, mode)) < TOL
        if mode == "exact":
            lhs = O.cdot(y.cpu().numpy().astype(complex), d)
            rhs = O.cdot(x, g.cpu().numpy().astype(complex))
            assert abs(lhs - rhs) < 1e-5 * abs(lhs)


@pytest.mark.parametrize("up", [False, True])
@pytest.mark.parametrize("mode", ["compat", "exact"])
def test_lin_oneway_matches_oracle_with_add(up, mode):
    import torch
    from wemfwi.unit import LinDown, LinUp
    s = _slow(4)
    p = _p()
    ref = O.refsampler(s, PAR["nref"])
    bg = _rnd((NZ, NX), 5)
    m = (_rnd((NZ, NX), 6) * 1e-3).astype(np.complex64)
    d = _rnd((NZ, NX), 7)
    init = _rnd((NZ, NX), 8)
    cls = LinUp if up else LinDown
    op = cls(_vec(s), PAR, _c(bg), adjoint_mode=mode)
    op.setFreq(11.0)
    w = 2 * np.pi * np.float32(11.0)
    fwd = O.linup_fwd if up else O.lindown_fwd
    adj = O.linup_adj if up else O.lindown_adj
    out = _c(init)
    op.forward(_c(m), out, add=True)
    want = fwd(m.astype(complex), bg.astype(complex), w, s, ref, p, out=init.astype(complex), add=True)
    assert rel(out.cpu().numpy(), want) < TOL
    g = _c(init)
    op.adjoint(g, _c(d), add=True)
    want = adj(d.astype(complex), bg.astype(complex), w, s, ref, p, out=init.astype(complex), add=True, mode=mode)
    assert rel(g.cpu().numpy(), want) < TOL
    g0 = torch.zeros_like(g)
    op.adjoint(g0, _c(d))
    want0 = adj(d.astype(complex), bg.astype(complex), w, s, ref, p, mode=mode)
    assert rel(g0.cpu().numpy(), want0) < TOL


def test_reflect_dreflect_ic_match_oracle():
    import torch
    from wemfwi.unit import IC, Reflect, dReflect
    s = _slow(9)
    rng = np.random.default_rng(10)
    den = (1 + 0.3 * rng.random((NZ, NX))).astype(np.complex64)
    W = _rnd((2, NZ, NX), 11)
    r = O.reflect_coef(s, den)
    op = Reflect([_vec(s), _vec(den)])
    out = torch.zeros((2, NZ, NX), dtype=torch.complex64, device="cuda")
    op.forward(_c(W), out)
    assert rel(out.cpu().numpy(), r * W) < TOL
    op.adjoint(out, _c(W))
    assert rel(out.cpu().numpy(), np.conj(r) * W) < TOL
    dS = (_rnd((NZ, NX), 12) * 1e-8).astype(np.complex64)
    dR = (_rnd((NZ, NX), 13) * 0.1).astype(np.complex64)
    dop = dReflect([_vec(s), _vec(den)])
    dr = torch.zeros((NZ, NX), dtype=torch.complex64, device="cuda")
    dop.forward([_c(dS), _c(dR)], dr)
    assert rel(dr.cpu().numpy(), O.dreflect_fwd(s, den, dS, dR)) < TOL
    gS, gR = torch.zeros_like(dr), torch.zeros_like(dr)
    v = _rnd((NZ, NX), 14)
    dop.adjoint([gS, gR], _c(v))
    wS, wR = O.dreflect_adj(s, den, v)
    assert rel(gS.cpu().numpy(), wS) < TOL and rel(gR.cpu().numpy(), wR) < TOL
    bg = _rnd((NZ, NX), 15)
    ic = IC(_c(bg))
    mm = _rnd((NZ, NX), 16)
    dd = torch.zeros_like(dr)
    ic.forward(_c(mm), dd)
    assert rel(dd.cpu().numpy(), O.ic_fwd(bg, mm)) < TOL
    m2 = torch.zeros_like(dr)
    ic.adjoint(m2, _c(mm))
    assert rel(m2.cpu().numpy(), O.ic_adj(bg, mm)) < TOL
    ic.setDepth(3)                       # row form (IC::setDepth)
    row = torch.zeros(NX, dtype=torch.complex64, device="cuda")
    ic.forward(_c(mm[3]), row)
    assert rel(row.cpu().numpy(), bg[3].astype(complex) * mm[3]) < TOL


def test_injection_forms_match_oracle_and_adjoint_zeroes_model():
    import torch
    from wemfwi.unit import Injection
    nz, nx, tap, ng = 12, 60, 4, 3
    iz, ix = [0, 2, 5, 11], [1, 17, 30, 58]
    inj = Injection(iz, ix, ng, tap)
    data = _rnd((3, len(ix), 9), 17)
    inj.setShot(1)
    inj.setStep(4)
    W = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
    inj.forward(_c(data), W)
    want = O.inj_rec_fwd(data[1, :, 4].astype(complex), iz, ix, ng, tap, nz, nx)
    assert rel(W.cpu().numpy(), want) < TOL
    wf = _rnd((nz, nx), 18)
    got = _c(np.ones((3, len(ix), 9)))
    inj.adjoint(got, _c(wf))
    g = got.cpu().numpy()
    assert np.all(g[0] == 0) and np.all(g[2] == 0) and np.all(g[1, :, :4] == 0)   # whole model zeroed
    assert rel(g[1, :, 4], O.inj_rec_adj(wf.astype(complex), iz, ix, ng, tap)) < TOL
    wave = _rnd((9,), 19)
    W.zero_()
    inj.forward(_c(wave), W)
    assert rel(W.cpu().numpy(), O.inj_wavelet_fwd(complex(wave[4]), iz, ix, ng, tap, nz, nx)) < TOL


def test_unit_ssf_and_down_match_oracle():
    import torch
    from wemfwi.unit import SSF, Down
    s = _slow(20)
    p = _p()
    ref = O.refsampler(s, PAR["nref"])
    op = SSF(_vec(s), PAR)
    op.setFreq(9.0)
    op.setDepth(6)
    x = _rnd((3, NX), 21)
    y = torch.zeros((3, NX), dtype=torch.complex64, device="cuda")
    op.forward(_c(x), y)
    w = 2 * np.pi * np.float32(9.0)
    assert rel(y.cpu().numpy(), O.onestep_fwd(x.astype(complex), 6, w, s, ref, p)) < TOL
    dn = Down(_vec(s), PAR)
    dn.setFreq(9.0)
    Wt = _c(_rnd((NZ, NX), 22))
    W0 = Wt.cpu().numpy().astype(complex)
    dn.forward(Wt, Wt, add=True)          # in place: the recursion
    assert rel(Wt.cpu().numpy(), O.down_fwd(W0, w, s, ref, p)) < TOL


def test_tables_on_any_grid_and_propagation_rejects_unbuilt_fft_length():
    import torch
    from wemfwi._lib import WemfwiError
    from wemfwi.unit import Reflect, SSF
    rng = np.random.default_rng(23)
    nz, nx = 6, 134  # 2 x 67: a prime factor above the largest planned radix
    s = (1 / (2000.0 + 100 * rng.random((nz, nx))) ** 2).astype(np.complex64)
    den = np.ones((nz, nx), np.complex64)
    W = _rnd((nz, nx), 24)
    out = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
    Reflect([_vec(s), _vec(den)]).forward(_c(W), out)
    assert rel(out.cpu().numpy(), O.reflect_coef(s, den) * W) < TOL
    op = SSF(_vec(s), {"nref": 1})
    with pytest.raises(WemfwiError) as e:
        op.forward(_c(W[0]), out[0])
    assert e.value.code == 3
