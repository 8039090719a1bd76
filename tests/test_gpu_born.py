"""GPU parity of the shot x frequency operators (WEM, BornFull, BornRefl,
extWEM, extBornFull, subsurface-offset image) against the FP64 oracle, plus
the reference's own test invariants (dot products, ext == plain,
propagator/test/test_wem_ops.py:256-295)."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _case(nx=100, nz=40, pad=14, tap=0, nw=5, ns=2, nref=1, prop="ssf", seed=11, nr=12):
    rng = np.random.default_rng(seed)
    nt = 64
    freq = (2.0 + np.arange(nw) * 3.25).astype(np.float32)
    index = np.arange(nw) + 3
    wave_f = (rng.standard_normal(nt) + 1j * rng.standard_normal(nt))
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freq, index=index, wave_f=wave_f,
                  src_iz=[0] * ns, src_ix=list(np.linspace(10, nx - 11, ns).astype(int)),
                  rec_iz=[0] * (nr // 2) + [3] * (nr - nr // 2),
                  rec_ix=list(np.linspace(0, nx - 1, nr).astype(int)),
                  nt_f=nt, tap=tap, pad=pad, nref=nref, eps=0.04, prop=prop, ntaylor=1,
                  ngs=2, ngr=1, ng=2)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.04 * rng.standard_normal((nz, nx)))
    slow = (1 / v ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    return sv, slow, den


def _gpu(sv, slow, den, refl=False):
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey
    if refl:
        m = Model(slow, sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
                  eps=sv.eps, ntaylor=sv.ntaylor)
        s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index,
                   sv.wave_f[sv.index], sv.nt_f, ngs=sv.ng, ngr=0, offset=0)
    else:
        sp, dp = O.edge_pad(slow, sv.pad), O.edge_pad(den, sv.pad)
        m = Model(sp.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref,
                  prop=sv.prop, eps=sv.eps, ntaylor=sv.ntaylor, den=dp.astype(np.complex64))
        s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index,
                   sv.wave_f[sv.index], sv.nt_f, ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    return torch, m, s


def _c(x, torch):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).cuda()


@pytest.mark.parametrize("nref,prop", [(1, "ssf"), (2, "pspi"), (3, "ssf")])
def test_wem_forward_matches_oracle(nref, prop):
    sv, slow, den = _case(nref=nref, prop=prop)
    torch, m, s = _gpu(sv, slow, den)
    from wemfwi.ops import wem_forward
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m, s, d)
    want = O.wem_forward(sv, slow, den)
    got = d.cpu().numpy()
    assert rel(got[:, :, sv.index], want[:, :, sv.index]) < TOL
    mask = np.ones(sv.nt_f, bool)
    mask[sv.index] = False
    assert np.all(got[:, :, mask] == 0)


def test_born_forward_matches_oracle():
    sv, slow, den = _case()
    torch, m, s = _gpu(sv, slow, den)
    from wemfwi.ops import born_forward
    rng = np.random.default_rng(3)
    dS = (rng.standard_normal((sv.nz, sv.nx)) * 1e-8).astype(np.complex64)
    dR = (rng.standard_normal((sv.nz, sv.nx)) * 0.1).astype(np.complex64)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(O.zero_pad(dS, sv.pad), torch), _c(O.zero_pad(dR, sv.pad), torch), d)
    want = O.born_forward(sv, slow, den, dS, dR)
    assert rel(d.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL


@pytest.mark.parametrize("exact", [False, True])
def test_born_adjoint_matches_oracle(exact):
    sv, slow, den = _case(nref=2, prop="pspi")
    torch, m, s = _gpu(sv, slow, den)
    from wemfwi.ops import born_adjoint
    rng = np.random.default_rng(4)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, len(sv.index))) + 1j * rng.standard_normal(
        (sv.ns, sv.nr, len(sv.index)))
    nxp = sv.nx + 2 * sv.pad
    gS = torch.zeros((sv.nz, nxp), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, _c(data, torch), exact=exact)
    wS, wR = O.born_adjoint(sv, slow, den, data, mode="exact" if exact else "compat")
    sl = slice(sv.pad, sv.pad + sv.nx)
    assert rel(gS.cpu().numpy()[:, sl], wS) < TOL
    assert rel(gR.cpu().numpy()[:, sl], wR) < TOL


def test_born_dot_product_exact_on_gpu():
    sv, slow, den = _case(nw=4, ns=3)
    torch, m, s = _gpu(sv, slow, den)
    from wemfwi.ops import born_adjoint, born_forward
    rng = np.random.default_rng(5)
    nxp = sv.nx + 2 * sv.pad
    dS = np.zeros((sv.nz, nxp), complex)
    dR = np.zeros((sv.nz, nxp), complex)
    dS[:, sv.pad:sv.pad + sv.nx] = rng.standard_normal((sv.nz, sv.nx)) + 1j * rng.standard_normal((sv.nz, sv.nx))
    dR[:, sv.pad:sv.pad + sv.nx] = rng.standard_normal((sv.nz, sv.nx)) + 1j * rng.standard_normal((sv.nz, sv.nx))
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, len(sv.index))) + 1j * rng.standard_normal(
        (sv.ns, sv.nr, len(sv.index)))
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(dS, torch), _c(dR, torch), d)
    gS = torch.zeros((sv.nz, nxp), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, _c(data, torch), exact=True)
    lhs = O.cdot(d.cpu().numpy().astype(complex), data)
    sl = slice(sv.pad, sv.pad + sv.nx)
    rhs = O.cdot(dS[:, sl], gS.cpu().numpy()[:, sl].astype(complex)) + O.cdot(
        dR[:, sl], gR.cpu().numpy()[:, sl].astype(complex))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_bornrefl_matches_oracle_and_dot_product():
    sv, slow, den = _case(nx=120, pad=6, tap=4, nw=4, ns=3)
    torch, m, s = _gpu(sv, slow, den, refl=True)
    from wemfwi.ops import bornrefl_adjoint, bornrefl_forward
    rng = np.random.default_rng(6)
    nw = len(sv.freq)
    img = rng.standard_normal((nw, sv.nz, sv.nx)) + 1j * rng.standard_normal((nw, sv.nz, sv.nx))
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, nw)) + 1j * rng.standard_normal((sv.ns, sv.nr, nw))
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    bornrefl_forward(m, s, _c(img, torch), d)
    want = O.bornrefl_forward(sv, slow, img)
    assert rel(d.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL
    gi = torch.zeros((nw, sv.nz, sv.nx), dtype=torch.complex64, device="cuda")
    bornrefl_adjoint(m, s, gi, _c(data, torch))
    wi = O.bornrefl_adjoint(sv, slow, data)
    assert rel(gi.cpu().numpy(), wi) < TOL
    lhs = O.cdot(d.cpu().numpy().astype(complex), data)
    rhs = O.cdot(img, gi.cpu().numpy().astype(complex))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_offset_image_zero_lag_oracle_and_dot_product():
    sv, slow, den = _case(nx=120, pad=6, tap=4, nw=3, ns=2)
    torch, m, s = _gpu(sv, slow, den, refl=True)
    from wemfwi.ops import bornrefl_adjoint, offset_image_adjoint, offset_image_forward
    nh = 5
    rng = np.random.default_rng(7)
    nw = len(sv.freq)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, nw)) + 1j * rng.standard_normal((sv.ns, sv.nr, nw))
    I = torch.zeros((nh, sv.nz, sv.nx), dtype=torch.complex64, device="cuda")
    offset_image_adjoint(m, s, nh, I, _c(data, torch))
    want = np.zeros((nh, sv.nz, sv.nx), complex)
    for _, _, bg, sc in O.bornrefl_wavefields(sv, slow, data):
        want += O.offset_ic_adj(bg, sc, nh)
    got = I.cpu().numpy()
    assert rel(got, want) < TOL
    # h = 0 equals the zero-lag image stacked over frequency
    gi = torch.zeros((nw, sv.nz, sv.nx), dtype=torch.complex64, device="cuda")
    bornrefl_adjoint(m, s, gi, _c(data, torch))
    assert rel(got[nh // 2], gi.cpu().numpy().sum(0)) < TOL
    # forward / adjoint pair
    img = rng.standard_normal((nh, sv.nz, sv.nx)) + 1j * rng.standard_normal((nh, sv.nz, sv.nx))
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    offset_image_forward(m, s, nh, _c(img, torch), d)
    lhs = O.cdot(d.cpu().numpy().astype(complex), data)
    rhs = O.cdot(img, got.astype(complex))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_extwem_equals_wem_exactly():
    # propagator/test/test_wem_ops.py:268-282: bitwise equality
    sv, slow, den = _case(nw=3, ns=2)
    torch, m, s = _gpu(sv, slow, den)
    from wemfwi import Model
    from wemfwi.ops import extwem_forward, wem_forward
    sp, dp = O.edge_pad(slow, sv.pad), O.edge_pad(den, sv.pad)
    ms = [Model(sp.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
                eps=sv.eps, den=dp.astype(np.complex64)) for _ in sv.freq]
    d1 = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    d2 = torch.zeros_like(d1)
    wem_forward(m, s, d1)
    extwem_forward(ms, s, d2)
    assert torch.equal(d1, d2)


@pytest.mark.parametrize("exact", [False, True])
def test_extborn_matches_oracle(exact):
    sv, slow, den = _case(nw=3, ns=3, nz=30)
    rng = np.random.default_rng(9)
    nw = len(sv.freq)
    s3 = np.stack([slow * (1 + 0.01 * i) for i in range(nw)]).astype(np.complex64)
    d3 = np.stack([den] * nw)
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey, extborn_adjoint, extborn_forward
    ms = [Model(O.edge_pad(s3[i], sv.pad).astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad,
                nref=sv.nref, prop=sv.prop, eps=sv.eps, den=O.edge_pad(d3[i], sv.pad).astype(np.complex64))
          for i in range(nw)]
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index],
               sv.nt_f, ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    dS = rng.standard_normal((nw, sv.nz, sv.nx)) * 1e-8
    dR = rng.standard_normal((nw, sv.nz, sv.nx)) * 0.1
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    extborn_forward(ms, s, _c(O.zero_pad(dS, sv.pad), torch), _c(O.zero_pad(dR, sv.pad), torch), d)
    want = O.extborn_forward(sv, s3, d3, dS, dR)
    assert rel(d.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, nw)) + 1j * rng.standard_normal((sv.ns, sv.nr, nw))
    nxp = sv.nx + 2 * sv.pad
    gS = torch.zeros((nw, sv.nz, nxp), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    extborn_adjoint(ms, s, gS, gR, _c(data, torch), exact=exact)
    wS, wR = O.extborn_adjoint(sv, s3, d3, data, mode="exact" if exact else "compat")
    sl = slice(sv.pad, sv.pad + sv.nx)
    assert rel(gS.cpu().numpy()[:, :, sl], wS) < TOL
    assert rel(gR.cpu().numpy()[:, :, sl], wR) < TOL
