"""Calls cut into several chunks of (shot, frequency) rows (WEMFWI_CHUNK_ROWS
caps the rows per chunk) give the same results as one chunk: bitwise for
the per-row operators and, because reductions sum rows in runs of 64 of the
call's row order and chunks hold whole runs, bitwise for the adjoints too
(also with wavefield checkpointing).  A cap below one run splits runs and
only changes the rounding of the FP64 sums."""

import os

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu


def _case(nx=100, nz=24, pad=14, nw=16, ns=10, seed=31, nr=12):
    rng = np.random.default_rng(seed)
    nt = 64
    freq = (2.0 + np.arange(nw) * 1.5).astype(np.float32)
    index = np.arange(nw) + 2
    wave_f = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freq, index=index, wave_f=wave_f,
                  src_iz=[0] * ns, src_ix=list(np.linspace(5, nx - 6, ns).astype(int)),
                  rec_iz=[0] * (nr - 3) + [4] * 3, rec_ix=list(np.linspace(0, nx - 1, nr).astype(int)),
                  nt_f=nt, tap=0, pad=pad, nref=2, eps=0.04, prop="pspi", ntaylor=1, ngs=1, ngr=1, ng=1)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1200 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    slow = (1 / v ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    data = np.zeros((ns, nr, nt), complex)
    data[:, :, index] = rng.standard_normal((ns, nr, nw)) + 1j * rng.standard_normal((ns, nr, nw))
    return sv, slow, den, data


def _setup(sv, slow, den):
    from wemfwi import Model
    from wemfwi.ops import Survey
    m = Model(O.edge_pad(slow, sv.pad).astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref,
              prop=sv.prop, eps=sv.eps, ntaylor=sv.ntaylor, den=O.edge_pad(den, sv.pad).astype(np.complex64))
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
               ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    return m, s


def _with_cap(cap, fn):
    old = os.environ.get("WEMFWI_CHUNK_ROWS")
    if cap is None:
        os.environ.pop("WEMFWI_CHUNK_ROWS", None)
    else:
        os.environ["WEMFWI_CHUNK_ROWS"] = str(cap)
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("WEMFWI_CHUNK_ROWS", None)
        else:
            os.environ["WEMFWI_CHUNK_ROWS"] = old


def _c(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).cuda()


def test_forward_operators_bitwise_over_chunks():
    import torch
    from wemfwi.ops import born_forward, wem_forward
    sv, slow, den, _ = _case()
    m, s = _setup(sv, slow, den)
    rng = np.random.default_rng(1)
    nxp = sv.nx + 2 * sv.pad
    dS = _c(O.zero_pad(rng.standard_normal((sv.nz, sv.nx)) * 1e-8, sv.pad))
    dR = _c(O.zero_pad(rng.standard_normal((sv.nz, sv.nx)) * 0.1, sv.pad))
    assert dS.shape == (sv.nz, nxp)

    def run():
        d1 = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
        d2 = torch.zeros_like(d1)
        wem_forward(m, s, d1)
        born_forward(m, s, dS, dR, d2)
        return d1, d2

    base = _with_cap(None, run)
    for cap in (7, 64, 100):
        got = _with_cap(cap, run)
        print(f"forward cap={cap}", flush=True)
        assert torch.equal(got[0], base[0]) and torch.equal(got[1], base[1]), cap


@pytest.mark.parametrize("every", [0, 4])
def test_born_adjoint_bitwise_over_whole_runs(every):
    import torch
    from wemfwi.ops import born_adjoint
    sv, slow, den, data = _case()
    m, s = _setup(sv, slow, den)
    s.set_checkpoint(every)
    d = _c(data)
    shape = (sv.nz, sv.nx + 2 * sv.pad)

    def run():
        gS = torch.zeros(shape, dtype=torch.complex64, device="cuda")
        gR = torch.zeros_like(gS)
        born_adjoint(m, s, gS, gR, d, exact=True)
        return gS, gR

    base = _with_cap(None, run)
    # 160 rows: chunks of 64 / 128 rows hold whole runs; 100 rounds down to 64
    for cap in (64, 100, 128):
        got = _with_cap(cap, run)
        print(f"born_adjoint every={every} cap={cap}", flush=True)
        assert torch.equal(got[0], base[0]) and torch.equal(got[1], base[1]), cap
    # chunks of 30 rows split runs: the FP64 sums round differently only
    got = _with_cap(30, run)
    for g, b in zip(got, base):
        assert float((g - b).abs().max()) <= 1e-6 * float(b.abs().max())


def test_extborn_and_image_adjoints_bitwise_over_whole_runs():
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey, bornrefl_adjoint, extborn_adjoint, offset_image_adjoint
    sv, slow, den, data = _case(nw=8, ns=20)
    nw = len(sv.freq)
    ms = [Model(O.edge_pad(slow * (1 + 0.01 * i), sv.pad).astype(np.complex64), sv.dx, sv.dz, tap=sv.tap,
                pad=sv.pad, nref=sv.nref, prop=sv.prop, eps=sv.eps, den=O.edge_pad(den, sv.pad).astype(np.complex64))
          for i in range(nw)]
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
               ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    # the imaging grid is unpadded: a taper of 14 makes its FFT length 128
    mr = Model(slow, sv.dx, sv.dz, tap=14, pad=sv.pad, nref=sv.nref, prop=sv.prop, eps=sv.eps)
    sr = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
                ngs=sv.ng, ngr=0, offset=0)
    d = _c(data)
    nxp = sv.nx + 2 * sv.pad

    def run():
        gS = torch.zeros((nw, sv.nz, nxp), dtype=torch.complex64, device="cuda")
        gR = torch.zeros_like(gS)
        extborn_adjoint(ms, s, gS, gR, d, exact=False)
        img = torch.zeros((nw, sv.nz, sv.nx), dtype=torch.complex64, device="cuda")
        bornrefl_adjoint(mr, sr, img, d)
        oimg = torch.zeros((3, sv.nz, sv.nx), dtype=torch.complex64, device="cuda")
        offset_image_adjoint(mr, sr, 3, oimg, d)
        return gS, gR, img, oimg

    base = _with_cap(None, run)
    for cap in (64, 128):
        got = _with_cap(cap, run)
        print(f"ext/image adjoints cap={cap}", flush=True)
        for g, b in zip(got, base):
            assert torch.equal(g, b), cap
