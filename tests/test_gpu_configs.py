"""The two parity-test configurations at their full sizes.

Modeling + migration grid: nx = 256, nz = 200, dx = dz = 10 m, 64 frequencies
2 + i 18/63 Hz, 10 Hz Ricker source at the surface centre, 256 surface
receivers, velocity 1500 -> 3500 m/s plus a seeded Gaussian-smoothed (5 cells)
5 % field.  WEM data and the zero-lag migration image (BornRefl adjoint)
against the FP64 oracle at relative L2 <= 1e-5.

Born pair on the same grid: smoothed background (5 cells), a sparse
reflectivity (5 % nonzero, N(0, 0.1) S); Born forward against the oracle,
the exact dot test <Lm, d> = <m, L*d> on random complex m, d (seed 33; the
difference within 1e-5 of |Lm| |d|), and
the nh = 21 subsurface-offset image, whose h = 0 slice is the zero-lag image
summed over frequency and whose other lags obey the image's adjoint pairing
with offset_image_forward."""

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NX, NZ, NW = 256, 200, 64


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _velocity(seed=1):
    rng = np.random.default_rng(seed)
    z = np.arange(NZ)[:, None] * 10.0
    v = 1500.0 + 2000.0 * z / ((NZ - 1) * 10.0) + np.zeros((NZ, NX))
    pert = gaussian_filter(rng.standard_normal((NZ, NX)), 5.0)
    return v + 0.05 * v * pert / np.max(np.abs(pert))


def _ricker(freqs, f0=10.0, t0=0.1):
    f = np.asarray(freqs, float)
    a = (f / f0) ** 2
    return (2 / np.sqrt(np.pi) * a / f0 * np.exp(-a) * np.exp(-2j * np.pi * f * t0)).astype(np.complex64)


def _survey(refl=False):
    freq = (2.0 + np.arange(NW) * 18.0 / 63.0).astype(np.float32)
    return O.Survey(nx=NX, nz=NZ, dx=10.0, dz=10.0, freq=freq, index=np.arange(NW), wave_f=_ricker(freq),
                    src_iz=[0], src_ix=[NX // 2], rec_iz=[0] * NX, rec_ix=list(range(NX)), nt_f=NW,
                    tap=0, pad=0, nref=1, eps=0.04, prop="ssf", ntaylor=1, ngs=0, ngr=0, ng=0)


def _gpu(sv, slow, den=None):
    from wemfwi import Model
    from wemfwi.ops import Survey
    m = Model(slow.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
              eps=sv.eps, ntaylor=sv.ntaylor, den=None if den is None else den.astype(np.complex64))
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
               ngs=sv.ngs, ngr=sv.ngr, offset=0)
    return m, s


def _c(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex64))).cuda()


def test_modeling_and_migration_full_grid():
    import torch
    from wemfwi.ops import bornrefl_adjoint, wem_forward
    sv = _survey()
    slow = (1 / _velocity() ** 2).astype(np.complex64)
    den = np.ones((NZ, NX), np.complex64)
    m, s = _gpu(sv, slow, den)
    print("modeling 256 x 200, 64 frequencies", flush=True)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m, s, d)
    want = O.wem_forward(sv, slow, den)
    assert rel(d.cpu().numpy(), want) < TOL
    print("migration (zero-lag image, stacked over frequency)", flush=True)
    img = torch.zeros((NW, NZ, NX), dtype=torch.complex64, device="cuda")
    bornrefl_adjoint(m, s, img, d)
    wimg = O.bornrefl_adjoint(sv, slow, d.cpu().numpy().astype(complex))
    assert rel(img.sum(0).cpu().numpy(), wimg.sum(0)) < TOL
    assert rel(img.cpu().numpy(), wimg) < TOL


def test_born_pair_full_grid_oracle_dot_and_offset_image():
    import torch
    from wemfwi.ops import (born_adjoint, born_forward, bornrefl_adjoint, offset_image_adjoint,
                            offset_image_forward)
    sv = _survey()
    v = _velocity()
    slow0 = (1 / gaussian_filter(v, 5.0) ** 2).astype(np.complex64)
    den = np.ones((NZ, NX), np.complex64)
    rng = np.random.default_rng(3)
    mask = rng.random((NZ, NX)) < 0.05
    dS = (mask * rng.normal(0.0, 0.1, (NZ, NX)) * slow0.real).astype(np.complex64)
    dR = np.zeros_like(dS)
    m, s = _gpu(sv, slow0, den)
    print("Born forward, sparse reflectivity", flush=True)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(dS), _c(dR), d)
    want = O.born_forward(sv, slow0, den, dS, dR)
    assert rel(d.cpu().numpy(), want) < TOL
    # dot test on random complex m, d (seed 33), exact adjoint
    r33 = np.random.default_rng(33)
    mS = r33.standard_normal((NZ, NX)) + 1j * r33.standard_normal((NZ, NX))
    mR = r33.standard_normal((NZ, NX)) + 1j * r33.standard_normal((NZ, NX))
    dd = r33.standard_normal(s.data_shape) + 1j * r33.standard_normal(s.data_shape)
    Lm = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(mS), _c(mR), Lm)
    gS = torch.zeros((NZ, NX), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, _c(dd), exact=True)
    lhs = O.cdot(Lm.cpu().numpy().astype(complex), dd.astype(np.complex64).astype(complex))
    rhs = O.cdot(mS.astype(np.complex64).astype(complex), gS.cpu().numpy().astype(complex)) + O.cdot(
        mR.astype(np.complex64).astype(complex), gR.cpu().numpy().astype(complex))
    # random m and d make <Lm, d> a sum of 16 k random-phase terms, |lhs| ~
    # |Lm| |d| / 128: the FP32 rounding of L and L* is measured against the
    # Cauchy-Schwarz bound |Lm| |d| (1e-5 of it), the plain ratio printed
    scale = float(np.linalg.norm(Lm.cpu().numpy())) * float(np.linalg.norm(dd))
    print(f"dot test |lhs - rhs| / |lhs| = {abs(lhs - rhs) / abs(lhs):.2e}, "
          f"/ (|Lm| |d|) = {abs(lhs - rhs) / scale:.2e}", flush=True)
    assert abs(lhs - rhs) <= 1e-5 * scale
    # subsurface-offset image, nh = 21, of the Born data
    nh = 21
    img = torch.zeros((nh, NZ, NX), dtype=torch.complex64, device="cuda")
    offset_image_adjoint(m, s, nh, img, d)
    zl = torch.zeros((NW, NZ, NX), dtype=torch.complex64, device="cuda")
    bornrefl_adjoint(m, s, zl, d)
    assert rel(img[nh // 2].cpu().numpy(), zl.sum(0).cpu().numpy()) < TOL
    Ih = (r33.standard_normal((nh, NZ, NX)) + 1j * r33.standard_normal((nh, NZ, NX))).astype(np.complex64)
    dI = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    offset_image_forward(m, s, nh, _c(Ih), dI)
    lhs = O.cdot(dI.cpu().numpy().astype(complex), d.cpu().numpy().astype(complex))
    rhs = O.cdot(Ih.astype(complex), img.cpu().numpy().astype(complex))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
