"""The two parity-test configurations at their full sizes.

Modeling + migration grid: nx = 256, nz = 200, dx = dz = 10 m, 64 frequencies
2 + i 18/63 Hz, 10 Hz Ricker source at the surface centre, 256 surface
receivers, velocity 1500 -> 3500 m/s plus a seeded Gaussian-smoothed (5 cells)
5 % field.  WEM data and the zero-lag migration image (BornRefl adjoint)
against the FP64 oracle at relative L2 <= 1e-5.

Born pair on the same grid: smoothed background (5 cells), a sparse
reflectivity (5 % nonzero, N(0, 0.1) S); Born forward against the oracle,
the exact dot test <Lm, d> = <m, L*d> on random complex m (seed 33) and d =
Lm plus a random part of equal norm (relative difference <= 1e-5), and
the nh = 21 subsurface-offset image, whose h = 0 slice is the zero-lag image
summed over frequency and whose other lags obey the image's adjoint pairing
with offset_image_forward.

The uneven survey (64 shots, apertures 256-2048, 32-192 frequencies, nz = 400,
nh = 11) at full size through size-independent properties of the extended
least-squares gradient."""

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
    Lm = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(mS), _c(mR), Lm)
    # d = Lm (scaled) plus an equally large random part: <Lm, d> is then of
    # the order of |Lm| |d|, so the plain relative difference measures the
    # FP32 rounding of L and L* (with an uncorrelated d, <Lm, d> is a sum of
    # 16 k random-phase terms, about |Lm| |d| / 128, and the ratio would
    # show that rounding 128 times over)
    noise = r33.standard_normal(s.data_shape) + 1j * r33.standard_normal(s.data_shape)
    lm = Lm.cpu().numpy().astype(complex)
    dd = lm / np.linalg.norm(lm) * np.linalg.norm(noise) + noise
    gS = torch.zeros((NZ, NX), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, _c(dd), exact=True)
    lhs = O.cdot(Lm.cpu().numpy().astype(complex), dd.astype(np.complex64).astype(complex))
    rhs = O.cdot(mS.astype(np.complex64).astype(complex), gS.cpu().numpy().astype(complex)) + O.cdot(
        mR.astype(np.complex64).astype(complex), gR.cpu().numpy().astype(complex))
    scale = float(np.linalg.norm(lm)) * float(np.linalg.norm(dd))
    print(f"dot test |lhs - rhs| / |lhs| = {abs(lhs - rhs) / abs(lhs):.2e}, "
          f"|lhs| / (|Lm| |d|) = {abs(lhs) / scale:.2f}", flush=True)
    assert abs(lhs) >= 0.5 * scale
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)
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


def _uneven_full():
    """The uneven survey at full size: 64 shots over a 4096-wide, 400-deep
    model (1500 -> 3500 m/s plus a smoothed 5 % field, seed 5), apertures
    drawn from {256, 512, 1024, 2048}, 32-192 frequencies from 2-5 Hz in
    0.1 Hz steps, receivers on every second aperture column, random data
    (seed 55): 7308 (shot, frequency) rows."""
    import torch
    from wemfwi import Model
    from wemfwi.ops import ImageShot
    rng = np.random.default_rng(5)
    nz, W0 = 400, 4096
    z = np.arange(nz)[:, None] * 10.0
    v = 1500.0 + 2000.0 * z / ((nz - 1) * 10.0) + np.zeros((nz, W0))
    pert = gaussian_filter(rng.standard_normal((nz, W0)), 8.0)
    slow = (1.0 / (v + 0.05 * v * pert / np.max(np.abs(pert))) ** 2).astype(np.complex64)
    geo = []
    for _ in range(64):
        nx = int(rng.choice((256, 512, 1024, 2048)))
        nw = int(rng.integers(32, 193))
        f0 = 2.0 + 3.0 * rng.random()
        cx = int(rng.integers(nx // 2, W0 - nx // 2 + 1))
        geo.append((cx - nx // 2, nx, [float(np.float32(f0 + 0.1 * i)) for i in range(nw)]))
    drng = np.random.default_rng(55)
    shots = []
    for x0, nx, freqs in geo:
        m = Model(np.ascontiguousarray(slow[:, x0:x0 + nx]), 10.0, 10.0, nref=1, prop="ssf", eps=0.04)
        rix = np.arange(0, nx, 2)
        nw = len(freqs)
        d = (drng.standard_normal((len(rix), nw)) + 1j * drng.standard_normal((len(rix), nw))).astype(np.complex64)
        shots.append(ImageShot(m, x0, freqs, _ricker(freqs), 0, nx // 2, np.zeros_like(rix), rix,
                               torch.from_numpy(d).cuda()))
    return shots, nz, W0


def test_uneven_extended_gradient_full_size_properties():
    """At the full uneven size (nh = 11) the oracle is out of reach; the
    gradient is checked through properties that hold at any size: at I = 0
    it is minus the migration of the observed data, bitwise, with J = 1/2
    ||d||^2; with zero observed data <I, L^H L I> = 2 J (1e-5 relative, the
    FP32 rounding of L and L^H); repeated calls give the same J and
    gradient bitwise."""
    import torch
    from wemfwi.ops import uneven_offset_gradient, uneven_offset_image_adjoint
    shots, nz, W0 = _uneven_full()
    nh = 11
    print("uneven full size: migration and gradient at I = 0", flush=True)
    mig = torch.zeros((nh, nz, W0), dtype=torch.complex64, device="cuda")
    uneven_offset_image_adjoint(shots, nh, mig, ngs=0)
    g0 = torch.zeros_like(mig)
    J0 = uneven_offset_gradient(shots, nh, torch.zeros_like(mig), g0, ngs=0)
    assert torch.equal(g0, -mig)
    want = 0.5 * sum(float(torch.sum(sh.data.to(torch.complex128).abs() ** 2)) for sh in shots)
    assert abs(J0 - want) <= 1e-12 * want
    print("uneven full size: <I, L^H L I> = 2 J", flush=True)
    I = mig / torch.linalg.vector_norm(mig)
    for sh in shots:
        sh.data.zero_()
    g = torch.zeros_like(mig)
    J = uneven_offset_gradient(shots, nh, I, g, ngs=0)
    ip = float((I.to(torch.complex128) * g.to(torch.complex128).conj()).real.sum())
    print(f"<I, g> = {ip:.10e}, 2 J = {2 * J:.10e}", flush=True)
    assert abs(ip - 2 * J) <= 1e-5 * 2 * J
    g2 = torch.zeros_like(mig)
    assert uneven_offset_gradient(shots, nh, I, g2, ngs=0) == J and torch.equal(g, g2)
    for sh in shots:
        sh.model.close()
