"""Subsurface-offset migration of shots with uneven apertures and bands in
one batched call (wemfwi_uneven_offset_image_adjoint): parity with the FP64
oracle (BornRefl's source and receiver wavefields on each shot's aperture,
ext_born_refl.cpp:60-110, with the offset imaging condition) placed at each
shot's image column, relative L2 <= 1e-5; FFT lengths 128, 256 and a
non-power-of-two aperture in one call; receivers below the surface and a
repeated receiver cell; argument checks.  The extended least-squares
objective of the same shots (wemfwi_uneven_offset_gradient): misfit and
gradient against the oracle's offset_objective, the gradient at I = 0 equal
to minus the migration bitwise, and <I, L^H L I> = 2 J."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NZ, NXI, DX, DZ = 20, 300, 10.0, 10.0


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _world(seed=3):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((NZ, NXI)))
    return (1 / v ** 2).astype(np.complex64)


SHOTS = [
    # x0, nx, freqs, src_ix (aperture), nref, prop
    (0, 100, [3.0, 6.5, 11.0], 30, 2, "pspi"),
    (40, 200, [2.5, 4.0, 7.25, 12.0, 16.5], 120, 1, "ssf"),
    (150, 150, [5.0, 9.5], 70, 2, "pspi"),      # FFT length 150 (runtime plan)
]


def _build(nh, ngs=1, seed=4):
    import torch
    from wemfwi import Model
    from wemfwi.ops import ImageShot
    world = _world()
    rng = np.random.default_rng(seed)
    shots, ref = [], []
    for (x0, nx, freqs, six, nref, prop) in SHOTS:
        s = np.ascontiguousarray(world[:, x0:x0 + nx])
        m = Model(s, DX, DZ, nref=nref, prop=prop, eps=0.04)
        nw = len(freqs)
        wav = (rng.standard_normal(nw) + 1j * rng.standard_normal(nw)).astype(np.complex64)
        rix = list(range(0, nx, 4)) + [nx // 2, nx // 2]      # a repeated cell
        riz = [0] * (len(rix) - 3) + [3, 5, 5]
        data = (rng.standard_normal((len(rix), nw)) + 1j * rng.standard_normal((len(rix), nw))).astype(np.complex64)
        shots.append(ImageShot(m, x0, freqs, wav, 0, six, riz, rix, torch.from_numpy(data).cuda()))
        sv = O.Survey(nx=nx, nz=NZ, dx=DX, dz=DZ, freq=np.asarray(freqs, np.float32), index=np.arange(nw),
                      wave_f=wav.astype(complex), src_iz=[0], src_ix=[six], rec_iz=riz, rec_ix=rix, nt_f=nw,
                      tap=0, pad=0, nref=nref, eps=0.04, prop=prop, ntaylor=1, ng=ngs)
        img = np.zeros((nh, NZ, nx), complex)
        for _, _, bg, sc in O.bornrefl_wavefields(sv, s, data[None].astype(complex)):
            img += O.offset_ic_adj(bg, sc, nh)
        ref.append((x0, nx, img))
    want = np.zeros((nh, NZ, NXI), complex)
    for x0, nx, img in ref:
        want[:, :, x0:x0 + nx] += img
    return shots, want


@pytest.mark.parametrize("nh", [1, 5, 11])
def test_uneven_offset_image_matches_oracle(nh):
    import torch
    from wemfwi.ops import uneven_offset_image_adjoint
    shots, want = _build(nh)
    print(f"uneven offset image nh={nh}", flush=True)
    img = torch.full((nh, NZ, NXI), 3.0, dtype=torch.complex64, device="cuda")
    uneven_offset_image_adjoint(shots, nh, img, ngs=1)
    assert rel(img.cpu().numpy(), want) < TOL
    # add = True accumulates
    uneven_offset_image_adjoint(shots, nh, img, ngs=1, add=True)
    assert rel(img.cpu().numpy(), 2 * want) < TOL


def test_uneven_offset_image_shot_by_shot_sums_to_the_batch():
    import torch
    from wemfwi.ops import uneven_offset_image_adjoint
    nh = 3
    shots, _ = _build(nh)
    full = torch.zeros((nh, NZ, NXI), dtype=torch.complex64, device="cuda")
    uneven_offset_image_adjoint(shots, nh, full, ngs=1)
    parts = torch.zeros_like(full)
    for sh in shots:
        uneven_offset_image_adjoint([sh], nh, parts, ngs=1, add=True)
    assert rel(parts.cpu().numpy(), full.cpu().numpy()) < 1e-6


def test_uneven_offset_image_rejects_bad_shots():
    import torch
    from wemfwi import Model
    from wemfwi._lib import WemfwiError
    from wemfwi.ops import ImageShot, uneven_offset_image_adjoint
    world = _world()
    m = Model(np.ascontiguousarray(world[:, :128]), DX, DZ)
    d = torch.zeros((2, 1), dtype=torch.complex64, device="cuda")
    img = torch.zeros((3, NZ, NXI), dtype=torch.complex64, device="cuda")
    for x0, rix in ((250, [0, 1]), (0, [0, 200])):  # aperture past the image; receiver off the aperture
        sh = ImageShot(m, x0, [5.0], [1.0], 0, 10, [0, 0], rix, d)
        with pytest.raises(WemfwiError) as e:
            uneven_offset_image_adjoint([sh], 3, img)
        assert e.value.code == 1
    with pytest.raises(WemfwiError):
        uneven_offset_image_adjoint([ImageShot(m, 0, [5.0], [1.0], 0, 10, [0, 0], [0, 1], d)], 4,
                                    torch.zeros((4, NZ, NXI), dtype=torch.complex64, device="cuda"))


def _oracle_objective(shots_np, nh, model, dobs):
    """Sum over shots of O.offset_objective on each shot's aperture."""
    J = 0.0
    grad = np.zeros((nh, NZ, NXI), complex)
    for (x0, nx, freqs, six, nref, prop, wav, riz, rix), d in zip(shots_np, dobs):
        world = _world()
        s = np.ascontiguousarray(world[:, x0:x0 + nx])
        nw = len(freqs)
        sv = O.Survey(nx=nx, nz=NZ, dx=DX, dz=DZ, freq=np.asarray(freqs, np.float32), index=np.arange(nw),
                      wave_f=wav.astype(complex), src_iz=[0], src_ix=[six], rec_iz=riz, rec_ix=rix, nt_f=nw,
                      tap=0, pad=0, nref=nref, eps=0.04, prop=prop, ntaylor=1, ng=1)
        j, g = O.offset_objective(sv, s, model[:, :, x0:x0 + nx], d[None].astype(complex))
        J += j
        grad[:, :, x0:x0 + nx] += g
    return J, grad


def _gradient_case(nh, seed=6):
    import torch
    from wemfwi import Model
    from wemfwi.ops import ImageShot
    world = _world()
    rng = np.random.default_rng(seed)
    shots, shots_np, dobs = [], [], []
    for (x0, nx, freqs, six, nref, prop) in SHOTS:
        m = Model(np.ascontiguousarray(world[:, x0:x0 + nx]), DX, DZ, nref=nref, prop=prop, eps=0.04)
        nw = len(freqs)
        wav = (rng.standard_normal(nw) + 1j * rng.standard_normal(nw)).astype(np.complex64)
        rix = list(range(1, nx, 5)) + [nx // 3, nx // 3]
        riz = [0] * (len(rix) - 2) + [4, 4]
        d = (rng.standard_normal((len(rix), nw)) + 1j * rng.standard_normal((len(rix), nw))).astype(np.complex64)
        shots.append(ImageShot(m, x0, freqs, wav, 0, six, riz, rix, torch.from_numpy(d).cuda()))
        shots_np.append((x0, nx, freqs, six, nref, prop, wav, riz, rix))
        dobs.append(d)
    # a model image on the image grid whose data are of the order of the observed data
    I = (rng.standard_normal((nh, NZ, NXI)) + 1j * rng.standard_normal((nh, NZ, NXI))).astype(np.complex64)
    return shots, shots_np, dobs, I


@pytest.mark.parametrize("nh", [1, 5, 11])
def test_uneven_offset_gradient_matches_oracle(nh):
    import torch
    from wemfwi.ops import uneven_offset_gradient
    shots, shots_np, dobs, I = _gradient_case(nh)
    print(f"uneven offset gradient nh={nh}", flush=True)
    g = torch.full((nh, NZ, NXI), 2.0, dtype=torch.complex64, device="cuda")
    J = uneven_offset_gradient(shots, nh, torch.from_numpy(I).cuda(), g, ngs=1)
    Jw, gw = _oracle_objective(shots_np, nh, I.astype(complex), dobs)
    assert abs(J - Jw) <= TOL * Jw, (J, Jw)
    assert rel(g.cpu().numpy(), gw) < TOL
    J2 = uneven_offset_gradient(shots, nh, torch.from_numpy(I).cuda(), g, ngs=1, add=True)
    assert J2 == J                                   # deterministic FP64 misfit
    assert rel(g.cpu().numpy(), 2 * gw) < TOL


def test_uneven_offset_gradient_properties():
    import torch
    from wemfwi.ops import uneven_offset_gradient, uneven_offset_image_adjoint
    nh = 5
    shots, _, dobs, I = _gradient_case(nh, seed=7)
    # I = 0: the residual is -d, the gradient minus the migration of d, bitwise
    g0 = torch.zeros((nh, NZ, NXI), dtype=torch.complex64, device="cuda")
    J0 = uneven_offset_gradient(shots, nh, torch.zeros_like(g0), g0, ngs=1)
    mig = torch.zeros_like(g0)
    uneven_offset_image_adjoint(shots, nh, mig, ngs=1)
    assert torch.equal(g0, -mig)
    want = 0.5 * sum(float(np.sum(np.abs(d.astype(complex)) ** 2)) for d in dobs)
    assert abs(J0 - want) <= 1e-12 * want
    # zero observed data: <I, L^H L I> = ||L I||^2 = 2 J
    for sh in shots:
        sh.data.zero_()
    It = torch.from_numpy(I).cuda()
    g = torch.zeros_like(g0)
    J = uneven_offset_gradient(shots, nh, It, g, ngs=1)
    ip = float((It.to(torch.complex128) * g.to(torch.complex128).conj()).real.sum())
    assert abs(ip - 2 * J) <= 1e-5 * 2 * J, (ip, 2 * J)
    from wemfwi._lib import WemfwiError
    with pytest.raises(WemfwiError):
        uneven_offset_gradient(shots, nh, It, It, ngs=1)           # model and grad alias
    with pytest.raises(ValueError):
        uneven_offset_gradient(shots, nh, It, torch.zeros((nh, NZ, NXI - 1), dtype=torch.complex64,
                                                          device="cuda"), ngs=1)
