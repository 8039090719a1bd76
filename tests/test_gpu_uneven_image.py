"""Subsurface-offset migration of shots with uneven apertures and bands in
one batched call (wemfwi_uneven_offset_image_adjoint): parity with the FP64
oracle (BornRefl's source and receiver wavefields on each shot's aperture,
ext_born_refl.cpp:60-110, with the offset imaging condition) placed at each
shot's image column, relative L2 <= 1e-5; FFT lengths 128, 256 and a
non-power-of-two aperture in one call; receivers below the surface and a
repeated receiver cell; argument checks."""

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
