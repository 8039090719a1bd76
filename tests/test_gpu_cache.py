"""The library's cache of per-call device buffers (wemfwi_cache_bytes /
wemfwi_cache_release): buffers freed by a call are kept and reused, reused
blocks (holding the previous call's values) give bitwise the same results as
fresh ones, and releasing the cache returns every block."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NXI = 16, 200


def _shots(torch):
    from wemfwi import Model
    from wemfwi.ops import ImageShot
    rng = np.random.default_rng(2)
    z = np.linspace(0, 1, NZ)[:, None]
    world = (1 / ((1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((NZ, NXI)))) ** 2).astype(np.complex64)
    shots = []
    for x0, nx, nw in ((0, 128, 3), (60, 140, 4)):
        m = Model(np.ascontiguousarray(world[:, x0:x0 + nx]), 10.0, 10.0, nref=2, prop="pspi", eps=0.04)
        rix = list(range(0, nx, 3))
        d = (rng.standard_normal((len(rix), nw)) + 1j * rng.standard_normal((len(rix), nw))).astype(np.complex64)
        shots.append(ImageShot(m, x0, 3.0 + 2.0 * np.arange(nw), np.ones(nw), 0, nx // 2, [0] * len(rix), rix,
                               torch.from_numpy(d).cuda()))
    return shots


def test_cached_blocks_are_reused_and_released():
    import torch
    import wemfwi
    from wemfwi.ops import uneven_offset_gradient
    shots = _shots(torch)
    nh = 5
    I = (torch.randn((nh, NZ, NXI), dtype=torch.complex64, device="cuda", generator=None) * 0.1)
    g1 = torch.zeros((nh, NZ, NXI), dtype=torch.complex64, device="cuda")
    wemfwi.empty_cache()
    assert wemfwi.cache_bytes() == 0
    J1 = uneven_offset_gradient(shots, nh, I, g1)
    held = wemfwi.cache_bytes()
    print(f"cached after one call: {held} bytes", flush=True)
    assert held > 0
    for _ in range(3):                       # reused blocks, stale contents
        g2 = torch.zeros_like(g1)
        J2 = uneven_offset_gradient(shots, nh, I, g2)
        assert J2 == J1 and torch.equal(g1, g2)
        assert wemfwi.cache_bytes() == held  # the same blocks went back
    wemfwi.empty_cache()
    assert wemfwi.cache_bytes() == 0
    g3 = torch.zeros_like(g1)
    assert uneven_offset_gradient(shots, nh, I, g3) == J1 and torch.equal(g1, g3)


def test_stale_blocks_with_devices_below_the_surface():
    """The sweeps read only the source / receiver depths as sources and the
    drivers clear only those rows: with the cache holding blocks full of
    non-zero values (a warm-up call of the same size fills them), WEM, the
    Born adjoint (store-all and checkpointed) and the BornRefl pair with
    sources at depths 2 and 5 and receivers at depths 3-6 still match the
    FP64 oracle (relative L2 <= 1e-5); a first row of a sweep that holds no
    source must be stored as zero, not left as found."""
    import torch
    import wemfwi
    from wemfwi import Model
    from wemfwi.ops import Survey, born_adjoint, bornrefl_adjoint, bornrefl_forward, wem_forward
    from oracle import wem_oracle as O
    rng = np.random.default_rng(21)
    nx, nz, nw, nt, pad = 100, 24, 4, 32, 14
    freq = (3.0 + np.arange(nw) * 2.5).astype(np.float32)
    index = np.arange(nw) + 2
    wave_f = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    rix = list(range(0, nx, 7))
    riz = [3 + (i % 4) for i in range(len(rix))]
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freq, index=index, wave_f=wave_f, src_iz=[2, 5],
                  src_ix=[30, 70], rec_iz=riz, rec_ix=rix, nt_f=nt, tap=0, pad=pad, nref=2, eps=0.04,
                  prop="pspi", ntaylor=1, ngs=1, ngr=1, ng=1)
    z = np.linspace(0, 1, nz)[:, None]
    slow = (1 / ((1500 + 1500 * z) * (1 + 0.04 * rng.standard_normal((nz, nx)))) ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    sp, dp = O.edge_pad(slow, pad), O.edge_pad(den, pad)
    m = Model(sp.astype(np.complex64), 10.0, 10.0, pad=pad, nref=2, prop="pspi", eps=0.04,
              den=dp.astype(np.complex64))
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], nt, ngs=1,
               ngr=1, offset=pad)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m, s, d)                 # warm-up: the freed buffers go to the cache non-zero
    gS = torch.zeros((nz, nx + 2 * pad), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, d)
    assert wemfwi.cache_bytes() > 0
    print("WEM and Born adjoint on reused blocks", flush=True)
    d2 = torch.zeros_like(d)
    wem_forward(m, s, d2)
    want = O.wem_forward(sv, slow, den)
    assert rel(d2.cpu().numpy()[:, :, index], want[:, :, index]) < TOL
    res = np.zeros(want.shape, complex)
    res[:, :, index] = rng.standard_normal((2, len(rix), nw)) + 1j * rng.standard_normal((2, len(rix), nw))
    wS, wR = O.born_adjoint(sv, slow, den, res, mode="compat")
    for k in (0, 5):
        s.set_checkpoint(k)
        gS.zero_()
        gR.zero_()
        born_adjoint(m, s, gS, gR, torch.from_numpy(res.astype(np.complex64)).cuda())
        assert rel(gS.cpu().numpy()[:, pad:nx + pad], wS) < TOL, k
        assert rel(gR.cpu().numpy()[:, pad:nx + pad], wR) < TOL, k
    s.set_checkpoint(0)
    print("BornRefl pair on reused blocks", flush=True)
    mr = Model(slow, 10.0, 10.0, pad=pad, nref=2, prop="pspi", eps=0.04)   # BornRefl: taper pad, no edge pad
    sr = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], nt, ngs=1,
                ngr=0, offset=0)
    img = (rng.standard_normal((nw, nz, nx)) + 1j * rng.standard_normal((nw, nz, nx))).astype(np.complex64)
    dr = torch.zeros(sr.data_shape, dtype=torch.complex64, device="cuda")
    bornrefl_forward(mr, sr, torch.from_numpy(img).cuda(), dr)
    wantr = O.bornrefl_forward(sv, slow, img.astype(complex))
    assert rel(dr.cpu().numpy()[:, :, index], wantr[:, :, index]) < TOL
    ia = torch.zeros((nw, nz, nx), dtype=torch.complex64, device="cuda")
    bornrefl_adjoint(mr, sr, ia, torch.from_numpy(res.astype(np.complex64)).cuda())
    assert rel(ia.cpu().numpy(), O.bornrefl_adjoint(sv, slow, res)) < TOL
    for mm in (m, mr):
        mm.close()


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


TOL = 1e-5
