"""GPU parity of the fused one-way depth step and the Down/Up recursions
against the FP64 oracle (relative L2 <= 1e-5, the tolerance of the hot path),
through the C ABI."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _slow(nz, nx, seed, lateral=0.05):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 2000 * z) * (1 + lateral * rng.standard_normal((nz, nx)))
    return (1 / v ** 2).astype(np.complex64)


CASES = [
    # nx, tap, pad(taper), nref, prop
    (128, 0, 0, 1, "ssf"),
    (200, 28, 10, 1, "ssf"),       # N = 256 with FFT padding and taper
    (256, 0, 0, 2, "pspi"),
    (500, 6, 20, 3, "ssf"),        # N = 512
    (1024, 0, 0, 4, "pspi"),       # N = 1024 (the PSPI downward-continuation shape)
    (1000, 12, 8, 1, "ssf"),       # N = 1024
    (2048, 0, 0, 1, "ssf"),        # N = 2048
]


def _setup(nx, tap, pad, nref, prop, nz, seed=0):
    import torch
    from wemfwi import Model
    s = _slow(nz, nx, seed)
    m = Model(s, 10.0, 10.0, tap=tap, pad=pad, nref=nref, prop=prop, eps=0.04)
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=pad, nref=nref, eps=0.04, prop=prop)
    ref = O.refsampler(s, nref)
    return torch, m, s, p, ref


@pytest.mark.parametrize("nx,tap,pad,nref,prop", CASES)
@pytest.mark.parametrize("adjoint", [False, True])
def test_onestep_matches_oracle(nx, tap, pad, nref, prop, adjoint):
    nz = 12
    torch, m, s, p, ref = _setup(nx, tap, pad, nref, prop, nz)
    rng = np.random.default_rng(1)
    nrows = 6
    x = (rng.standard_normal((nrows, nx)) + 1j * rng.standard_normal((nrows, nx))).astype(np.complex64)
    freqs = np.array([2.0, 5.5, 11.0, 17.25, 23.0, 30.0], dtype=np.float32)
    depth = np.array([0, 3, 5, 7, 9, 11], dtype=np.int32)
    out = torch.zeros((nrows, nx), dtype=torch.complex64, device="cuda")
    m.onestep(adjoint, freqs, depth, torch.from_numpy(x).cuda(), out)
    got = out.cpu().numpy()
    f = O.onestep_adj if adjoint else O.onestep_fwd
    for i in range(nrows):
        want = f(x[i].astype(complex), int(depth[i]), 2 * np.pi * float(freqs[i]), s, ref, p)
        assert rel(got[i], want) < TOL, (i, rel(got[i], want))


@pytest.mark.parametrize("nx,tap,pad,nref,prop", CASES)
@pytest.mark.parametrize("up", [False, True])
@pytest.mark.parametrize("adjoint", [False, True])
def test_oneway_inplace_matches_oracle(nx, tap, pad, nref, prop, up, adjoint):
    nz = 40
    torch, m, s, p, ref = _setup(nx, tap, pad, nref, prop, nz, seed=2)
    rng = np.random.default_rng(3)
    nrows = 3
    W = np.zeros((nrows, nz, nx), dtype=np.complex64)
    # sparse sources at a few depths, as injected wavefields are
    for r in range(nrows):
        for z0 in (0, nz // 2, nz - 1):
            W[r, z0, nx // 3: nx // 3 + 5] = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    freqs = np.array([3.0, 12.5, 27.0], dtype=np.float32)
    t = torch.from_numpy(W).cuda()
    m.oneway(up, adjoint, freqs, t)
    got = t.cpu().numpy()
    fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
          (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adjoint)]
    for i in range(nrows):
        want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
        assert rel(got[i], want) < TOL, (i, rel(got[i], want))


@pytest.mark.parametrize("up,adjoint", [(False, False), (True, True), (False, True)])
def test_oneway_out_of_place_and_add_false(up, adjoint):
    nx, nz = 256, 16
    torch, m, s, p, ref = _setup(nx, 0, 6, 2, "ssf", nz, seed=4)
    rng = np.random.default_rng(5)
    W = (rng.standard_normal((2, nz, nx)) + 1j * rng.standard_normal((2, nz, nx))).astype(np.complex64)
    freqs = np.array([4.0, 9.0], dtype=np.float32)
    src = torch.from_numpy(W).cuda()
    out = torch.full_like(src, 7.0)
    m.oneway(up, adjoint, freqs, src, out)
    fn = {(False, False): O.down_fwd, (True, True): O.up_adj, (False, True): O.down_adj}[(up, adjoint)]
    got = out.cpu().numpy()
    for i in range(2):
        want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p, inplace=False)
        assert rel(got[i], want) < TOL
    # in place with add=False: the reference zeroes the shared vector first
    m.oneway(up, adjoint, freqs, src, src, add=False)
    assert float(src.abs().max()) == 0.0


def test_oneway_pspi_reference_compat_partition_deep():
    """nz > nx/nref: the reference buffer fills and the partition becomes
    history dependent (RefSampler.cpp:63-81); the GPU path consumes the host
    RefSampler tables, so it must follow it at every depth."""
    nx, nz, nref = 128, 90, 4
    torch, m, s, p, ref = _setup(nx, 0, 0, nref, "pspi", nz, seed=6)
    assert len(np.unique(ref.refid[-1])) > 1
    W = np.zeros((2, nz, nx), dtype=np.complex64)
    W[:, 0, 60:68] = 1.0
    freqs = np.array([6.0, 19.0], dtype=np.float32)
    t = torch.from_numpy(W).cuda()
    m.oneway(False, False, freqs, t)
    got = t.cpu().numpy()
    for i in range(2):
        want = O.down_fwd(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
        assert rel(got[i], want) < TOL


@pytest.mark.parametrize("up,adjoint", [(False, False), (True, False), (False, True), (True, True)])
def test_oneway_recursive_equals_inplace_on_copy(up, adjoint):
    nx, nz = 512, 24
    torch, m, s, p, ref = _setup(nx, 0, 0, 3, "pspi", nz, seed=7)
    rng = np.random.default_rng(8)
    W = (rng.standard_normal((4, nz, nx)) + 1j * rng.standard_normal((4, nz, nx))).astype(np.complex64)
    freqs = np.array([2.0, 7.0, 15.0, 29.0], dtype=np.float32)
    src = torch.from_numpy(W).cuda()
    out = torch.zeros_like(src)
    m.oneway_recursive(up, adjoint, freqs, src, out)
    inpl = src.clone()
    m.oneway(up, adjoint, freqs, inpl)
    assert torch.equal(out, inpl)
    assert torch.equal(src, torch.from_numpy(W).cuda())   # sources untouched
