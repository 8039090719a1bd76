"""BornFull forward and adjoint of shots with uneven apertures and bands in
one batched call (wemfwi_uneven_born_forward / _adjoint):

* every shot's data, and every shot's gradient slice before the sum over
  shots, bitwise equal to wemfwi_born_forward / wemfwi_born_adjoint called
  on that shot alone (a one-shot survey);
* parity with the FP64 oracle's BornFull (born_prop.cpp:11-235 per shot on
  its aperture, the perturbation windowed to the aperture and zero outside
  the grid), relative L2 <= 1e-5;
* the exact-mode dot product <L m, d> = <m, L^H d> with d correlated with
  L m, relative to |<L m, d>|;
* FFT lengths 128 and 256 (full rows), runtime-planned 100 and 200 in one
  call, SSF and PSPI, a shot hanging over the right edge of the grid, source
  and receiver spreads, receivers below the surface; argument checks."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NZ, NXG, DX, DZ = 24, 300, 10.0, 10.0
NGS, NGR = 2, 1

SHOTS = [
    # x0, nx, freqs, src_ix (aperture), nref, prop
    (0, 128, [3.0, 6.5, 11.0], 40, 1, "ssf"),
    (30, 256, [2.5, 4.0, 7.25, 12.0, 16.5], 120, 2, "pspi"),
    (120, 100, [5.0, 9.5], 50, 1, "ssf"),       # FFT length 100 (runtime plan)
    (230, 100, [4.0, 8.0, 13.0], 30, 2, "pspi"),  # columns 300..329 lie outside the grid
    (60, 200, [6.0, 10.0], 100, 1, "ssf"),      # FFT length 200 in a 256-lane row
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _world(seed=3):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    # 30 columns of edge padding on the right for the overhanging shot
    v = (1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((NZ, NXG + 30)))
    den = 1 + 0.2 * rng.random((NZ, NXG + 30))
    return (1 / v ** 2).astype(np.complex64), den.astype(np.complex64)


def _window(a, x0, nx):
    out = np.zeros((a.shape[0], nx), dtype=a.dtype)
    lo, hi = max(x0, 0), min(x0 + nx, a.shape[1])
    out[:, lo - x0:hi - x0] = a[:, lo:hi]
    return out


def _build(seed=4):
    import torch
    from wemfwi import Model
    from wemfwi.ops import BornShot
    slow, den = _world()
    rng = np.random.default_rng(seed)
    shots, info = [], []
    for (x0, nx, freqs, six, nref, prop) in SHOTS:
        s = np.ascontiguousarray(slow[:, x0:x0 + nx])
        d = np.ascontiguousarray(den[:, x0:x0 + nx])
        m = Model(s, DX, DZ, nref=nref, prop=prop, eps=0.04, ntaylor=1, den=d)
        nw = len(freqs)
        wav = (rng.standard_normal(nw) + 1j * rng.standard_normal(nw)).astype(np.complex64)
        rix = list(range(0, nx, 5)) + [nx // 2, nx // 2]
        riz = [0] * (len(rix) - 3) + [3, 5, 5]
        data = torch.zeros((len(rix), nw), dtype=torch.complex64, device="cuda")
        shots.append(BornShot(m, x0, freqs, wav, 0, six, riz, rix, data))
        info.append(dict(x0=x0, nx=nx, freqs=np.asarray(freqs, np.float32), wav=wav, six=six, riz=riz, rix=rix,
                         nref=nref, prop=prop, slow=s, den=d, model=m))
    return shots, info


def _pert(seed=5):
    rng = np.random.default_rng(seed)
    ds = (rng.standard_normal((NZ, NXG)) + 1j * rng.standard_normal((NZ, NXG))) * 1e-8
    dr = (rng.standard_normal((NZ, NXG)) + 1j * rng.standard_normal((NZ, NXG))) * 1e-2
    return ds.astype(np.complex64), dr.astype(np.complex64)


def _oracle_survey(it):
    nw = len(it["freqs"])
    return O.Survey(nx=it["nx"], nz=NZ, dx=DX, dz=DZ, freq=it["freqs"], index=np.arange(nw),
                    wave_f=it["wav"].astype(complex), src_iz=[0], src_ix=[it["six"]], rec_iz=it["riz"],
                    rec_ix=it["rix"], nt_f=nw, tap=0, pad=0, nref=it["nref"], eps=0.04, prop=it["prop"],
                    ntaylor=1, ngs=NGS, ngr=NGR)


def _gpu_survey(it):
    from wemfwi.ops import Survey
    nw = len(it["freqs"])
    return Survey([0], [it["six"]], it["riz"], it["rix"], it["freqs"], np.arange(nw), it["wav"], nw,
                  ngs=NGS, ngr=NGR, offset=0)


def _c(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).cuda()


def test_uneven_born_forward_bitwise_per_shot_and_oracle():
    import torch
    from wemfwi.ops import born_forward, uneven_born_forward
    shots, info = _build()
    ds, dr = _pert()
    print("uneven born forward: 5 shots, apertures 128/256/100/100/200", flush=True)
    # stale values in the data must be overwritten (add = False)
    for sh in shots:
        sh.data.fill_(7.0)
    uneven_born_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR)
    for sh, it in zip(shots, info):
        nw = len(it["freqs"])
        # the same shot alone through the one-grid operator
        one = torch.zeros((1, len(it["rix"]), nw), dtype=torch.complex64, device="cuda")
        sv = _gpu_survey(it)
        born_forward(it["model"], sv, _c(_window(ds, it["x0"], it["nx"])), _c(_window(dr, it["x0"], it["nx"])), one)
        sv.close()
        assert torch.equal(one[0], sh.data), f"shot x0={it['x0']}: not bitwise the one-shot call"
        want = O.born_forward(_oracle_survey(it), it["slow"], it["den"], _window(ds, it["x0"], it["nx"]),
                              _window(dr, it["x0"], it["nx"]))[0]
        assert rel(sh.data.cpu().numpy(), want) < TOL, f"shot x0={it['x0']}"
    # add = True accumulates
    before = [sh.data.clone() for sh in shots]
    uneven_born_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR, add=True)
    for b, sh in zip(before, shots):
        assert rel(sh.data.cpu().numpy(), 2 * b.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("exact", [False, True])
def test_uneven_born_adjoint_bitwise_per_shot_and_oracle(exact):
    import torch
    from wemfwi.ops import born_adjoint, uneven_born_adjoint
    shots, info = _build()
    rng = np.random.default_rng(8)
    for sh in shots:
        nr, nw = sh.data.shape
        sh.data.copy_(_c(rng.standard_normal((nr, nw)) + 1j * rng.standard_normal((nr, nw))))
    print(f"uneven born adjoint exact={exact}", flush=True)
    gS = torch.full((NZ, NXG), 5.0, dtype=torch.complex64, device="cuda")
    gR = torch.full((NZ, NXG), 5.0, dtype=torch.complex64, device="cuda")
    uneven_born_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=exact)
    accS = np.zeros((NZ, NXG), np.complex128)
    accR = np.zeros((NZ, NXG), np.complex128)
    wS = np.zeros((NZ, NXG), np.complex128)
    wR = np.zeros((NZ, NXG), np.complex128)
    mode = "exact" if exact else "compat"
    for sh, it in zip(shots, info):
        x0, nx = it["x0"], it["nx"]
        lo, hi = max(x0, 0), min(x0 + nx, NXG)
        a = torch.zeros((NZ, nx), dtype=torch.complex64, device="cuda")
        b = torch.zeros((NZ, nx), dtype=torch.complex64, device="cuda")
        sv = _gpu_survey(it)
        born_adjoint(it["model"], sv, a, b, sh.data[None].contiguous(), exact=exact)
        sv.close()
        # the shot's float slice, summed over shots in shot order in FP64
        accS[:, lo:hi] += a.cpu().numpy()[:, lo - x0:hi - x0].astype(np.complex128)
        accR[:, lo:hi] += b.cpu().numpy()[:, lo - x0:hi - x0].astype(np.complex128)
        oS, oR = O.born_adjoint(_oracle_survey(it), it["slow"], it["den"],
                                sh.data.cpu().numpy()[None].astype(complex), mode=mode)
        wS[:, lo:hi] += oS[:, lo - x0:hi - x0]
        wR[:, lo:hi] += oR[:, lo - x0:hi - x0]
    assert torch.equal(gS, _c(accS.astype(np.complex64))), "not bitwise the sum of one-shot calls"
    assert torch.equal(gR, _c(accR.astype(np.complex64)))
    assert rel(gS.cpu().numpy(), wS) < TOL
    assert rel(gR.cpu().numpy(), wR) < TOL
    # add = True accumulates
    g0 = gS.clone()
    uneven_born_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=exact, add=True)
    assert rel(gS.cpu().numpy(), 2 * g0.cpu().numpy()) < 1e-6


def test_uneven_born_dot_product_exact():
    import torch
    from wemfwi.ops import uneven_born_adjoint, uneven_born_forward
    shots, info = _build(seed=9)
    ds, dr = _pert(seed=10)
    print("uneven born dot test", flush=True)
    uneven_born_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR)
    lm = [sh.data.cpu().numpy().astype(np.complex128) for sh in shots]
    rng = np.random.default_rng(11)
    # d correlated with L m, so that |<Lm, d>| is of the order of |Lm| |d|
    dd = []
    for x in lm:
        scale = np.linalg.norm(x) / np.sqrt(x.size)
        dd.append(x + 0.5 * scale * (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)))
    for sh, d in zip(shots, dd):
        sh.data.copy_(_c(d))
    dd = [sh.data.cpu().numpy().astype(np.complex128) for sh in shots]
    gS = torch.zeros((NZ, NXG), dtype=torch.complex64, device="cuda")
    gR = torch.zeros((NZ, NXG), dtype=torch.complex64, device="cuda")
    uneven_born_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=True)
    lhs = sum(np.vdot(d, x) for d, x in zip(dd, lm))
    rhs = np.vdot(gS.cpu().numpy().astype(np.complex128), ds.astype(np.complex128)) + \
        np.vdot(gR.cpu().numpy().astype(np.complex128), dr.astype(np.complex128))
    # <d, L m> = <L^H d, m>
    assert abs(lhs - rhs) / abs(lhs) < TOL


def test_uneven_born_argument_checks():
    import torch
    from wemfwi import WemfwiError
    from wemfwi.ops import BornShot, uneven_born_forward
    shots, info = _build()
    ds, dr = _pert()
    it = info[0]
    bad = BornShot(it["model"], 0, it["freqs"], it["wav"], NZ, it["six"], it["riz"], it["rix"], shots[0].data)
    with pytest.raises(WemfwiError):
        uneven_born_forward([bad], _c(ds), _c(dr))
    bad = BornShot(it["model"], 0, it["freqs"], it["wav"], 0, it["six"], [NZ + 1] * len(it["riz"]), it["rix"],
                   shots[0].data)
    with pytest.raises(WemfwiError):
        uneven_born_forward([bad], _c(ds), _c(dr))
    with pytest.raises(ValueError):
        uneven_born_forward(shots, _c(ds), _c(dr[:, :10]))
    # a model of another depth count
    from wemfwi import Model
    m = Model(np.ascontiguousarray(it["slow"][:NZ - 2]), DX, DZ)
    bad = BornShot(m, 0, it["freqs"], it["wav"], 0, it["six"], it["riz"], it["rix"], shots[0].data)
    with pytest.raises(WemfwiError):
        uneven_born_forward([bad], _c(ds), _c(dr))
    torch.cuda.synchronize()
