"""extBornFull forward and adjoint of shots with uneven apertures and bands in
one batched call (wemfwi_uneven_extborn_forward / _adjoint): every frequency
of a shot has its own background and its own perturbation slice (w0 + i of
[nslice][nz][nx_model]).

* every shot's data, and (exact mode) every shot's contribution to each
  slice before the sum over shots, bitwise equal to wemfwi_extborn_forward /
  wemfwi_extborn_adjoint called on that shot alone;
* parity with the FP64 oracle's extBornFull (ext_born_prop.cpp:8-194),
  relative L2 <= 1e-5, including the compat mode's last-row rule of the
  reference's serial shot loop on shots that cover the whole grid;
* compat mode: the serial loop of one-shot compat calls, bitwise (the last
  depth row of a slice keeps the last covering shot's value);
* the exact-mode dot product <L m, d> = <m, L^H d>;
* FFT lengths 128 and 256, runtime-planned 100 and 200, SSF and PSPI, a shot
  hanging over the right edge of the grid."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5
NZ, NXG, DX, DZ = 24, 300, 10.0, 10.0
NGS, NGR = 2, 1
NSLICE = 8
FREQ = 2.5 + 1.5 * np.arange(NSLICE)

SHOTS = [
    # x0, nx, w0, nw, src_ix (aperture), nref, prop
    (0, 128, 0, 3, 40, 1, "ssf"),
    (30, 256, 2, 4, 120, 2, "pspi"),
    (120, 100, 1, 2, 50, 1, "ssf"),       # FFT length 100 (runtime plan)
    (230, 100, 5, 3, 30, 2, "pspi"),      # columns 300..329 lie outside the grid
    (60, 200, 3, 2, 100, 1, "ssf"),       # FFT length 200 in a 256-lane row
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _world(seed=3):
    """One background per slice (the frequency-extended model) and one
    density, with 30 columns of edge padding on the right."""
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, NZ)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((NZ, NXG + 30)))
    slow = np.stack([(1 / (v * (1 + 0.02 * q)) ** 2) for q in range(NSLICE)]).astype(np.complex64)
    den = (1 + 0.2 * rng.random((NZ, NXG + 30))).astype(np.complex64)
    return slow, den


def _window(a, x0, nx):
    out = np.zeros(a.shape[:-1] + (nx,), dtype=a.dtype)
    lo, hi = max(x0, 0), min(x0 + nx, a.shape[-1])
    out[..., lo - x0:hi - x0] = a[..., lo:hi]
    return out


def _build(shots_spec=SHOTS, seed=4, common=False):
    """common: every shot with the same receivers and source spectrum (an
    oracle survey of several shots)."""
    import torch
    from wemfwi import Model
    from wemfwi.ops import ExtBornShot
    slow, den = _world()
    rng = np.random.default_rng(seed)
    shots, info = [], []
    for (x0, nx, w0, nw, six, nref, prop) in shots_spec:
        s3 = np.ascontiguousarray(slow[w0:w0 + nw, :, x0:x0 + nx])
        d = np.ascontiguousarray(den[:, x0:x0 + nx])
        ms = [Model(np.ascontiguousarray(s3[i]), DX, DZ, nref=nref, prop=prop, eps=0.04, ntaylor=1, den=d)
              for i in range(nw)]
        freqs = FREQ[w0:w0 + nw].astype(np.float32)
        wav = (rng.standard_normal(nw) + 1j * rng.standard_normal(nw)).astype(np.complex64)
        if common:
            wav = np.linspace(1.0, 2.0, nw).astype(np.complex64) * (1 - 0.5j)
            rix = list(range(0, nx, 7))
            riz = [0] * len(rix)
        else:
            rix = list(range(0, nx, 5)) + [nx // 2, nx // 2]
            riz = [0] * (len(rix) - 3) + [3, 5, 5]
        data = torch.zeros((len(rix), nw), dtype=torch.complex64, device="cuda")
        shots.append(ExtBornShot(ms, x0, w0, freqs, wav, 0, six, riz, rix, data))
        info.append(dict(x0=x0, nx=nx, w0=w0, nw=nw, freqs=freqs, wav=wav, six=six, riz=riz, rix=rix,
                         nref=nref, prop=prop, slow3=s3, den=d, models=ms))
    return shots, info


def _pert(seed=5):
    rng = np.random.default_rng(seed)
    shape = (NSLICE, NZ, NXG)
    ds = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 1e-8
    dr = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) * 1e-2
    return ds.astype(np.complex64), dr.astype(np.complex64)


def _oracle_survey(it, src_ix=None):
    nw = it["nw"]
    six = [it["six"]] if src_ix is None else list(src_ix)
    return O.Survey(nx=it["nx"], nz=NZ, dx=DX, dz=DZ, freq=it["freqs"], index=np.arange(nw),
                    wave_f=it["wav"].astype(complex), src_iz=[0] * len(six), src_ix=six, rec_iz=it["riz"],
                    rec_ix=it["rix"], nt_f=nw, tap=0, pad=0, nref=it["nref"], eps=0.04, prop=it["prop"],
                    ntaylor=1, ngs=NGS, ngr=NGR)


def _gpu_survey(it):
    from wemfwi.ops import Survey
    nw = it["nw"]
    return Survey([0], [it["six"]], it["riz"], it["rix"], it["freqs"], np.arange(nw), it["wav"], nw,
                  ngs=NGS, ngr=NGR, offset=0)


def _c(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.complex64))).cuda()


def _den3(it):
    return np.repeat(it["den"][None], it["nw"], axis=0)


def test_uneven_extborn_forward_bitwise_per_shot_and_oracle():
    import torch
    from wemfwi.ops import extborn_forward, uneven_extborn_forward
    shots, info = _build()
    ds, dr = _pert()
    print("uneven extborn forward: 5 shots, apertures 128/256/100/100/200, 8 slices", flush=True)
    for sh in shots:
        sh.data.fill_(7.0)   # overwritten (add = False)
    uneven_extborn_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR)
    for sh, it in zip(shots, info):
        x0, nx, w0, nw = it["x0"], it["nx"], it["w0"], it["nw"]
        wS = _window(ds[w0:w0 + nw], x0, nx)
        wR = _window(dr[w0:w0 + nw], x0, nx)
        one = torch.zeros((1, len(it["rix"]), nw), dtype=torch.complex64, device="cuda")
        sv = _gpu_survey(it)
        extborn_forward(it["models"], sv, _c(wS), _c(wR), one)
        sv.close()
        assert torch.equal(one[0], sh.data), f"shot x0={x0}: not bitwise the one-shot call"
        want = O.extborn_forward(_oracle_survey(it), it["slow3"], _den3(it), wS, wR)[0]
        assert rel(sh.data.cpu().numpy(), want) < TOL, f"shot x0={x0}"
    before = [sh.data.clone() for sh in shots]
    uneven_extborn_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR, add=True)
    for b, sh in zip(before, shots):
        assert rel(sh.data.cpu().numpy(), 2 * b.cpu().numpy()) < 1e-6


def _random_data(shots, seed):
    rng = np.random.default_rng(seed)
    for sh in shots:
        nr, nw = sh.data.shape
        sh.data.copy_(_c(rng.standard_normal((nr, nw)) + 1j * rng.standard_normal((nr, nw))))


def test_uneven_extborn_adjoint_exact_bitwise_per_shot_and_oracle():
    import torch
    from wemfwi.ops import extborn_adjoint, uneven_extborn_adjoint
    shots, info = _build()
    _random_data(shots, 8)
    print("uneven extborn adjoint exact", flush=True)
    gS = torch.full((NSLICE, NZ, NXG), 5.0, dtype=torch.complex64, device="cuda")
    gR = torch.full((NSLICE, NZ, NXG), 5.0, dtype=torch.complex64, device="cuda")
    uneven_extborn_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=True)
    accS = np.zeros((NSLICE, NZ, NXG), np.complex128)
    accR = np.zeros((NSLICE, NZ, NXG), np.complex128)
    wS = np.zeros((NSLICE, NZ, NXG), np.complex128)
    wR = np.zeros((NSLICE, NZ, NXG), np.complex128)
    for sh, it in zip(shots, info):
        x0, nx, w0, nw = it["x0"], it["nx"], it["w0"], it["nw"]
        lo, hi = max(x0, 0), min(x0 + nx, NXG)
        a = torch.zeros((nw, NZ, nx), dtype=torch.complex64, device="cuda")
        b = torch.zeros((nw, NZ, nx), dtype=torch.complex64, device="cuda")
        sv = _gpu_survey(it)
        extborn_adjoint(it["models"], sv, a, b, sh.data[None].contiguous(), exact=True)
        sv.close()
        accS[w0:w0 + nw, :, lo:hi] += a.cpu().numpy()[:, :, lo - x0:hi - x0].astype(np.complex128)
        accR[w0:w0 + nw, :, lo:hi] += b.cpu().numpy()[:, :, lo - x0:hi - x0].astype(np.complex128)
        oS, oR = O.extborn_adjoint(_oracle_survey(it), it["slow3"], _den3(it),
                                   sh.data.cpu().numpy()[None].astype(complex), mode="exact")
        wS[w0:w0 + nw, :, lo:hi] += oS[:, :, lo - x0:hi - x0]
        wR[w0:w0 + nw, :, lo:hi] += oR[:, :, lo - x0:hi - x0]
    assert torch.equal(gS, _c(accS.astype(np.complex64))), "not bitwise the sum of one-shot calls"
    assert torch.equal(gR, _c(accR.astype(np.complex64)))
    assert rel(gS.cpu().numpy(), wS) < TOL
    assert rel(gR.cpu().numpy(), wR) < TOL
    g0 = gS.clone()
    uneven_extborn_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=True, add=True)
    assert rel(gS.cpu().numpy(), 2 * g0.cpu().numpy()) < 1e-6


def test_uneven_extborn_adjoint_compat_last_row_rule():
    """compat (the reference's serial shot loop): every depth row but the
    last is the FP64 sum over shots of the one-shot compat calls, bitwise; on
    the last row each column of a slice keeps the value of the last shot that
    covers it (LinDown::adjoint's cleared row, LinOneWay.cpp:64-65; dReflect
    adds nothing on that row).  On shots covering the whole grid with one
    band, the FP64 oracle's serial-loop extBornFull adjoint."""
    import torch
    from wemfwi.ops import extborn_adjoint, uneven_extborn_adjoint
    shots, info = _build()
    _random_data(shots, 12)
    print("uneven extborn adjoint compat", flush=True)
    gS = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    gR = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    uneven_extborn_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=False)
    accS = np.zeros((NSLICE, NZ, NXG), np.complex128)
    accR = np.zeros((NSLICE, NZ, NXG), np.complex128)
    overwritten = np.zeros((NSLICE, NXG), bool)
    for sh, it in zip(shots, info):
        x0, nx, w0, nw = it["x0"], it["nx"], it["w0"], it["nw"]
        lo, hi = max(x0, 0), min(x0 + nx, NXG)
        a = torch.zeros((nw, NZ, nx), dtype=torch.complex64, device="cuda")
        b = torch.zeros((nw, NZ, nx), dtype=torch.complex64, device="cuda")
        sv = _gpu_survey(it)
        extborn_adjoint(it["models"], sv, a, b, sh.data[None].contiguous(), exact=False)
        sv.close()
        a = a.cpu().numpy()[:, :, lo - x0:hi - x0].astype(np.complex128)
        b = b.cpu().numpy()[:, :, lo - x0:hi - x0].astype(np.complex128)
        accS[w0:w0 + nw, :-1, lo:hi] += a[:, :-1]
        overwritten[w0:w0 + nw, lo:hi] |= accS[w0:w0 + nw, -1, lo:hi] != 0
        accS[w0:w0 + nw, -1, lo:hi] = a[:, -1]
        accR[w0:w0 + nw, :, lo:hi] += b
    assert overwritten.any(), "the survey must have shots overlapping in aperture and band"
    assert torch.equal(gS, _c(accS.astype(np.complex64))), "compat: not the serial loop of one-shot calls"
    assert torch.equal(gR, _c(accR.astype(np.complex64)))
    # whole-grid shots, one band: the oracle's serial loop
    spec = [(0, NXG, 0, 3, six, 1, "ssf") for six in (40, 150, 260)]
    shots, info = _build(spec, seed=6, common=True)
    _random_data(shots, 13)
    gS = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    gR = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    uneven_extborn_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=False)
    it = info[0]
    data = np.stack([sh.data.cpu().numpy() for sh in shots]).astype(complex)
    wS, wR = O.extborn_adjoint(_oracle_survey(it, src_ix=[x["six"] for x in info]), it["slow3"], _den3(it), data,
                               mode="compat")
    got = gS.cpu().numpy()[:3]
    assert rel(got, wS) < TOL
    # the last row alone is one shot's LinUp term (no sum over shots to
    # average the FP32 rounding): 3e-5 here; summing the shots instead of
    # keeping the last would be off by O(1)
    assert rel(got[:, -1], wS[:, -1]) < 1e-4
    assert rel(gR.cpu().numpy()[:3], wR) < TOL


def test_uneven_extborn_dot_product_exact():
    import torch
    from wemfwi.ops import uneven_extborn_adjoint, uneven_extborn_forward
    shots, info = _build(seed=9)
    ds, dr = _pert(seed=10)
    print("uneven extborn dot test", flush=True)
    uneven_extborn_forward(shots, _c(ds), _c(dr), ngs=NGS, ngr=NGR)
    lm = [sh.data.cpu().numpy().astype(np.complex128) for sh in shots]
    rng = np.random.default_rng(11)
    dd = []
    for x in lm:
        scale = np.linalg.norm(x) / np.sqrt(x.size)
        dd.append(x + 0.5 * scale * (rng.standard_normal(x.shape) + 1j * rng.standard_normal(x.shape)))
    for sh, d in zip(shots, dd):
        sh.data.copy_(_c(d))
    dd = [sh.data.cpu().numpy().astype(np.complex128) for sh in shots]
    gS = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    gR = torch.zeros((NSLICE, NZ, NXG), dtype=torch.complex64, device="cuda")
    uneven_extborn_adjoint(shots, gS, gR, ngs=NGS, ngr=NGR, exact=True)
    lhs = sum(np.vdot(d, x) for d, x in zip(dd, lm))
    rhs = np.vdot(gS.cpu().numpy().astype(np.complex128), ds.astype(np.complex128)) + \
        np.vdot(gR.cpu().numpy().astype(np.complex128), dr.astype(np.complex128))
    assert abs(lhs - rhs) / abs(lhs) < TOL


def test_uneven_extborn_argument_checks():
    import torch
    from wemfwi import WemfwiError
    from wemfwi.ops import ExtBornShot, uneven_extborn_forward
    shots, info = _build()
    ds, dr = _pert()
    it = info[3]
    # band beyond the slices
    bad = ExtBornShot(it["models"], it["x0"], NSLICE - 1, it["freqs"], it["wav"], 0, it["six"], it["riz"],
                      it["rix"], shots[3].data)
    with pytest.raises(WemfwiError):
        uneven_extborn_forward([bad], _c(ds), _c(dr))
    # models of different widths within a shot
    mixed = [info[0]["models"][0]] + info[3]["models"][1:]
    bad = ExtBornShot(mixed, 0, 0, it["freqs"], it["wav"], 0, 10, it["riz"], it["rix"], shots[3].data)
    with pytest.raises(WemfwiError):
        uneven_extborn_forward([bad], _c(ds), _c(dr))
    with pytest.raises(ValueError):
        ExtBornShot(it["models"][:2], 0, 0, it["freqs"], it["wav"], 0, 10, it["riz"], it["rix"], shots[3].data)
    with pytest.raises(ValueError):
        uneven_extborn_forward(shots, _c(ds[0]), _c(dr[0]))
    torch.cuda.synchronize()
