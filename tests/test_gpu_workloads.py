"""The two benchmark workloads at their own shapes, on sampled rows, against
the FP64 oracle.

PSPI downward continuation (nref = 4, nx = 1024, nz = 500, 256 frequencies
2-30 Hz, 16 shots, the six-layer model of seed 2): rows at 2, ~16 and 30 Hz
of shots 0 and 15 through the full 499-step Down recursion, forward from the
surface sources and adjoint from random data, relative L2 <= 1e-5 per row.
The recursion crosses the depths where RefSampler's stale buffer activates
several references (RefSampler.cpp:63-81), so the multi-reference step runs
inside the sampled sweeps.

FWI gradient (nx = 512, nz = 300, 128 frequencies 2-15 Hz, 32 shots, SSF;
true model = gradient + two anomalies, start = true smoothed): the misfit J
and the exact Born-adjoint gradient of 4 (shot, frequency) rows, store-all
and with wavefield rows kept every 25 depths, against the oracle's
wem_forward / born_adjoint(mode="exact") of the same rows, relative L2 <=
1e-5 (the gradient) and relative 1e-5 (J)."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_pspi_headline_sampled_rows_full_depth():
    import torch
    from bench import CONFIGS, ricker_spectrum, velocity
    from wemfwi import Model
    cfg = CONFIGS["pspi1024"]
    nx, nz, nw, ns, nref = cfg["nx"], cfg["nz"], cfg["nw"], cfg["ns"], cfg["nref"]
    freqs = (cfg["fmin"] + np.arange(nw) * (cfg["fmax"] - cfg["fmin"]) / (nw - 1)).astype(np.float32)
    src_ix = [32 + 64 * s for s in range(ns)]
    slow = (1.0 / velocity(cfg) ** 2).astype(np.complex64)
    ref = O.refsampler(slow, nref)
    multi = sum(len(set(np.unique(ref.refid[iz]).tolist()) - {-1}) > 1 for iz in range(nz - 1))
    print(f"pspi nref={nref} {nx} x {nz}: {multi} depth steps with several active references", flush=True)
    assert multi > 0
    rows = [(s, i) for s in (0, ns - 1) for i in (0, nw // 2, nw - 1)]
    f_rows = np.array([freqs[i] for _, i in rows], dtype=np.float32)
    spec = ricker_spectrum(freqs)
    src = np.zeros((len(rows), nz, nx), np.complex64)
    for k, (s, i) in enumerate(rows):
        src[k, 0, src_ix[s]] = spec[i]
    m = Model(slow, 10.0, 10.0, nref=nref, prop=cfg["prop"], eps=0.04)
    S = torch.from_numpy(src).cuda()
    W = torch.empty_like(S)
    print("Down forward, 6 rows x 499 steps", flush=True)
    m.oneway_recursive(False, False, f_rows, S, W)
    got = W.cpu().numpy()
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, nref=nref, prop=cfg["prop"], eps=0.04)
    w = 2 * np.pi * f_rows.astype(np.float64)
    want = O.down_fwd(src.astype(complex), w, slow, ref, p)
    errs = [rel(got[k], want[k]) for k in range(len(rows))]
    print("forward rel L2 per row:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= TOL
    rng = np.random.default_rng(2)
    d = (rng.standard_normal((len(rows), nz, nx)) + 1j * rng.standard_normal((len(rows), nz, nx))).astype(np.complex64)
    D = torch.from_numpy(d).cuda()
    print("Down adjoint, 6 rows x 499 steps", flush=True)
    m.oneway(False, True, f_rows, D)
    got = D.cpu().numpy()
    want = O.down_adj(d.astype(complex), w, slow, ref, p)
    errs = [rel(got[k], want[k]) for k in range(len(rows))]
    print("adjoint rel L2 per row:", " ".join(f"{e:.2e}" for e in errs), flush=True)
    assert max(errs) <= TOL
    m.close()


def test_fwi_gradient_sampled_rows_vs_fp64_oracle():
    import torch
    from bench import CKPT, FWI, fwi_models, ricker_spectrum
    from wemfwi import Model
    from wemfwi.ops import Survey, born_adjoint, wem_forward
    nx, nz, nw, ns = FWI["nx"], FWI["nz"], FWI["nw"], FWI["ns"]
    freqs = (FWI["fmin"] + np.arange(nw) * (FWI["fmax"] - FWI["fmin"]) / (nw - 1)).astype(np.float32)
    wav = ricker_spectrum(freqs)
    src_ix = [int(round(8 + s * (nx - 16) / (ns - 1))) for s in range(ns)]
    shots, fidx = [0, ns - 1], [0, nw - 1]
    s_true, s_start = fwi_models()
    den = np.ones((nz, nx), np.complex64)
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freqs[fidx], index=np.arange(len(fidx)),
                  wave_f=wav[fidx], src_iz=[0] * len(shots), src_ix=[src_ix[s] for s in shots],
                  rec_iz=[0] * nx, rec_ix=list(range(nx)), nt_f=len(fidx), tap=0, pad=0, nref=1, eps=0.04,
                  prop="ssf", ntaylor=1, ngs=0, ngr=0, ng=0)
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f, sv.nt_f)
    m_true = Model(s_true, 10.0, 10.0, den=den)
    m0 = Model(s_start, 10.0, 10.0, den=den)
    print("FWI gradient 512 x 300: 2 shots x 2 frequencies", flush=True)
    obs = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m_true, s, obs)
    pred = torch.zeros_like(obs)
    wem_forward(m0, s, pred)
    res = pred - obs
    J = 0.5 * float((res.to(torch.complex128).abs() ** 2).sum())
    # the oracle end to end in FP64: observed and predicted data, residual,
    # misfit and the exact Born adjoint of the residual
    obs_o = O.wem_forward(sv, s_true, den)
    pred_o = O.wem_forward(sv, s_start, den)
    res_o = pred_o - obs_o
    J_o = 0.5 * float(np.sum(np.abs(res_o) ** 2))
    print(f"data rel L2 {rel(obs.cpu().numpy(), obs_o):.2e} / {rel(pred.cpu().numpy(), pred_o):.2e}, "
          f"|res|/|pred| = {np.linalg.norm(res_o) / np.linalg.norm(pred_o):.3f}, "
          f"J rel {abs(J - J_o) / J_o:.2e}", flush=True)
    assert rel(obs.cpu().numpy(), obs_o) <= TOL and rel(pred.cpu().numpy(), pred_o) <= TOL
    assert abs(J - J_o) <= TOL * J_o
    gS_o, gR_o = O.born_adjoint(sv, s_start, den, res_o, mode="exact")
    # the adjoint alone: the oracle's Born adjoint of the GPU's residual
    gS_r, gR_r = O.born_adjoint(sv, s_start, den, res.cpu().numpy().astype(complex), mode="exact")
    grads = []
    for k in (0, CKPT):
        s.set_checkpoint(k)
        gS = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
        gR = torch.zeros_like(gS)
        born_adjoint(m0, s, gS, gR, res, exact=True)
        a, b = gS.cpu().numpy(), gR.cpu().numpy()
        e_end, e_adj = rel(a, gS_o), rel(a, gS_r)
        print(f"checkpoint {k}: gradient rel L2 vs oracle end to end {e_end:.2e} (density {rel(b, gR_o):.2e}), "
              f"adjoint of the same residual {e_adj:.2e}", flush=True)
        assert e_end <= TOL and rel(b, gR_o) <= TOL and e_adj <= TOL
        grads.append((gS, gR))
    assert torch.equal(grads[0][0], grads[1][0]) and torch.equal(grads[0][1], grads[1][1])
    for mm in (m_true, m0):
        mm.close()
