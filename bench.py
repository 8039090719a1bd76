#!/usr/bin/env python
"""Throughput of the fused one-way depth step on one MI355X.

Workload (default): 2-D PSPI downward continuation with 4 reference
slownesses per depth, nx = 1024, nz = 500, dx = dz = 10 m, 256 frequencies
2-30 Hz, 16 shots at ix = 32 + 64*s, complex64; velocity = 6 seeded random flat
layers 1500-4500 m/s plus a smooth +-10% perturbation (seed 2).  All
frequency x shot rows go to the GPU in one launch per sweep.

One step = the phase tables of the 256 frequencies (Phshift::setFreq for all
of them) + Down forward over all 4096 rows (sources -> wavefield) + Down
adjoint over all rows (in place), i.e. 2 sweeps of 1024 x 499 x 4096 wavefield
updates.  The timed steps replay that sequence as one captured hipGraph of
batch-plan launches; eager steps before it warm up (the sweep autotuner
decides there) and give the per-launch kernel times of the roofline.  The
metric is wavefield updates per second (x.z.w, summed over the shots), the
reference's unit of work.

Also reported in the same line: forward and adjoint sweep rates on their own;
the same grid, rows and model with SSF and one reference (two FFTs and the
split-step correction at every depth step) as a sweep line against the HBM
roofline; and the FWI gradient (WEM forward + residual + exact Born adjoint,
512 x 300, 128 frequencies, 32 shots) with all rows batched against the same
library called once per frequency with all shots of it (a timed sample of
frequencies, scaled), with the relative difference of the two gradients on
the sampled frequencies; the batched gradient again with wavefield rows kept
every 25 depth steps (recomputed segment by segment), which must equal the
store-all gradient bitwise.  The CPU leg also checks 2 x 2 rows of that
gradient and its misfit against the FP64 oracle.

Also the uneven batch: 64 shots with their own apertures (256-2048) and
bands (32-192 frequencies), nz = 400, Down forward of all rows in one batched
call against one call per shot and one call per row (sampled); and the
subsurface-offset migration (nh = 11) of the same survey in one call against
one call per (shot, frequency) row (sampled).

Also Down sweeps of 1024 rows at the reference's own FFT lengths nk = 500
and 600 (compile-time pass chains, and the runtime chain for comparison)
beside the powers of two 512 and 1024.

The gradient line also carries a central-difference check on 20 random
directions (--fd-directions 0 skips it).

Prints one JSON line.  --cpu-sample 0 skips the CPU baseline, --fwi-sample 0
the gradient comparison, --uneven-sample 0 the uneven batch,
--any-length-rows 0 the nk = 500 / 600 sweeps.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "wem-fwi_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0     # MI355X HBM3E
BYTES_PER_UPDATE = 16     # read W[iz] c64 + write W[iz+1] c64 (slowness row amortised)
VALU_PEAK_TFLOPS = 157.3  # MI355X FP32 (published; packed FP32 counted)
VALU_SCALAR_TFLOPS = 78.6  # the same without packed FP32 instructions, which this build does not emit


def sweep_flops(slow, nref: int, prop: str, nrows: int) -> float:
    """Algorithmic FP32 flops of one Down sweep over nrows rows: per depth
    step one forward FFT, and per active reference of the depth a phase
    multiply and an inverse FFT; SSF adds the split-step/taper multiply; the
    source add closes the step.  FFT = 5 N log2 N, complex multiply = 6,
    complex add = 2 flops per sample."""
    import wemfwi
    _, refid = wemfwi.refsampler(slow, nref)
    nz, nx = refid.shape
    n = nx
    fft = 5.0 * n * math.log2(n)
    total = 0.0
    for iz in range(nz - 1):
        slots = len(set(np.unique(refid[iz]).tolist()) - {-1})
        total += fft + slots * (6.0 * n + fft) + (6.0 * n if prop == "ssf" else 0.0) + 2.0 * n
    return total * nrows

CONFIGS = {
    # PSPI downward continuation (the headline configuration)
    "pspi1024": dict(nx=1024, nz=500, nw=256, fmin=2.0, fmax=30.0, ns=16, nref=4, prop="pspi",
                     seed=2, model="layers"),
    # the PSPI workload's grid, rows and model with one reference per depth
    # and the split-step correction (SSF): two FFTs per depth step, every step
    "ssf1024": dict(nx=1024, nz=500, nw=256, fmin=2.0, fmax=30.0, ns=16, nref=1, prop="ssf",
                    seed=2, model="layers"),
    # split-step Fourier, single shot (the CPU-runnable configuration)
    "ssf256": dict(nx=256, nz=200, nw=64, fmin=2.0, fmax=20.0, ns=1, nref=1, prop="ssf",
                   seed=1, model="gradient"),
}


def velocity(cfg) -> np.ndarray:
    from scipy.ndimage import gaussian_filter
    rng = np.random.default_rng(cfg["seed"])
    nz, nx = cfg["nz"], cfg["nx"]
    if cfg["model"] == "layers":
        tops = np.sort(rng.integers(1, nz - 1, size=5))
        vel = rng.uniform(1500.0, 4500.0, size=6)
        v = np.empty((nz, nx))
        bounds = [0, *tops.tolist(), nz]
        for i in range(6):
            v[bounds[i]:bounds[i + 1]] = vel[i]
        pert = gaussian_filter(rng.standard_normal((nz, nx)), 10.0)
        v *= 1.0 + 0.10 * pert / np.max(np.abs(pert))
    else:
        z = np.arange(nz)[:, None] * 10.0
        v = 1500.0 + 2000.0 * z / ((nz - 1) * 10.0) + np.zeros((nz, nx))
        pert = gaussian_filter(rng.standard_normal((nz, nx)), 5.0)
        v = v + 0.05 * v * pert / np.max(np.abs(pert))
    return v


def ricker_spectrum(freqs, f0=10.0, t0=0.1):
    f = np.asarray(freqs, dtype=np.float64)
    amp = 2.0 / math.sqrt(math.pi) * f ** 2 / f0 ** 3 * np.exp(-(f / f0) ** 2)
    return (amp * np.exp(-2j * math.pi * f * t0)).astype(np.complex64)


def cpu_baseline(cfg, slow, freqs, src_ix, nrows_sample):
    """FP32 C restatement (oracle/cpu_wem.c) on a bounded sample of rows."""
    from oracle import cpu_wem
    from oracle import wem_oracle as O
    nz, nx = cfg["nz"], cfg["nx"]
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, nref=cfg["nref"], prop=cfg["prop"], eps=0.04)
    ref = O.refsampler(slow, cfg["nref"])
    nw = len(freqs)
    stride = max(1, (nw * len(src_ix)) // nrows_sample)
    rows = [(s, w) for s in range(len(src_ix)) for w in range(nw)][::stride][:nrows_sample]
    W = np.zeros((len(rows), nz, nx), dtype=np.complex64)
    spec = ricker_spectrum(freqs)
    fr = np.array([freqs[w] for _, w in rows], dtype=np.float32)
    for i, (s, w) in enumerate(rows):
        W[i, 0, src_ix[s]] = spec[w]
    cores = os.cpu_count() or 1
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    t0 = time.perf_counter()
    cpu_wem.oneway(W, fr, slow, ref, p, up=False, adjoint=False, nthreads=cores)
    cpu_wem.oneway(W, fr, slow, ref, p, up=False, adjoint=True, nthreads=cores)
    dt = time.perf_counter() - t0
    upd = 2.0 * nx * (nz - 1) * len(rows)
    # one thread on a smaller sample (the per-core rate)
    n1 = max(1, len(rows) // 64)
    W1 = W[:n1].copy()
    t0 = time.perf_counter()
    cpu_wem.oneway(W1, fr[:n1], slow, ref, p, up=False, adjoint=False, nthreads=1)
    cpu_wem.oneway(W1, fr[:n1], slow, ref, p, up=False, adjoint=True, nthreads=1)
    dt1 = time.perf_counter() - t0
    return {"value": upd / dt / 1e9, "unit": "Gupdates/s", "cores": cores, "kind": "port",
            "sample": f"{len(rows)} of {nw * len(src_ix)} (shot, frequency) rows, Down forward + "
                      f"adjoint, FP32 C restatement (oracle/cpu_wem.c, OpenMP, radix-4 Stockham FFT), "
                      f"{dt:.2f} s",
            "single_thread_value": 2.0 * nx * (nz - 1) * n1 / dt1 / 1e9,
            "single_thread_sample": f"{n1} rows, {dt1:.2f} s",
            "host": host_cpu()}


def fwi_oracle_check():
    """The CPU leg's check of the FWI gradient workload against the FP64
    oracle (the oracle is the checker here, never the thing timed): misfit
    and exact Born-adjoint gradient of 2 shots x 2 frequencies (shots 0 and
    31, 2 and 15 Hz) of the 512 x 300 workload, batched on the GPU, store-all
    and checkpointed every 25 depth steps, against wem_forward /
    born_adjoint(mode="exact") of the oracle on the same rows."""
    import torch
    import wemfwi
    from oracle import wem_oracle as O
    from wemfwi import ops
    c = FWI
    nx, nz, nw, ns = c["nx"], c["nz"], c["nw"], c["ns"]
    freqs = (c["fmin"] + np.arange(nw) * (c["fmax"] - c["fmin"]) / (nw - 1)).astype(np.float32)
    wav = ricker_spectrum(freqs)
    src_ix = [int(round(8 + s * (nx - 16) / (ns - 1))) for s in range(ns)]
    shots, fidx = [0, ns - 1], [0, nw - 1]
    s_true, s_start = fwi_models()
    den = np.ones((nz, nx), np.complex64)
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freqs[fidx], index=np.arange(len(fidx)),
                  wave_f=wav[fidx], src_iz=[0] * len(shots), src_ix=[src_ix[s] for s in shots],
                  rec_iz=[0] * nx, rec_ix=list(range(nx)), nt_f=len(fidx), tap=0, pad=0, nref=1, eps=0.04,
                  prop="ssf", ntaylor=1, ngs=0, ngr=0, ng=0)
    g = ops.Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f, sv.nt_f)
    m_true = wemfwi.Model(s_true, 10.0, 10.0, den=den)
    m0 = wemfwi.Model(s_start, 10.0, 10.0, den=den)
    obs = torch.zeros(g.data_shape, dtype=torch.complex64, device="cuda")
    ops.wem_forward(m_true, g, obs)
    res = torch.zeros_like(obs)
    ops.wem_forward(m0, g, res)
    res.sub_(obs)
    J = 0.5 * float((res.to(torch.complex128).abs() ** 2).sum())
    t0 = time.perf_counter()
    res_o = O.wem_forward(sv, s_start, den) - O.wem_forward(sv, s_true, den)
    J_o = 0.5 * float(np.sum(np.abs(res_o) ** 2))
    gS_o, _ = O.born_adjoint(sv, s_start, den, res_o, mode="exact")
    t_oracle = time.perf_counter() - t0
    out = {"rows": "shots 0, 31 x 2, 15 Hz of the 512 x 300 workload", "misfit_rel_err": abs(J - J_o) / J_o,
           "oracle_s": round(t_oracle, 1)}
    for k in (0, CKPT):
        g.set_checkpoint(k)
        gS = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
        gR = torch.zeros_like(gS)
        ops.born_adjoint(m0, g, gS, gR, res, exact=True)
        a = gS.cpu().numpy().astype(np.complex128)
        out["gradient_rel_err" + ("" if k == 0 else f"_checkpoint_{k}")] = float(
            np.linalg.norm(a - gS_o) / np.linalg.norm(gS_o))
    g.close()
    m_true.close()
    m0.close()
    return out


def host_cpu() -> str:
    """CPU model and topology of the host (lscpu when present)."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    topo = ""
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keep = ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)")
        topo = ", ".join(l.split(":", 1)[0].strip() + " " + l.split(":", 1)[1].strip()
                         for l in out.splitlines() if l.split(":", 1)[0].strip() in keep)
    except Exception:
        pass
    return f"{model}; {topo}" if topo else model


# FWI gradient configuration: 512 x 300 grid, 128 frequencies 2-15 Hz, 32 shots,
# SSF; true model = linear gradient 1500 -> 3500 m/s with a seeded +-5%
# smooth field regridded to 512 x 300 plus two Gaussian anomalies of +-300 m/s
# (radius 200 m); start model = the true model smoothed over 10 cells.
FWI = dict(nx=512, nz=300, nw=128, fmin=2.0, fmax=15.0, ns=32, seed=4)
CKPT = 25  # depth steps between kept wavefield rows in the checkpointed gradient


def fwi_models():
    from scipy.ndimage import gaussian_filter
    c = FWI
    nz, nx = c["nz"], c["nx"]
    rng = np.random.default_rng(c["seed"])
    z = np.arange(nz)[:, None] * 10.0
    v = 1500.0 + 2000.0 * z / ((nz - 1) * 10.0) + np.zeros((nz, nx))
    pert = gaussian_filter(rng.standard_normal((nz, nx)), 5.0)
    v = v + 0.05 * v * pert / np.max(np.abs(pert))
    X, Z = np.meshgrid(np.arange(nx) * 10.0, np.arange(nz) * 10.0)
    for (x0, z0, a) in ((0.3 * nx * 10, 0.5 * nz * 10, 300.0), (0.7 * nx * 10, 0.6 * nz * 10, -300.0)):
        v = v + a * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (200.0 ** 2))
    v0 = gaussian_filter(v, 10.0)
    return (1 / v ** 2).astype(np.complex64), (1 / v0 ** 2).astype(np.complex64)


def fd_check(sv, obs, s0, den, gS, ndir, seed=44, rel_step=1e-3):
    """Central differences of J along ndir smooth random slowness directions
    (seed 44) against <grad J, delta>.  The gradient is the reference's
    Born^H residual; the reference's Born is not the derivative of its WEM
    (dReflect's coefficients carry an extra (a+b)^4 factor, dReflect.h:18-24,
    and Scatter ignores the reference slownesses' dependence on S), so the two
    only correlate; the line reports how well."""
    import torch
    import wemfwi
    from wemfwi import ops
    from scipy.ndimage import gaussian_filter
    nz, nx = s0.shape
    g = gS.cpu().numpy().astype(np.complex128)

    def J(s):
        m = wemfwi.Model(s.astype(np.complex64), 10.0, 10.0, den=den)
        d = torch.zeros_like(obs)
        ops.wem_forward(m, sv, d)
        m.close()
        r = (d - obs).to(torch.complex128)
        return 0.5 * float((r.real ** 2 + r.imag ** 2).sum())

    rng = np.random.default_rng(seed)
    fd, an = [], []
    for _ in range(ndir):
        dl = gaussian_filter(rng.standard_normal((nz, nx)), 3.0)
        dl = dl / np.abs(dl).max() * float(np.abs(s0).mean()) * rel_step
        fd.append((J(s0 + dl) - J(s0 - dl)) / 2)
        an.append(float(np.real(np.vdot(g, dl))))
    fd, an = np.array(fd), np.array(an)
    return {"directions": ndir, "seed": seed, "step": f"{rel_step:g} x mean slowness, smoothed 3 cells",
            "cosine_fd_vs_gradient": round(float(np.dot(fd, an) / np.linalg.norm(fd) / np.linalg.norm(an)), 4),
            "median_rel_err": round(float(np.median(np.abs(fd - an) / np.abs(fd))), 4),
            "note": "Born^H r is not the derivative of WEM in the reference: dReflect as written is the "
                    "derivative of Reflect times (a+b)^4, and Scatter does not follow the reference "
                    "slownesses' dependence on S"}


def fwi_gradient_bench(sample_freqs: int, fd_dirs: int = 20):
    """Wall time of one FWI gradient (WEM forward of the start model, residual,
    exact Born adjoint) with all (shot, frequency) rows batched, eagerly and
    replayed as one captured hipGraph, against the same library called once
    per frequency (all shots of it in one call) on a sample of frequencies."""
    import torch
    import wemfwi
    from wemfwi import ops
    c = FWI
    nx, nz, nw, ns = c["nx"], c["nz"], c["nw"], c["ns"]
    freqs = (c["fmin"] + np.arange(nw) * (c["fmax"] - c["fmin"]) / (nw - 1)).astype(np.float32)
    wav = ricker_spectrum(freqs)
    src_ix = [int(round(8 + s * (nx - 16) / (ns - 1))) for s in range(ns)]
    rec_ix = list(range(nx))
    s_true, s_start = fwi_models()
    den = np.ones((nz, nx), np.complex64)
    bins = np.arange(nw)

    def survey(shots, fidx):
        return ops.Survey([0] * len(shots), [src_ix[s] for s in shots], [0] * nx, rec_ix, freqs[fidx],
                          np.arange(len(fidx)), wav[fidx], len(fidx))

    m_true = wemfwi.Model(s_true, 10.0, 10.0, den=den)
    m0 = wemfwi.Model(s_start, 10.0, 10.0, den=den)
    sv = survey(list(range(ns)), bins)
    dev = torch.device("cuda:0")
    obs = torch.zeros(sv.data_shape, dtype=torch.complex64, device=dev)
    ops.wem_forward(m_true, sv, obs)
    pred = torch.zeros_like(obs)
    gS = torch.zeros((nz, nx), dtype=torch.complex64, device=dev)
    gR = torch.zeros_like(gS)

    def batched():
        ops.wem_forward(m0, sv, pred)
        pred.sub_(obs)
        ops.born_adjoint(m0, sv, gS, gR, pred, exact=True)

    batched()  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batched()
    torch.cuda.synchronize()
    t_batched = time.perf_counter() - t0
    # the same gradient as one replayed hipGraph (wemfwi_gradient_create):
    # WEM, residual, FP64 misfit and Born adjoint captured once, replayed
    g_all = (gS.clone(), gR.clone())
    plan = ops.GradientPlan(m0, sv, obs)
    plan.run()  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        plan.run()
    torch.cuda.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    graph_equal = bool(torch.equal(plan.gS, g_all[0]) and torch.equal(plan.gR, g_all[1]))
    plan.close()
    del plan
    # the same gradient with wavefield rows kept every CKPT depth steps
    sv.set_checkpoint(CKPT)
    batched()  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batched()
    torch.cuda.synchronize()
    t_ckpt = time.perf_counter() - t0
    ckpt_equal = bool(torch.equal(gS, g_all[0]) and torch.equal(gR, g_all[1]))
    sv.set_checkpoint(0)
    # the plain per-frequency loop: one gradient call per frequency with all
    # shots of that frequency (one survey per frequency, built beforehand),
    # timed on a sample of frequencies and scaled
    total = ns * nw
    stride = max(1, nw // max(1, sample_freqs))
    fsample = list(range(nw))[::stride][:sample_freqs]
    f_sv = [survey(list(range(ns)), [i]) for i in fsample]
    f_obs = [obs[:, :, i:i + 1].contiguous() for i in fsample]
    d1 = torch.zeros((ns, nx, 1), dtype=torch.complex64, device=dev)
    pS = torch.zeros_like(gS)
    pR = torch.zeros_like(gS)

    def plain():
        pS.zero_()
        pR.zero_()
        for k in range(len(fsample)):
            ops.wem_forward(m0, f_sv[k], d1)
            d1.sub_(f_obs[k])
            ops.born_adjoint(m0, f_sv[k], pS, pR, d1, add=True, exact=True)

    plain()  # warmup
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain()
    torch.cuda.synchronize()
    t_plain_sample = time.perf_counter() - t0
    # the batched gradient restricted to the same frequencies, for the check
    bS = torch.zeros_like(gS)
    bR = torch.zeros_like(gS)
    svs = survey(list(range(ns)), fsample)
    d = torch.zeros(svs.data_shape, dtype=torch.complex64, device=dev)
    ops.wem_forward(m0, svs, d)
    d.sub_(obs[:, :, fsample].contiguous())
    ops.born_adjoint(m0, svs, bS, bR, d, exact=True)
    torch.cuda.synchronize()
    rel = float(torch.linalg.vector_norm((pS - bS).to(torch.complex128)) /
                torch.linalg.vector_norm(bS.to(torch.complex128)))
    t_plain = t_plain_sample * nw / len(fsample)
    # sweeps per (shot, frequency) row: executed here (WEM: Down, Up; Born
    # adjoint: Down, Up^H, Down^H, LinDown^H, Up, LinUp^H) and in the
    # reference's orchestration (wem_prop.cpp:52-74, born_prop.cpp:121-232)
    sweeps_exec, sweeps_ref = 8, 12
    upd_row = nx * (nz - 1) * total
    fd = fd_check(sv, obs, s_start, den, gS, fd_dirs) if fd_dirs > 0 else None
    return {"config": f"ssf nx={nx} nz={nz} nw={nw} {c['fmin']:g}-{c['fmax']:g} Hz ns={ns} nr={nx}",
            "batched_ms": round(t_batched * 1e3, 2),
            "graph_ms": round(t_graph * 1e3, 2),
            "graph_bitwise_equal": graph_equal,
            "sweeps_executed": sweeps_exec,
            "sweeps_reference": sweeps_ref,
            "executed_gupdates_per_s": round(sweeps_exec * upd_row / t_graph / 1e9, 2),
            "reference_equivalent_gupdates_per_s": round(sweeps_ref * upd_row / t_graph / 1e9, 2),
            "checkpoint_every": CKPT,
            "checkpointed_ms": round(t_ckpt * 1e3, 2),
            "checkpointed_bitwise_equal": ckpt_equal,
            "adjoint_row_bytes": {"store_all": 4 * nz * nx * 8,
                                  "checkpointed": (4 * (CKPT + 2) + 2 * (-(-(nz - 1) // CKPT)) + 2) * nx * 8},
            "per_frequency_loop_ms": round(t_plain * 1e3, 1),
            "per_frequency_loop_sample": f"{len(fsample)} of {nw} frequencies timed, one call per frequency "
                                         f"with all {ns} shots ({t_plain_sample:.2f} s), scaled to all",
            "speedup_vs_per_frequency_loop": round(t_plain / t_graph, 2),
            "rel_err_per_frequency_vs_batched": rel,
            "fd_check": fd}


# Uneven batch: 64 shots over a 4096-wide model, shot s with its own aperture
# nx_s in {256, 512, 1024, 2048} (the window centred on the shot) and its own
# band: nw_s in 32..192 frequencies 0.1 Hz apart from a start in [2, 5) Hz;
# nz = 400, SSF, velocity = linear gradient + 5% smooth field (seed 5).
UNEVEN = dict(ns=64, nz=400, width=4096, widths=(256, 512, 1024, 2048), seed=5)


def uneven_survey():
    """The uneven survey: slowness of the 4096-wide model and, per shot, its
    aperture (x0, nx) and band (f0 + 0.1 Hz steps, nw frequencies)."""
    from scipy.ndimage import gaussian_filter
    c = UNEVEN
    rng = np.random.default_rng(c["seed"])
    nz, W0 = c["nz"], c["width"]
    z = np.arange(nz)[:, None] * 10.0
    v = 1500.0 + 2000.0 * z / ((nz - 1) * 10.0) + np.zeros((nz, W0))
    pert = gaussian_filter(rng.standard_normal((nz, W0)), 8.0)
    v = v + 0.05 * v * pert / np.max(np.abs(pert))
    slow = (1.0 / v ** 2).astype(np.complex64)
    shots = []
    for _ in range(c["ns"]):
        nx = int(rng.choice(c["widths"]))
        nw = int(rng.integers(32, 193))
        f0 = 2.0 + 3.0 * rng.random()
        cx = int(rng.integers(nx // 2, W0 - nx // 2 + 1))
        freqs = [float(np.float32(f0 + 0.1 * i)) for i in range(nw)]
        shots.append((cx - nx // 2, nx, freqs))
    return slow, shots


def uneven_batch_bench(sample_rows: int):
    """Down forward of every (shot, frequency) row of the uneven survey as one
    batch plan run (wemfwi_batch_run; also the one-call form that plans each
    time) against one call per row on a sample of rows, and against one call
    per shot; the sampled rows must be bitwise equal between the batch and the
    per-row calls."""
    import torch
    import wemfwi
    from wemfwi.core import BatchPlan, oneway_batch
    c = UNEVEN
    nz = c["nz"]
    slow, survey = uneven_survey()
    dev = torch.device("cuda:0")
    models, rows_m, rows_f, rows_shot = [], [], [], []
    nxs, nws = [], []
    for s, (x0, nx, freqs) in enumerate(survey):
        nxs.append(nx)
        nws.append(len(freqs))
        win = slow[:, x0:x0 + nx]
        models.append(wemfwi.Model(np.ascontiguousarray(win), 10.0, 10.0, nref=1, prop="ssf", eps=0.04))
        for f in freqs:
            rows_m.append(s)
            rows_f.append(f)
            rows_shot.append(s)
    n = len(rows_m)
    src, out = [], []
    spec = ricker_spectrum(rows_f)
    for r in range(n):
        nx = nxs[rows_m[r]]
        a = torch.zeros((nz, nx), dtype=torch.complex64, device=dev)
        a[0, nx // 2] = complex(spec[r])
        src.append(a)
        out.append(torch.empty_like(a))
    ms = [models[m] for m in rows_m]
    upd = sum(nxs[m] * (nz - 1) for m in rows_m)

    plan = BatchPlan(ms, rows_f, src, out)
    plan.run(False, False)
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        plan.run(False, False)
    torch.cuda.synchronize()
    t_b = (time.perf_counter() - t0) / reps
    # the one-call form (plan built and dropped inside the call)
    oneway_batch(False, False, ms, rows_f, src, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    oneway_batch(False, False, ms, rows_f, src, out)
    torch.cuda.synchronize()
    t_once = time.perf_counter() - t0
    # one call per shot (all its frequencies in one launch)
    by_shot = {}
    for r in range(n):
        by_shot.setdefault(rows_shot[r], []).append(r)
    stacked = {s: (torch.stack([src[r] for r in rr]), torch.empty((len(rr), nz, nxs[s]), dtype=torch.complex64,
                                                                  device=dev)) for s, rr in by_shot.items()}

    def per_shot():
        for s, rr in by_shot.items():
            models[s].oneway_recursive(False, False, [rows_f[r] for r in rr], stacked[s][0], stacked[s][1])

    per_shot()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_shot()
    torch.cuda.synchronize()
    t_s = time.perf_counter() - t0
    # one call per row on a sample
    stride = max(1, n // sample_rows)
    sample = list(range(0, n, stride))[:sample_rows]
    one = [torch.empty_like(out[r]) for r in sample]

    def per_row():
        for k, r in enumerate(sample):
            models[rows_m[r]].oneway_recursive(False, False, [rows_f[r]], src[r][None], one[k][None])

    per_row()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    per_row()
    torch.cuda.synchronize()
    t_r = (time.perf_counter() - t0) / len(sample) * n
    equal = all(bool(torch.equal(one[k], out[r])) for k, r in enumerate(sample))
    plan.close()
    for m in models:
        m.close()
    return {"config": f"ssf ns=64 nz=400 nx_s in {{256,512,1024,2048}} nw_s in 32..192, {n} rows",
            "updates": upd, "batched_ms": round(t_b * 1e3, 2),
            "batched_gupdates_per_s": round(upd / t_b / 1e9, 2),
            "one_call_ms": round(t_once * 1e3, 2),
            "per_shot_ms": round(t_s * 1e3, 2),
            "per_row_loop_ms": round(t_r * 1e3, 1),
            "per_row_sample": f"{len(sample)} of {n} rows timed ({t_r * len(sample) / n:.2f} s), scaled",
            "speedup_vs_row_loop": round(t_r / t_b, 1), "speedup_vs_shot_loop": round(t_s / t_b, 2),
            "bitwise_equal_on_sample": equal}


UNEVEN_NH = 11  # subsurface offsets of the uneven migration


def uneven_image_bench(sample_rows: int):
    """Subsurface-offset migration (nh = 11) of the uneven survey: every
    (shot, frequency) row of the 64 shots in one call
    (wemfwi_uneven_offset_image_adjoint: batched Down and Up^H sweeps, FP64
    image) against one call per (shot, frequency) row on a sample of rows
    (scaled to all rows); the sampled rows' image from the per-row calls is
    compared with one batched call over the same rows."""
    import torch
    import wemfwi
    from wemfwi.ops import ImageShot, uneven_offset_image_adjoint
    c = UNEVEN
    nz, W0, nh = c["nz"], c["width"], UNEVEN_NH
    slow, survey = uneven_survey()
    rng = np.random.default_rng(55)
    dev = torch.device("cuda:0")
    shots, parts = [], []
    for (x0, nx, freqs) in survey:
        m = wemfwi.Model(np.ascontiguousarray(slow[:, x0:x0 + nx]), 10.0, 10.0, nref=1, prop="ssf", eps=0.04)
        rix = np.arange(0, nx, 2)
        riz = np.zeros_like(rix)
        nw = len(freqs)
        data = torch.from_numpy((rng.standard_normal((len(rix), nw)) + 1j * rng.standard_normal((len(rix), nw)))
                                .astype(np.complex64)).to(dev)
        wav = ricker_spectrum(freqs)
        shots.append(ImageShot(m, x0, freqs, wav, 0, nx // 2, riz, rix, data))
        parts.append((m, x0, nx, freqs, wav, riz, rix, data))
    rows = [(s, i) for s in range(len(survey)) for i in range(len(survey[s][2]))]
    upd = 2.0 * sum(survey[s][1] * (nz - 1) for s, _ in rows)
    img = torch.zeros((nh, nz, W0), dtype=torch.complex64, device=dev)
    # warm-up: plans, phase tables, and the batch plans' launch orders (the
    # in-place sweeps of the call time one candidate order per call until
    # each has run twice)
    for _ in range(9):
        uneven_offset_image_adjoint(shots, nh, img)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    uneven_offset_image_adjoint(shots, nh, img)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    # one call per (shot, frequency) row, on a sample
    stride = max(1, len(rows) // sample_rows)
    sample = rows[::stride][:sample_rows]
    one = []
    for s, i in sample:
        m, x0, nx, freqs, wav, riz, rix, data = parts[s]
        one.append(ImageShot(m, x0, [freqs[i]], [wav[i]], 0, nx // 2, riz, rix, data[:, i:i + 1].contiguous()))
    loop_img = torch.zeros_like(img)

    def loop():
        loop_img.zero_()
        for sh in one:
            uneven_offset_image_adjoint([sh], nh, loop_img, add=True)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop()
    torch.cuda.synchronize()
    t_loop_sample = time.perf_counter() - t0
    # the same rows batched in one call
    by_shot = {}
    for s, i in sample:
        by_shot.setdefault(s, []).append(i)
    grouped = []
    for s, ii in by_shot.items():
        m, x0, nx, freqs, wav, riz, rix, data = parts[s]
        grouped.append(ImageShot(m, x0, [freqs[i] for i in ii], [wav[i] for i in ii], 0, nx // 2, riz, rix,
                                 data[:, ii].contiguous()))
    bimg = torch.zeros_like(img)
    uneven_offset_image_adjoint(grouped, nh, bimg)
    torch.cuda.synchronize()
    rel = float(torch.linalg.vector_norm((loop_img - bimg).to(torch.complex128)) /
                torch.linalg.vector_norm(bimg.to(torch.complex128)))
    t_loop = t_loop_sample * len(rows) / len(sample)
    grad = extended_gradient_bench(shots, one, grouped, img, nh, len(rows), len(sample),
                                   3.0 * sum(survey[s][1] * (nz - 1) for s, _ in rows))
    for m, *_ in parts:
        m.close()
    return {"config": f"ssf ns={len(survey)} nz={nz} nx_s in {{256,512,1024,2048}} nw_s in 32..192, "
                      f"{len(rows)} rows, nh={nh}, image {nh}x{nz}x{W0} (FP64 reduction)",
            "updates": upd, "batched_ms": round(t_b * 1e3, 2),
            "batched_gupdates_per_s": round(upd / t_b / 1e9, 2),
            "per_row_loop_ms": round(t_loop * 1e3, 1),
            "per_row_sample": f"{len(sample)} of {len(rows)} rows timed ({t_loop_sample:.2f} s), scaled",
            "speedup_vs_row_loop": round(t_loop / t_b, 1),
            "rel_err_row_loop_vs_batched_on_sample": rel,
            "extended_fwi_gradient": grad}


def extended_gradient_bench(shots, one, grouped, model_img, nh, nrows, nsample, upd):
    """The extended least-squares objective and gradient of the uneven survey
    (wemfwi_uneven_offset_gradient: Down, offset scatter, Up, residual, Up^H,
    offset image; the shots' data taken as observed data, the migration image
    as the model) in one call, against one call per (shot, frequency) row on
    the rows sampled for the migration (scaled); the sampled rows' summed
    misfit and gradient from the per-row calls are compared with one batched
    call over the same rows."""
    import torch
    from wemfwi.ops import uneven_offset_gradient
    I = model_img.clone()
    scale = float(torch.linalg.vector_norm(I)) or 1.0
    I.mul_(1.0 / scale)
    g = torch.zeros_like(I)
    for _ in range(9):                        # warm-up, launch orders as above
        uneven_offset_gradient(shots, nh, I, g)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    J = uneven_offset_gradient(shots, nh, I, g)
    torch.cuda.synchronize()
    t_b = time.perf_counter() - t0
    lg = torch.zeros_like(I)

    def loop():
        lg.zero_()
        return sum(uneven_offset_gradient([sh], nh, I, lg, add=True) for sh in one)

    loop()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Jl = loop()
    torch.cuda.synchronize()
    t_ls = time.perf_counter() - t0
    bg = torch.zeros_like(I)
    Jb = uneven_offset_gradient(grouped, nh, I, bg)
    torch.cuda.synchronize()
    rel = float(torch.linalg.vector_norm((lg - bg).to(torch.complex128)) /
                torch.linalg.vector_norm(bg.to(torch.complex128)))
    t_loop = t_ls * nrows / nsample
    return {"updates": upd, "sweeps_per_row": 3, "batched_ms": round(t_b * 1e3, 2),
            "batched_gupdates_per_s": round(upd / t_b / 1e9, 2), "misfit": J,
            "per_row_loop_ms": round(t_loop * 1e3, 1),
            "per_row_sample": f"{nsample} of {nrows} rows timed ({t_ls:.2f} s), scaled",
            "speedup_vs_row_loop": round(t_loop / t_b, 1),
            "rel_err_row_loop_vs_batched_on_sample": rel,
            "misfit_rel_err_row_loop_vs_batched_on_sample": abs(Jl - Jb) / Jb}


def sweep_line(config: str, reps: int):
    """Down forward (sources -> wavefield) and Down adjoint sweeps of a
    configuration's rows on their own: the sweep kernel's mean launch time
    (HIP events inside the library) against the HBM roofline."""
    import torch
    import wemfwi
    cfg = CONFIGS[config]
    nx, nz, nw, ns = cfg["nx"], cfg["nz"], cfg["nw"], cfg["ns"]
    freqs = (cfg["fmin"] + np.arange(nw) * (cfg["fmax"] - cfg["fmin"]) / (nw - 1)).astype(np.float32)
    src_ix = [32 + 64 * s for s in range(ns)]
    slow = (1.0 / velocity(cfg) ** 2).astype(np.complex64)
    model = wemfwi.Model(slow, 10.0, 10.0, nref=cfg["nref"], prop=cfg["prop"], eps=0.04)
    nrows = ns * nw
    f_rows = np.tile(freqs, ns)
    dev = torch.device("cuda:0")
    S = torch.zeros((nrows, nz, nx), dtype=torch.complex64, device=dev)
    spec = torch.from_numpy(np.tile(ricker_spectrum(freqs), ns)).to(dev)
    ix = torch.tensor(np.repeat(src_ix, nw), device=dev)
    S[torch.arange(nrows, device=dev), 0, ix] = spec
    W = torch.empty_like(S)

    def step():
        model.oneway_recursive(False, False, f_rows, S, W)
        model.oneway(False, True, f_rows, W)

    step()  # warm-up (the autotuner decides here)
    torch.cuda.synchronize()
    wemfwi.timing(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wemfwi.timing(False)
    k_ms, k_n = wemfwi.kernel_time("oneway_kernel")
    avg = k_ms / max(k_n, 1)
    upd = nx * (nz - 1) * nrows
    achieved = BYTES_PER_UPDATE * upd / (avg / 1e3) / 1e9
    finite = bool(torch.isfinite(W.abs()).all())
    model.close()
    del S, W
    return {"config": f"{cfg['prop']} nref={cfg['nref']} nx={nx} nz={nz} nw={nw} ns={ns} ({nrows} rows)",
            "kernel": "oneway_kernel", "kernel_ms": round(avg, 3),
            "gupdates_per_s": round(upd / (avg / 1e3) / 1e9, 2),
            "wall_gupdates_per_s": round(2 * upd * reps / (t1 - t0) / 1e9, 2),
            "achieved_gbs": round(achieved, 1), "frac": round(achieved / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": BYTES_PER_UPDATE * upd, "finite": finite}


def any_length_line(nrows: int):
    """Down sweeps (SSF nref = 1, nz = 300, full rows) at the reference's own
    FFT lengths nk = 500 and 600 (rows of 512 and 1024 lanes, compile-time
    pass chains; also timed with the runtime chain, WEMFWI_GEN_FIXED=0) and at
    the neighbouring powers of two: the sweep kernel's mean launch time (HIP
    events inside the library), nrows rows each."""
    import torch
    import wemfwi
    nz = 300
    out = {"config": f"ssf nref=1 nz={nz}, {nrows} rows of 2-30 Hz, Down forward in place"}
    for nk, chain in ((500, "fixed"), (500, "runtime"), (512, "pow2"), (600, "fixed"), (600, "runtime"),
                      (1024, "pow2")):
        os.environ["WEMFWI_GEN_FIXED"] = "0" if chain == "runtime" else "1"
        rng = np.random.default_rng(3)
        v = 1500.0 + 2000.0 * np.linspace(0, 1, nz)[:, None] + 50.0 * rng.standard_normal((nz, nk))
        model = wemfwi.Model((1.0 / v ** 2).astype(np.complex64), 10.0, 10.0, nref=1, prop="ssf", eps=0.04)
        f_rows = (2.0 + 28.0 * np.arange(nrows) / max(nrows - 1, 1)).astype(np.float32)
        W = torch.zeros((nrows, nz, nk), dtype=torch.complex64, device="cuda:0")
        W[:, 0, nk // 3] = 1.0
        model.oneway(False, False, f_rows, W)  # warm-up (the autotuner decides here)
        torch.cuda.synchronize()
        wemfwi.timing(True)
        for _ in range(3):
            model.oneway(False, False, f_rows, W)
        torch.cuda.synchronize()
        k_ms, k_n = wemfwi.kernel_time("oneway_kernel")
        wemfwi.timing(False)
        avg = k_ms / max(k_n, 1)
        out[f"nk{nk}_{chain}"] = {"kernel_ms": round(avg, 3),
                                  "gupdates_per_s": round(nk * (nz - 1) * nrows / (avg / 1e3) / 1e9, 1),
                                  "finite": bool(torch.isfinite(torch.view_as_real(W)).all())}
        model.close()
        del W
        torch.cuda.empty_cache()
    os.environ.pop("WEMFWI_GEN_FIXED", None)
    return out


def measured_traffic(config: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this configuration (mean over its launches, each
    counter scaled by its copy-kernel calibration factor), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{config}_traffic.md")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        for ln in f:
            if not ln.startswith("bench-traffic:"):
                continue
            kv = dict(tok.split("=", 1) for tok in ln.split()[1:])
            if kv.get("kernel") != kernel or kv.get("config") != config:
                continue
            fetch = [float(x) for x in kv["fetch_kib"].split(",")]
            write = [float(x) for x in kv["write_kib"].split(",")]
            ff, wf = float(kv["fetch_factor"]), float(kv["write_factor"])
            per = [1024.0 * (ff * a + wf * b) for a, b in zip(fetch, write)]
            return sum(per) / len(per), os.path.relpath(path, ROOT)
    return None, None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="pspi1024", choices=sorted(CONFIGS))
    ap.add_argument("--cpu-sample", type=int, default=4096,
                    help="rows of the CPU baseline sample (0: skip)")
    ap.add_argument("--fwi-sample", type=int, default=16,
                    help="frequencies of the per-frequency loop in the FWI gradient comparison (0: skip)")
    ap.add_argument("--fd-directions", type=int, default=20,
                    help="central-difference directions of the FWI gradient check (0: skip)")
    ap.add_argument("--uneven-sample", type=int, default=32,
                    help="rows of the per-row loop in the uneven-batch comparison (0: skip)")
    ap.add_argument("--any-length-rows", type=int, default=1024,
                    help="rows of the nk = 500 / 600 sweep comparison (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device (calibrates FETCH_SIZE/WRITE_SIZE) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the (shot, frequency) rows of this path are "
              "independent and would be split across GPUs with one final image reduction; "
              "only --gpus 1 is run here.", file=sys.stderr)
        sys.exit(2)

    import torch
    import wemfwi

    cfg = CONFIGS[args.config]
    nx, nz, nw, ns = cfg["nx"], cfg["nz"], cfg["nw"], cfg["ns"]
    freqs = (cfg["fmin"] + np.arange(nw) * (cfg["fmax"] - cfg["fmin"]) / (nw - 1)).astype(np.float32)
    src_ix = [32 + 64 * s for s in range(ns)] if ns > 1 else [nx // 2]
    v = velocity(cfg)
    slow = (1.0 / v ** 2).astype(np.complex64)
    model = wemfwi.Model(slow, 10.0, 10.0, nref=cfg["nref"], prop=cfg["prop"], eps=0.04)

    dev = torch.device("cuda:0")
    nrows = ns * nw
    # rows ordered (shot, frequency); sources injected at the surface
    f_rows = np.tile(freqs, ns)
    S = torch.zeros((nrows, nz, nx), dtype=torch.complex64, device=dev)
    spec = torch.from_numpy(np.tile(ricker_spectrum(freqs), ns)).to(dev)
    ix = torch.tensor(np.repeat(src_ix, nw), device=dev)
    S[torch.arange(nrows, device=dev), 0, ix] = spec
    W = torch.empty_like(S)

    def step():
        model.set_freqs(freqs)
        model.oneway_recursive(False, False, f_rows, S, W)
        model.oneway(False, True, f_rows, W)

    # eager steps: warm-up (the sweep autotuner decides here) and the
    # per-launch kernel timing of the roofline
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    wemfwi.timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wemfwi.timing(False)
    eager_ms = (t1 - t0) * 1e3 / args.steps
    # the timed steps: the same work (phase tables of the 256 frequencies in
    # place, Down forward S -> W, Down adjoint in place) as one captured
    # hipGraph of batch-plan launches, replayed
    from wemfwi.core import BatchGraph, BatchPlan
    rows_s = [S[i] for i in range(nrows)]
    rows_w = [W[i] for i in range(nrows)]
    p_fwd = BatchPlan([model] * nrows, f_rows, rows_s, rows_w)
    p_adj = BatchPlan([model] * nrows, f_rows, rows_w, rows_w)
    graph = BatchGraph([(p_fwd, False, False, True), (p_adj, False, True, False)])
    for _ in range(args.warmup):
        graph.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.launch()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    upd_sweep = nx * (nz - 1) * nrows
    value = 2 * upd_sweep / (ms_step / 1e3) / 1e9
    k_ms, k_n = wemfwi.kernel_time("oneway_kernel")
    t_ms, t_n = wemfwi.kernel_time("phase_table_kernel")
    avg = k_ms / max(k_n, 1)
    achieved = BYTES_PER_UPDATE * upd_sweep / (avg / 1e3) / 1e9
    if not np.isfinite(float(W.abs().max())):
        print("non-finite wavefield", file=sys.stderr)
        sys.exit(1)

    # forward and adjoint sweeps timed on their own (phase tables in place)
    def timed(fn, reps):
        torch.cuda.synchronize()
        a = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - a) / reps
    reps = max(1, args.steps)
    t_fwd = timed(lambda: model.oneway_recursive(False, False, f_rows, S, W), reps)
    t_adj = timed(lambda: model.oneway(False, True, f_rows, W), reps)
    traffic, traffic_src = measured_traffic(args.config, "oneway_kernel")
    line = {
        "metric": "Gupdates/s (x·z·ω) fwd+adj",
        "value": round(value, 3),
        "unit": "Gupdates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded velocity model, Ricker sources)",
        "config": {"workload": f"{cfg['prop']} downward continuation nref={cfg['nref']} "
                               f"nx={nx} nz={nz} nw={nw} {cfg['fmin']:g}-{cfg['fmax']:g} Hz "
                               f"ns={ns}",
                   "rows": nrows, "updates_per_step": 2 * upd_sweep,
                   "launch": "hipGraph of phase tables + batched Down forward + Down adjoint",
                   "eager_ms_per_step": round(eager_ms, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": traffic_src,
                     "algorithmic_bytes": BYTES_PER_UPDATE * upd_sweep,
                     "kernel": "oneway_kernel", "kernel_ms": round(avg, 3),
                     "phase_table_ms": round(t_ms / max(t_n, 1), 3)},
    }
    # the same kernel against the FP32 VALU roofline (PSPI with 4 references
    # sits near the ridge: both fractions are reported)
    flops = sweep_flops(slow, cfg["nref"], cfg["prop"], nrows)
    tfl = flops / (avg / 1e3) / 1e12
    line["roofline"]["valu"] = {"achieved": round(tfl, 2), "peak": VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": round(tfl / VALU_PEAK_TFLOPS, 4),
                                "frac_of_scalar_fp32": round(tfl / VALU_SCALAR_TFLOPS, 4),
                                "flops_per_launch": flops,
                                "flop_per_byte": round(flops / (BYTES_PER_UPDATE * upd_sweep), 2)}
    # the sweep autotuner's radix-plan and workgroup-tile choices for this workload's launches
    line["autotune"] = [e for e in wemfwi.autotune_report()
                        if (e["kernel"] == "oneway_kernel" and e["N"] == nx)
                        or (e["kernel"] == "oneway_tile" and e["N"] == nx and e["rows"] == nrows)]
    line["fwd_gupdates_per_s"] = round(upd_sweep / t_fwd / 1e9, 3)
    line["adj_gupdates_per_s"] = round(upd_sweep / t_adj / 1e9, 3)
    if args.config == "pspi1024":
        line["ssf_sweep"] = sweep_line("ssf1024", reps)
    if args.fwi_sample > 0:
        line["fwi_gradient"] = fwi_gradient_bench(args.fwi_sample, args.fd_directions)
    if args.uneven_sample > 0:
        line["uneven_batch"] = uneven_batch_bench(args.uneven_sample)
        # the batch plan's launch-order decision (groups, order, best time of each candidate)
        rep = wemfwi.autotune_report()
        line["uneven_batch"]["launch_order"] = [e for e in rep if e["kernel"] == "batch_order"]
        # the workgroup tile each FFT-length group of the batch runs with
        groups = {(g["N"], g["rows"]) for e in line["uneven_batch"]["launch_order"] for g in e["groups"]}
        line["uneven_batch"]["tiles"] = [{"N": e["N"], "rows": e["rows"], "tile": e["tile"], "ms": e["ms"]}
                                         for e in rep if e["kernel"] == "oneway_tile" and (e["N"], e["rows"]) in groups]
        line["uneven_offset_image"] = uneven_image_bench(args.uneven_sample)
    if args.any_length_rows > 0:
        line["any_length_sweeps"] = any_length_line(args.any_length_rows)
    if args.cpu_sample > 0:
        line["cpu_baseline"] = cpu_baseline(cfg, slow, freqs, src_ix, args.cpu_sample)
        if args.fwi_sample > 0:
            line["fwi_gradient"]["fp64_oracle_check"] = fwi_oracle_check()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
