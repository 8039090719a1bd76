"""Uneven batching on the GPU: Down/Up recursions of rows with different
models (apertures, depth counts, FFT lengths, SSF/PSPI) and frequencies in one
wemfwi_oneway_batch call must equal, bit for bit, the same rows swept one
model at a time through wemfwi_oneway_recursive (same kernels, same tables);
one row per model is also checked against the FP64 oracle (1e-5 relative L2)."""

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

# (nz, nx, tap, prop, nref): FFT lengths 128, 256 (full), 256 (padded), 512
MODELS = [(20, 100, 14, "ssf", 2), (24, 256, 0, "ssf", 1), (16, 200, 28, "pspi", 2), (18, 512, 0, "pspi", 3)]


def _slow(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.04 * rng.standard_normal((nz, nx)))
    return (1 / v ** 2).astype(np.complex64)


def _rows():
    rng = np.random.default_rng(3)
    rows = []
    for k in range(14):
        mi = int(rng.integers(len(MODELS)))
        rows.append((mi, float(np.float32(2.0 + 14.0 * rng.random()))))
    return rows


@pytest.mark.parametrize("up,adjoint", [(False, False), (True, False), (False, True), (True, True)])
def test_uneven_batch_equals_per_model_sweeps(up, adjoint):
    import torch
    from wemfwi import Model
    from wemfwi.core import oneway_batch
    slows = [_slow(nz, nx, 10 + i) for i, (nz, nx, _, _, _) in enumerate(MODELS)]
    models = [Model(s, 10.0, 10.0, tap=tap, pad=6, nref=nref, prop=prop, eps=0.04)
              for s, (nz, nx, tap, prop, nref) in zip(slows, MODELS)]
    rows = _rows()
    rng = np.random.default_rng(5)
    src = []
    for mi, f in rows:
        nz, nx = MODELS[mi][:2]
        a = (rng.standard_normal((nz, nx)) + 1j * rng.standard_normal((nz, nx))).astype(np.complex64)
        src.append(torch.from_numpy(a).cuda())
    out = [torch.zeros_like(a) for a in src]
    oneway_batch(up, adjoint, [models[mi] for mi, _ in rows], [f for _, f in rows], src, out)
    checked = set()
    for (mi, f), a, b in zip(rows, src, out):
        ref = torch.zeros_like(a)
        models[mi].oneway_recursive(up, adjoint, [f], a[None], ref[None])
        assert torch.equal(ref, b), (mi, f)
        if mi not in checked:
            checked.add(mi)
            nz, nx, tap, prop, nref = MODELS[mi]
            p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=6, nref=nref, eps=0.04, prop=prop)
            ref_t = O.refsampler(slows[mi], nref)
            fn = {(False, False): O.down_fwd, (True, False): O.up_fwd, (False, True): O.down_adj,
                  (True, True): O.up_adj}[(up, adjoint)]
            want = fn(a.cpu().numpy().astype(complex), 2 * np.pi * f, slows[mi], ref_t, p)
            got = b.cpu().numpy()
            assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    # in place: src is out
    inpl = [a.clone() for a in src]
    oneway_batch(up, adjoint, [models[mi] for mi, _ in rows], [f for _, f in rows], inpl, inpl)
    for a, b in zip(inpl, out):
        assert torch.equal(a, b)
    for m in models:
        m.close()


def test_batch_plan_runs_repeat_and_detect_rebuilt_tables():
    import torch
    from wemfwi import Model
    from wemfwi._lib import WemfwiError
    from wemfwi.core import BatchPlan, oneway_batch
    slows = [_slow(nz, nx, 20 + i) for i, (nz, nx, _, _, _) in enumerate(MODELS)]
    models = [Model(s, 10.0, 10.0, tap=tap, pad=6, nref=nref, prop=prop, eps=0.04)
              for s, (nz, nx, tap, prop, nref) in zip(slows, MODELS)]
    rows = _rows()
    rng = np.random.default_rng(8)
    src = [torch.from_numpy((rng.standard_normal(MODELS[mi][:2]) + 0j).astype(np.complex64)).cuda()
           for mi, _ in rows]
    ms, fs = [models[mi] for mi, _ in rows], [f for _, f in rows]
    want = [torch.zeros_like(a) for a in src]
    oneway_batch(True, False, ms, fs, src, want)
    out = [torch.zeros_like(a) for a in src]
    plan = BatchPlan(ms, fs, src, out)
    for _ in range(2):
        for o in out:
            o.zero_()
        plan.run(True, False)
        for a, b in zip(out, want):
            assert torch.equal(a, b)
    models[rows[0][0]].set_freqs([1.25])      # rebuilds that model's phase table
    with pytest.raises(WemfwiError):
        plan.run(True, False)
    plan.close()
    for m in models:
        m.close()


def test_batch_launch_order_is_timed_once_per_shape_and_reused():
    """A plan whose runs can be repeated times its candidate launch orders on
    its first run and keeps the fastest; a later plan of the same group shape
    (here the one-call form, and an in-place one that could not be timed)
    reuses that decision; every run stays bitwise equal to the per-model
    sweeps."""
    import torch
    import wemfwi
    from wemfwi import Model
    from wemfwi.core import BatchPlan, oneway_batch
    slows = [_slow(nz, nx, 30 + i) for i, (nz, nx, _, _, _) in enumerate(MODELS)]
    models = [Model(s, 10.0, 10.0, tap=tap, pad=6, nref=nref, prop=prop, eps=0.04)
              for s, (nz, nx, tap, prop, nref) in zip(slows, MODELS)]
    rows = _rows()[:11]     # a group shape no other test builds
    rng = np.random.default_rng(9)
    src = [torch.from_numpy((rng.standard_normal(MODELS[mi][:2]) + 0j).astype(np.complex64)).cuda()
           for mi, _ in rows]
    ms, fs = [models[mi] for mi, _ in rows], [f for _, f in rows]
    want = []
    for (mi, f), a in zip(rows, src):
        ref = torch.zeros_like(a)
        models[mi].oneway_recursive(False, False, [f], a[None], ref[None])
        want.append(ref)
    before = [e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"]
    out = [torch.zeros_like(a) for a in src]
    plan = BatchPlan(ms, fs, src, out)
    plan.run(False, False)
    after = [e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"]
    assert len(after) == len(before) + 1
    e = after[-1]
    ngroups = len(e["groups"])
    assert ngroups >= 2 and sorted(e["order"]) == list(range(ngroups))
    assert len(e["ms"]) == min(ngroups, 4) and all(v > 0 for v in e["ms"])
    assert e["order"][0] == int(np.argmin(e["ms"]))   # candidate c starts with group c
    for a, b in zip(out, want):
        assert torch.equal(a, b)
    plan.close()
    out2 = [torch.zeros_like(a) for a in src]
    oneway_batch(False, False, ms, fs, src, out2)
    inpl = [a.clone() for a in src]
    oneway_batch(False, False, ms, fs, inpl, inpl)
    assert len([e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"]) == len(after)
    for a, b, c in zip(out2, inpl, want):
        assert torch.equal(a, c) and torch.equal(b, c)
    for m in models:
        m.close()


def test_in_place_batch_times_one_order_per_call():
    """In-place plans cannot repeat a run: each call of a new group shape
    launches the candidate order tried least so far and times it; after two
    calls per candidate the shape has its decision.  Every call stays
    bitwise equal to the per-model sweeps."""
    import torch
    import wemfwi
    from wemfwi import Model
    from wemfwi.core import oneway_batch
    slows = [_slow(nz, nx, 40 + i) for i, (nz, nx, _, _, _) in enumerate(MODELS)]
    models = [Model(s, 10.0, 10.0, tap=tap, pad=6, nref=nref, prop=prop, eps=0.04)
              for s, (nz, nx, tap, prop, nref) in zip(slows, MODELS)]
    rows = _rows()[:9]     # a group shape no other test builds
    rng = np.random.default_rng(21)
    src = [torch.from_numpy((rng.standard_normal(MODELS[mi][:2]) + 0j).astype(np.complex64)).cuda()
           for mi, _ in rows]
    ms, fs = [models[mi] for mi, _ in rows], [f for _, f in rows]
    want = []
    for (mi, f), a in zip(rows, src):
        ref = torch.zeros_like(a)
        models[mi].oneway_recursive(True, True, [f], a[None], ref[None])
        want.append(ref)
    ngroups = len({(MODELS[mi][1], MODELS[mi][2], MODELS[mi][3]) for mi, _ in rows})
    before = len([e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"])
    calls = 2 * min(ngroups, 4)
    for k in range(calls):
        inpl = [a.clone() for a in src]
        oneway_batch(True, True, ms, fs, inpl, inpl)
        for a, b in zip(inpl, want):
            assert torch.equal(a, b), f"call {k}"
        n = len([e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"])
        assert n == before + (1 if k == calls - 1 else 0), f"call {k}: decision count {n}"
    e = [e for e in wemfwi.autotune_report() if e["kernel"] == "batch_order"][-1]
    assert len(e["ms"]) == min(ngroups, 4) and all(v > 0 for v in e["ms"])
    assert e["order"][0] == int(np.argmin(e["ms"]))
    for m in models:
        m.close()
