"""The sweep autotuner: every radix plan it may choose (FftPlan<N, K>, forced
with WEMFWI_FFT_PLAN) matches the FP64 oracle on the Down/Up recursions
(relative L2 <= 1e-5) for full and padded rows, SSF and PSPI, forward and
adjoint; the tuner records one decision per (length, variant) with a timing
for every plan and uses the fastest."""

import os

import numpy as np
import pytest

from oracle import wem_oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-5

CASES = [
    # nx, tap, pad(taper), nref, prop     -> N, plans
    (1024, 0, 0, 1, "ssf"),                # 1024, full row
    (1000, 12, 8, 2, "pspi"),              # 1024, padded
    (2048, 0, 0, 2, "pspi"),               # 2048, full row
    (2040, 4, 0, 1, "ssf"),                # 2048, padded
]


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _slow(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    return (1 / v ** 2).astype(np.complex64)


@pytest.fixture
def forced_plan():
    old = os.environ.get("WEMFWI_FFT_PLAN")
    yield
    if old is None:
        os.environ.pop("WEMFWI_FFT_PLAN", None)
    else:
        os.environ["WEMFWI_FFT_PLAN"] = old


@pytest.mark.parametrize("nx,tap,pad,nref,prop", CASES)
def test_every_plan_matches_oracle(nx, tap, pad, nref, prop, forced_plan):
    import torch
    from wemfwi import Model
    nz = 12
    s = _slow(nz, nx, 1)
    m = Model(s, 10.0, 10.0, tap=tap, pad=pad, nref=nref, prop=prop, eps=0.04)
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=pad, nref=nref, eps=0.04, prop=prop)
    ref = O.refsampler(s, nref)
    rng = np.random.default_rng(2)
    W = np.zeros((2, nz, nx), dtype=np.complex64)
    W[:, 0, nx // 3: nx // 3 + 7] = rng.standard_normal((2, 7)) + 1j * rng.standard_normal((2, 7))
    W[:, nz - 1, nx // 2: nx // 2 + 7] = rng.standard_normal((2, 7)) + 1j * rng.standard_normal((2, 7))
    freqs = np.array([4.0, 23.0], dtype=np.float32)
    nplans = 3 if nx + 2 * tap == 1024 else 2
    for k in range(nplans):
        os.environ["WEMFWI_FFT_PLAN"] = str(k)
        for up, adjoint in ((False, False), (False, True), (True, False), (True, True)):
            print(f"plan {k} nk={nx + 2 * tap} up={up} adjoint={adjoint}", flush=True)
            t = torch.from_numpy(W).cuda()
            m.oneway(up, adjoint, freqs, t)
            fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
                  (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adjoint)]
            got = t.cpu().numpy()
            for i in range(2):
                want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
                assert rel(got[i], want) < TOL, (k, up, adjoint, i, rel(got[i], want))


def test_tuner_records_one_decision_per_variant(forced_plan):
    import torch
    import wemfwi
    from wemfwi import Model
    os.environ.pop("WEMFWI_FFT_PLAN", None)
    nz, nx = 20, 2044   # padded 2048 row, PSPI adjoint: a key no other test of this file tunes
    s = _slow(nz, nx, 3)
    m = Model(s, 10.0, 10.0, tap=2, nref=2, prop="pspi", eps=0.04)
    W = torch.zeros((64, nz, nx), dtype=torch.complex64, device="cuda")
    W[:, 0, 100] = 1.0
    freqs = np.linspace(3, 20, 64).astype(np.float32)
    m.oneway(False, True, freqs, W)
    m.oneway(False, True, freqs, W)
    rep = [e for e in wemfwi.autotune_report() if e["kernel"] == "oneway_kernel" and e["N"] == 2048
           and e["full"] == 0 and e["ssf"] == 0 and e["adjoint"] == 1]
    assert len(rep) == 1
    e = rep[0]
    assert len(e["ms"]) == 2 and all(v > 0 for v in e["ms"])
    assert e["plan"] == int(np.argmin(e["ms"]))
    assert np.isfinite(float(W.abs().max()))


@pytest.fixture
def tile_env():
    keys = ("WEMFWI_SWEEP_TILE", "WEMFWI_TABLE_DMA", "WEMFWI_FFT_PLAN")
    old = {k: os.environ.get(k) for k in keys}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


TILE_CASES = [
    # nx, tap, nref, prop, table DMA, recursive (source-row prefetch)
    (1024, 0, 1, "ssf", "0", True),
    (1024, 0, 2, "pspi", "1", False),
    (1000, 12, 2, "pspi", "0", False),   # padded 1024 row
    (2048, 0, 2, "pspi", "1", True),
    (2040, 4, 1, "ssf", "0", False),     # padded 2048 row
    (600, 0, 1, "ssf", "0", False),      # nk = 600: runtime-planned passes in a 1024 row
    (1000, 0, 2, "pspi", "1", True),     # nk = 1000, runtime-planned
]


@pytest.mark.parametrize("nx,tap,nref,prop,tdma,recursive", TILE_CASES)
def test_half_tile_is_bitwise_the_full_tile(nx, tap, nref, prop, tdma, recursive, tile_env):
    """The workgroup tile only changes how rows spread over the CUs: both
    tiles give bit-identical sweeps (every direction), and the full tile's
    rows match the oracle."""
    import torch
    from wemfwi import Model
    os.environ["WEMFWI_TABLE_DMA"] = tdma
    nz, nrows = 10, 11   # 11 rows: a partial last workgroup in both tiles
    s = _slow(nz, nx, 5)
    m = Model(s, 10.0, 10.0, tap=tap, nref=nref, prop=prop, eps=0.04)
    rng = np.random.default_rng(6)
    W = np.zeros((nrows, nz, nx), dtype=np.complex64)
    W[:, 0, nx // 3: nx // 3 + 5] = rng.standard_normal((nrows, 5)) + 1j * rng.standard_normal((nrows, 5))
    W[:, nz - 1, nx // 2: nx // 2 + 5] = rng.standard_normal((nrows, 5)) + 1j * rng.standard_normal((nrows, 5))
    W[:, 4, 7] = 1.0
    freqs = np.linspace(3.0, 25.0, nrows).astype(np.float32)
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=0, nref=nref, eps=0.04, prop=prop)
    ref = O.refsampler(s, nref)
    for up, adjoint in ((False, False), (False, True), (True, False), (True, True)):
        got = {}
        for tile in ("full", "half"):
            os.environ["WEMFWI_SWEEP_TILE"] = tile
            print(f"tile {tile} nk={nx + 2 * tap} {prop} tdma={tdma} recursive={recursive} "
                  f"up={up} adjoint={adjoint}", flush=True)
            src = torch.from_numpy(W).cuda()
            if recursive:
                out = torch.empty_like(src)
                m.oneway_recursive(up, adjoint, freqs, src, out)
            else:
                out = src
                m.oneway(up, adjoint, freqs, out)
            got[tile] = out.cpu().numpy()
        assert np.array_equal(got["full"].view(np.uint32), got["half"].view(np.uint32)), (up, adjoint)
        if not recursive:
            fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
                  (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adjoint)]
            i = nrows - 1
            want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
            assert rel(got["full"][i], want) < TOL, (up, adjoint, rel(got["full"][i], want))


def test_tile_tuner_times_each_row_count_once(tile_env):
    import torch
    import wemfwi
    from wemfwi import Model
    for k in ("WEMFWI_SWEEP_TILE", "WEMFWI_TABLE_DMA", "WEMFWI_FFT_PLAN"):
        os.environ.pop(k, None)
    nz, nx = 24, 1000   # padded 1024 row, SSF adjoint
    s = _slow(nz, nx, 7)
    m = Model(s, 10.0, 10.0, tap=12, nref=1, prop="ssf", eps=0.04)

    def entries():
        return [e for e in wemfwi.autotune_report() if e["kernel"] == "oneway_tile" and e["N"] == 1024
                and e["full"] == 0 and e["ssf"] == 1 and e["adjoint"] == 1 and e["rows"] in (37, 300)]

    for nrows in (37, 300, 37, 300):
        W = torch.zeros((nrows, nz, nx), dtype=torch.complex64, device="cuda")
        W[:, 0, 100] = 1.0
        m.oneway(False, True, np.linspace(3, 20, nrows).astype(np.float32), W)
        assert np.isfinite(float(W.abs().max()))
    rep = entries()
    assert len(rep) == 2, rep
    for e in rep:
        assert e["steps"] >= 2 and all(v > 0 for v in e["ms"])
        assert e["tile"] == ("half" if e["ms"][1] < 0.98 * e["ms"][0] else "full")
