"""FFT lengths nk = nx + 2 tap that are not powers of two (the reference's
FFTW accepts any length; its own runs use 500 and 600): GPU parity of the
depth step, the Down/Up recursions, the scatterer, LinDown/LinUp and the
Born operators against the FP64 oracle (relative L2 <= 1e-5), the exact
dot test, and the error for lengths without a plan.  Lengths cover radix
4/2/3/5 register passes and direct-DFT prime radices (7, 11, 31, 61) in rows
of 128 to 2048 lanes."""

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


def _c(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.complex64))).cuda()


def _setup(nx, tap, pad, nref, prop, nz, seed=0, ntaylor=1):
    from wemfwi import Model
    s = _slow(nz, nx, seed)
    m = Model(s, 10.0, 10.0, tap=tap, pad=pad, nref=nref, prop=prop, eps=0.04, ntaylor=ntaylor)
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=pad, nref=nref, eps=0.04, prop=prop, ntaylor=ntaylor)
    return m, s, p, O.refsampler(s, nref)


CASES = [
    # nx, tap, pad(taper), nref, prop       nk and its passes; row lanes
    (100, 0, 0, 1, "ssf"),                 # 100 = 4.5.5; 128
    (90, 5, 6, 2, "pspi"),                 # 100 with FFT padding and taper; 128
    (300, 0, 0, 3, "pspi"),                # 300 = 4.3.5.5; 512
    (375, 5, 0, 1, "ssf"),                 # 385 = 5.7.11; 512
    (500, 0, 10, 2, "ssf"),                # 500 = 4.5.5.5 (a reference run length); 512
    (600, 0, 0, 4, "pspi"),                # 600 = 8.3.5.5 (a reference run length); 1024
    (1000, 0, 0, 1, "ssf"),                # 1000 = 8.5.5.5; 1024
    (1922, 15, 0, 2, "ssf"),               # 1952 = 16.2.61; 2048
    (1536, 0, 0, 1, "pspi"),               # 1536 = 16.16.2.3; 2048
]


@pytest.mark.parametrize("nx,tap,pad,nref,prop", CASES)
@pytest.mark.parametrize("adjoint", [False, True])
def test_onestep_any_length(nx, tap, pad, nref, prop, adjoint):
    import torch
    nz = 10
    m, s, p, ref = _setup(nx, tap, pad, nref, prop, nz)
    rng = np.random.default_rng(1)
    nrows = 4
    x = (rng.standard_normal((nrows, nx)) + 1j * rng.standard_normal((nrows, nx))).astype(np.complex64)
    freqs = np.array([2.0, 7.5, 17.25, 30.0], dtype=np.float32)
    depth = np.array([0, 3, 6, 9], dtype=np.int32)
    out = torch.zeros((nrows, nx), dtype=torch.complex64, device="cuda")
    print(f"onestep nk={nx + 2 * tap} adjoint={adjoint}", flush=True)
    m.onestep(adjoint, freqs, depth, _c(x), out)
    got = out.cpu().numpy()
    f = O.onestep_adj if adjoint else O.onestep_fwd
    for i in range(nrows):
        want = f(x[i].astype(complex), int(depth[i]), 2 * np.pi * float(freqs[i]), s, ref, p)
        assert rel(got[i], want) < TOL, (i, rel(got[i], want))


@pytest.mark.parametrize("nx,tap,pad,nref,prop", CASES)
@pytest.mark.parametrize("up,adjoint", [(False, False), (True, False), (False, True), (True, True)])
def test_oneway_any_length(nx, tap, pad, nref, prop, up, adjoint):
    nz = 24
    m, s, p, ref = _setup(nx, tap, pad, nref, prop, nz, seed=2)
    rng = np.random.default_rng(3)
    nrows = 2
    W = np.zeros((nrows, nz, nx), dtype=np.complex64)
    for r in range(nrows):
        for z0 in (0, nz // 2, nz - 1):
            W[r, z0, nx // 3: nx // 3 + 5] = rng.standard_normal(5) + 1j * rng.standard_normal(5)
    freqs = np.array([3.0, 21.0], dtype=np.float32)
    t = _c(W)
    print(f"oneway nk={nx + 2 * tap} up={up} adjoint={adjoint}", flush=True)
    m.oneway(up, adjoint, freqs, t)
    got = t.cpu().numpy()
    fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
          (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adjoint)]
    for i in range(nrows):
        want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
        assert rel(got[i], want) < TOL, (i, rel(got[i], want))


@pytest.mark.parametrize("nx,tap", [(100, 0), (288, 6), (600, 0)])
@pytest.mark.parametrize("mode", ["compat", "exact"])
def test_scatter_and_lin_any_length(nx, tap, mode):
    import torch
    from wemfwi.api import Hypercube, Vector
    from wemfwi.unit import LinDown, LinUp, Scatter
    nz = 12
    s = _slow(nz, nx, 4)
    par = {"tap": tap, "pad": 0, "nref": 2, "prop": "ssf", "eps": 0.04, "ntaylor": 2}
    p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=0, nref=2, eps=0.04, prop="ssf", ntaylor=2)
    ref = O.refsampler(s, 2)
    hyp = Hypercube([nx, nz], [0.0, 0.0], [10.0, 10.0])
    vs = Vector(hyp, arr=_c(s))
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(nx) + 1j * rng.standard_normal(nx)).astype(np.complex64)
    d = (rng.standard_normal(nx) + 1j * rng.standard_normal(nx)).astype(np.complex64)
    print(f"scatter/lin nk={nx + 2 * tap} mode={mode}", flush=True)
    op = Scatter(vs, 2, par, adjoint_mode=mode)
    op.setFreq(9.0)
    op.setDepth(4)
    w = 2 * np.pi * np.float32(9.0)
    y = torch.zeros(nx, dtype=torch.complex64, device="cuda")
    op.forward(_c(x), y)
    assert rel(y.cpu().numpy(), O.scatter_fwd(x.astype(complex), 4, w, s, p)) < TOL
    g = torch.zeros_like(y)
    op.adjoint(g, _c(d))
    assert rel(g.cpu().numpy(), O.scatter_adj(d.astype(complex), 4, w, s, p, mode)) < TOL
    bg = (rng.standard_normal((nz, nx)) + 1j * rng.standard_normal((nz, nx))).astype(np.complex64)
    dm = ((rng.standard_normal((nz, nx)) + 1j * rng.standard_normal((nz, nx))) * 1e-3).astype(np.complex64)
    dd = (rng.standard_normal((nz, nx)) + 1j * rng.standard_normal((nz, nx))).astype(np.complex64)
    for cls, fwd, adj in ((LinDown, O.lindown_fwd, O.lindown_adj), (LinUp, O.linup_fwd, O.linup_adj)):
        lop = cls(vs, par, _c(bg), adjoint_mode=mode)
        lop.setFreq(13.0)
        w = 2 * np.pi * np.float32(13.0)
        out = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
        lop.forward(_c(dm), out)
        assert rel(out.cpu().numpy(), fwd(dm.astype(complex), bg.astype(complex), w, s, ref, p)) < TOL
        gg = torch.zeros_like(out)
        lop.adjoint(gg, _c(dd))
        assert rel(gg.cpu().numpy(), adj(dd.astype(complex), bg.astype(complex), w, s, ref, p, mode=mode)) < TOL


def _born_case(nx, pad, tap, nz=20, nw=3, ns=2, nr=10, seed=17):
    rng = np.random.default_rng(seed)
    nt = 64
    freq = (2.0 + np.arange(nw) * 4.0).astype(np.float32)
    index = np.arange(nw) + 3
    wave_f = rng.standard_normal(nt) + 1j * rng.standard_normal(nt)
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=freq, index=index, wave_f=wave_f,
                  src_iz=[0] * ns, src_ix=list(np.linspace(10, nx - 11, ns).astype(int)),
                  rec_iz=[0] * nr, rec_ix=list(np.linspace(0, nx - 1, nr).astype(int)),
                  nt_f=nt, tap=tap, pad=pad, nref=2, eps=0.04, prop="pspi", ntaylor=1, ngs=2, ngr=1, ng=2)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 1500 * z) * (1 + 0.04 * rng.standard_normal((nz, nx)))
    slow = (1 / v ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    return sv, slow, den


@pytest.mark.parametrize("nx,pad,tap", [(80, 10, 0), (270, 15, 0), (500, 25, 25)])
def test_born_any_length_matches_oracle_and_dot(nx, pad, tap):
    """nk = 100, 300 and 600 with edge padding (and FFT padding for 600)."""
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey, born_adjoint, born_forward, wem_forward
    sv, slow, den = _born_case(nx, pad, tap)
    sp, dp = O.edge_pad(slow, sv.pad), O.edge_pad(den, sv.pad)
    m = Model(sp.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
              eps=sv.eps, ntaylor=sv.ntaylor, den=dp.astype(np.complex64))
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
               ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    print(f"born nk={nx + 2 * pad + 2 * tap}", flush=True)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    wem_forward(m, s, d)
    want = O.wem_forward(sv, slow, den)
    assert rel(d.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL
    rng = np.random.default_rng(6)
    dS = (rng.standard_normal((sv.nz, sv.nx)) * 1e-8).astype(np.complex64)
    dR = (rng.standard_normal((sv.nz, sv.nx)) * 0.1).astype(np.complex64)
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(O.zero_pad(dS, sv.pad)), _c(O.zero_pad(dR, sv.pad)), d)
    want = O.born_forward(sv, slow, den, dS, dR)
    assert rel(d.cpu().numpy()[:, :, sv.index], want[:, :, sv.index]) < TOL
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, len(sv.index))) + 1j * rng.standard_normal(
        (sv.ns, sv.nr, len(sv.index)))
    nxp = sv.nx + 2 * sv.pad
    gS = torch.zeros((sv.nz, nxp), dtype=torch.complex64, device="cuda")
    gR = torch.zeros_like(gS)
    born_adjoint(m, s, gS, gR, _c(data), exact=True)
    wS, wR = O.born_adjoint(sv, slow, den, data, mode="exact")
    sl = slice(sv.pad, sv.pad + sv.nx)
    assert rel(gS.cpu().numpy()[:, sl], wS) < TOL
    assert rel(gR.cpu().numpy()[:, sl], wR) < TOL
    # exact dot test on the padded model
    dSp = np.zeros((sv.nz, nxp), complex)
    dRp = np.zeros((sv.nz, nxp), complex)
    dSp[:, sl] = rng.standard_normal((sv.nz, sv.nx)) + 1j * rng.standard_normal((sv.nz, sv.nx))
    dRp[:, sl] = rng.standard_normal((sv.nz, sv.nx)) + 1j * rng.standard_normal((sv.nz, sv.nx))
    d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
    born_forward(m, s, _c(dSp), _c(dRp), d)
    lhs = O.cdot(d.cpu().numpy().astype(complex), data)
    rhs = O.cdot(dSp[:, sl], gS.cpu().numpy()[:, sl].astype(complex)) + O.cdot(
        dRp[:, sl], gR.cpu().numpy()[:, sl].astype(complex))
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)


def test_lengths_without_a_plan_are_rejected():
    import torch
    from wemfwi import Model
    from wemfwi._lib import WemfwiError
    for nx in (127, 134, 2049):  # a prime above 61, 2 x 67, longer than 2048 lanes
        m = Model(_slow(4, nx, 1), 10.0, 10.0)
        W = torch.zeros((1, 4, nx), dtype=torch.complex64, device="cuda")
        with pytest.raises(WemfwiError) as e:
            m.oneway(False, False, np.array([5.0], np.float32), W)
        assert e.value.code == 3


@pytest.fixture
def gen_env():
    import os
    keys = ("WEMFWI_GEN_FIXED", "WEMFWI_SWEEP_TILE")
    old = {k: os.environ.get(k) for k in keys}
    yield os.environ
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


@pytest.mark.parametrize("nx,tap,nref,prop,tile", [
    (500, 0, 1, "ssf", "full"),      # nk = 500: compile-time chain 4.5.5.5 in 512-lane rows
    (490, 5, 3, "pspi", "full"),     # nk = 500 with FFT padding
    (600, 0, 2, "pspi", "full"),     # nk = 600: compile-time chain 8.3.5.5 in 1024-lane rows
    (600, 0, 1, "ssf", "half"),      # the same in the half workgroup tile
    (100, 0, 1, "ssf", "full"),      # 4.5.5 in 128-lane rows
    (190, 5, 2, "pspi", "full"),     # nk = 200: 4.2.5.5 in 256-lane rows
    (300, 0, 1, "ssf", "full"),      # 4.3.5.5 in 512-lane rows
    (750, 0, 2, "pspi", "half"),     # 2.3.5.5.5 in 1024-lane rows
    (1000, 0, 1, "ssf", "full"),     # 8.5.5.5
    (1500, 0, 2, "pspi", "full"),    # 4.3.5.5.5 in 2048-lane rows
])
def test_fixed_chain_is_bitwise_the_runtime_chain(nx, tap, nref, prop, tile, gen_env):
    """The compile-time pass chains do the runtime chain's butterflies in the
    same order: bit-identical sweeps in every direction, and oracle parity."""
    nz = 12
    m, s, p, ref = _setup(nx, tap, 0, nref, prop, nz, seed=4)
    rng = np.random.default_rng(5)
    nrows = 9
    W = np.zeros((nrows, nz, nx), dtype=np.complex64)
    W[:, 0, nx // 3: nx // 3 + 6] = rng.standard_normal((nrows, 6)) + 1j * rng.standard_normal((nrows, 6))
    W[:, nz - 1, nx // 2: nx // 2 + 6] = rng.standard_normal((nrows, 6)) + 1j * rng.standard_normal((nrows, 6))
    freqs = np.linspace(3.0, 27.0, nrows).astype(np.float32)
    gen_env["WEMFWI_SWEEP_TILE"] = tile
    for up, adjoint in ((False, False), (False, True), (True, False), (True, True)):
        got = {}
        for fixed in ("1", "0"):
            gen_env["WEMFWI_GEN_FIXED"] = fixed
            print(f"nk={nx + 2 * tap} {prop} tile={tile} fixed={fixed} up={up} adjoint={adjoint}", flush=True)
            t = _c(W)
            m.oneway(up, adjoint, freqs, t)
            got[fixed] = t.cpu().numpy()
        assert np.array_equal(got["1"].view(np.uint32), got["0"].view(np.uint32)), (up, adjoint)
        fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
              (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adjoint)]
        for i in (0, nrows - 1):
            want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
            assert rel(got["1"][i], want) < TOL, (up, adjoint, i, rel(got["1"][i], want))


def test_born_fixed_chain_is_bitwise_the_runtime_chain(gen_env):
    """nk = 600 (a compile-time chain): WEM, Born forward and exact adjoint
    bit-identical with the runtime chain (sweeps and LinDown/LinUp)."""
    import torch
    from wemfwi import Model
    from wemfwi.ops import Survey, born_adjoint, born_forward
    sv, slow, den = _born_case(500, 25, 25)
    sp, dp = O.edge_pad(slow, sv.pad), O.edge_pad(den, sv.pad)
    m = Model(sp.astype(np.complex64), sv.dx, sv.dz, tap=sv.tap, pad=sv.pad, nref=sv.nref, prop=sv.prop,
              eps=sv.eps, ntaylor=sv.ntaylor, den=dp.astype(np.complex64))
    s = Survey(sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix, sv.freq, sv.index, sv.wave_f[sv.index], sv.nt_f,
               ngs=sv.ngs, ngr=sv.ngr, offset=sv.pad)
    rng = np.random.default_rng(8)
    nxp = sv.nx + 2 * sv.pad
    dS = _c(rng.standard_normal((sv.nz, nxp)) * 1e-8)
    dR = _c(rng.standard_normal((sv.nz, nxp)) * 0.1)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), complex)
    data[:, :, sv.index] = rng.standard_normal((sv.ns, sv.nr, len(sv.index)))
    got = {}
    for fixed in ("1", "0"):
        gen_env["WEMFWI_GEN_FIXED"] = fixed
        print(f"born nk=600 fixed={fixed}", flush=True)
        d = torch.zeros(s.data_shape, dtype=torch.complex64, device="cuda")
        born_forward(m, s, dS, dR, d)
        gS = torch.zeros((sv.nz, nxp), dtype=torch.complex64, device="cuda")
        gR = torch.zeros_like(gS)
        born_adjoint(m, s, gS, gR, _c(data), exact=True)
        got[fixed] = [a.cpu().numpy() for a in (d, gS, gR)]
    for a, b in zip(got["1"], got["0"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.mark.parametrize("nx,tap,nref,prop", [(500, 0, 2, "pspi"), (580, 10, 1, "ssf"), (1000, 0, 1, "ssf"),
                                                (1500, 0, 2, "pspi")])
def test_unit_ops_fixed_chain_is_bitwise_the_runtime_chain(nx, tap, nref, prop, gen_env):
    """Single depth steps and the scatterer at compile-time chain lengths:
    compile-time and runtime chains bit-identical (their oracle parity is
    checked above)."""
    import torch
    from wemfwi.api import Hypercube, Vector
    from wemfwi.unit import Scatter
    nz = 10
    m, s, p, ref = _setup(nx, tap, 0, nref, prop, nz)
    rng = np.random.default_rng(2)
    x = (rng.standard_normal((4, nx)) + 1j * rng.standard_normal((4, nx))).astype(np.complex64)
    freqs = np.array([2.0, 7.5, 17.25, 30.0], dtype=np.float32)
    depth = np.array([0, 3, 6, 9], dtype=np.int32)
    par = {"tap": tap, "pad": 0, "nref": nref, "prop": prop, "eps": 0.04, "ntaylor": 2}
    vs = Vector(Hypercube([nx, nz], [0.0, 0.0], [10.0, 10.0]), arr=_c(s))
    got = {}
    for fixed in ("1", "0"):
        gen_env["WEMFWI_GEN_FIXED"] = fixed
        print(f"unit ops nk={nx + 2 * tap} fixed={fixed}", flush=True)
        res = []
        for adjoint in (False, True):
            out = torch.zeros((4, nx), dtype=torch.complex64, device="cuda")
            m.onestep(adjoint, freqs, depth, _c(x), out)
            res.append(out.cpu().numpy())
        op = Scatter(vs, 2, par, adjoint_mode="exact")
        op.setFreq(9.0)
        op.setDepth(4)
        y = torch.zeros(nx, dtype=torch.complex64, device="cuda")
        op.forward(_c(x[0]), y)
        g = torch.zeros_like(y)
        op.adjoint(g, _c(x[1]))
        res += [y.cpu().numpy(), g.cpu().numpy()]
        got[fixed] = res
    for a, b in zip(got["1"], got["0"]):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
