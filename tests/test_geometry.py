"""Acquisition geometry on the CPU: irregular JSON devices (Device.h:65-85)
and the per-shot conformal maps (PropParam.h:89-131, conformal/src).

The oracle (oracle/cmap_oracle.py) is pinned first by what the mathematics
fixes -- identities, closed-form axes and Jacobians, and containment of the
exact mapped point in the cell a table names -- then the library's host
tables (wemfwi_cmap_plan) are compared with it: index tables and devices
exactly (a handful of last-bit libm ties allowed, counted), axes and
Jacobians to 1e-6 relative."""

import json
import math

import numpy as np
import pytest

from oracle import cmap_oracle as CO
from oracle import wem_oracle as O

NZ, NX, DX, DZ = 40, 64, 10.0, 10.0


def _plan(*a, **k):
    from wemfwi.params import cmap_plan
    return cmap_plan(*a, **k)


# --------------------------------------------------------------- JSON devices

def test_json_geometry_matches_oracle_and_regular_layout(tmp_path):
    from wemfwi.params import device_indices
    rng = np.random.default_rng(0)
    sx = (rng.random(7) * 600).tolist()
    sz = (rng.random(7) * 90).tolist()
    rx = [float(x) for x in np.arange(30) * 21.3 + 4.9]
    rz = [12.5] * 30
    f = tmp_path / "geom.json"
    f.write_text(json.dumps({"sx": sx, "sz": sz, "rx": rx, "rz": rz}))
    par = {"coord": str(f)}
    for name, (x, z) in (("s", (sx, sz)), ("r", (rx, rz))):
        iz, ix = device_indices(par, name, 2.0, -1.0, DX, DZ)
        wz, wx = CO.device_coords(x, z, 2.0, -1.0, DX, DZ)
        assert np.array_equal(iz, wz) and np.array_equal(ix, wx)
    # a regular layout written as coordinates gives the regular indices
    reg = {"nr": 30, "orx": 4.9, "drx": 21.3, "orz": 12.5, "drz": 0.0}
    iz0, ix0 = device_indices(reg, "r", 2.0, -1.0, DX, DZ)
    xs = [np.float32(np.float32(4.9) + np.float32(i) * np.float32(21.3)) for i in range(30)]
    f2 = tmp_path / "reg.json"
    f2.write_text(json.dumps({"rx": [float(v) for v in xs], "rz": [12.5] * 30}))
    iz1, ix1 = device_indices({"coord": str(f2)}, "r", 2.0, -1.0, DX, DZ)
    assert np.array_equal(iz0, iz1) and np.array_equal(ix0, ix1)
    # closed form of the rounding: nearest cell
    assert np.array_equal(ix1, np.floor((np.asarray(xs, np.float64) - 2.0) / DX + 0.5).astype(int))


def test_json_geometry_errors(tmp_path):
    from wemfwi._lib import WemfwiError
    from wemfwi.params import device_indices
    f = tmp_path / "bad.json"
    f.write_text(json.dumps({"sx": [10.0, -50.0], "sz": [0.0, 0.0]}))
    with pytest.raises(WemfwiError):
        device_indices({"coord": str(f)}, "s", 0.0, 0.0, DX, DZ)
    f.write_text(json.dumps({"sx": [10.0], "sz": [0.0, 0.0]}))
    with pytest.raises(ValueError):
        device_indices({"coord": str(f)}, "s", 0.0, 0.0, DX, DZ)


# ------------------------------------------------------------- oracle pins

SRC_IZ, SRC_IX = [0, 3, 10], [5, 30, 60]
REC_IZ, REC_IX = [0] * NX, list(range(NX))


def test_oracle_identity_cases():
    n = NZ * NX
    t = CO.cmap_tables(0, NZ, NX, 0.0, 0.0, DZ, DX, SRC_IZ, SRC_IX, REC_IZ, REC_IX)
    assert np.array_equal(t.fwd, np.tile(np.arange(n), (3, 1))) and np.array_equal(t.inv, t.fwd)
    assert np.all(t.jac == 1)
    # rotation by 0 about a grid with its origin at 0: the identity
    t = CO.cmap_tables(2, NZ, NX, 0.0, 0.0, DZ, DX, SRC_IZ, SRC_IX, REC_IZ, REC_IX, angles=[0.0] * 3)
    assert np.array_equal(t.fwd[0], np.arange(n)) and np.array_equal(t.inv[0], np.arange(n))
    assert np.allclose(t.axes[0], [0.0, DX, 0.0, DZ])
    assert np.array_equal(t.rec_iz[0], REC_IZ) and np.array_equal(t.rec_ix[0], REC_IX)
    # ShiftAndScale with a = 1, b = 0 on integer coordinates: the identity
    g = CO.Grid(NX, NZ, np.float32(0), np.float32(DX), np.float32(0), np.float32(DZ))
    m = CO.shift_scale(g, 1.0, 0.0)
    assert np.array_equal(CO.forward_gather(m), np.arange(n))
    assert np.array_equal(CO.inverse_gather(m), np.arange(n))


def _cell_slack(frac_exact, cell, slack_low=1):
    """cell is the floor of the exact coordinate, or up to slack_low below it
    (a first truncation at an exact integer, where rounding picks a side)."""
    f = np.floor(frac_exact + 1e-9)
    return (cell <= f) & (cell >= f - slack_low)


def test_oracle_log_map_closed_forms():
    eps = 0.01
    t = CO.cmap_tables(1, NZ, NX, 0.0, 0.0, DZ, DX, SRC_IZ, SRC_IX, REC_IZ, REC_IX, eps_log=eps)
    e = math.log(eps)
    for s in range(3):
        o1, d1, o2, d2 = t.axes[s]
        assert o1 == 0 and abs(d1 - 2 * math.pi / (NX - 1)) < 1e-6 and abs(o2 - e) < 1e-6
        assert abs(d2 + e / (NZ - 1)) < 1e-6
        # sources: the first row of the log grid
        assert np.array_equal(t.src[s][0], np.zeros(NX)) and np.array_equal(t.src[s][1], np.arange(NX))
    # scale of each shot: 1 / (farthest corner distance), with the previous
    # shot's (tiny) scale in the running maximum (PropParam.h:91,96-103)
    zb = [complex(SRC_IZ[s] * DZ, SRC_IX[s] * DX) for s in range(3)]
    corners = [complex(z, x) for z in (0.0, (NZ - 1) * DZ) for x in (0.0, (NX - 1) * DX)]
    a = 0.0
    for s in range(3):
        a = 1.0 / max([a] + [abs(c - zb[s]) for c in corners])
        assert abs(t.jac[s, 0, 0] - 1.0 / a) < 1e-5 / a
        i2 = np.arange(NZ * NX) // NX
        assert np.allclose(t.jac[s, 1], np.exp(e - i2 * e / (NZ - 1)), rtol=1e-6)
        lg_grid = t.maps[s][1].out

        def log_cells(iz, ix):
            """Exact log-grid cells of model sample (iz, ix) through both
            truncations: the shift-and-scale cell is (iz, ix) exactly, which
            rounding may turn into the neighbour below in either axis."""
            cands = set()
            for dz_ in (0, 1):
                for dx_ in (0, 1):
                    w = a * (complex((iz - dz_) * DZ, (ix - dx_) * DX) - zb[s])
                    # inside the small circle: radius e^eps, angle kept (ComplexLog.h:51-53)
                    u = complex(math.log(eps), math.atan2(w.imag, w.real)) if abs(w) < eps else np.log(w)
                    ang = u.imag + 2 * math.pi if u.imag < 0 else u.imag
                    cz = math.floor((u.real - float(lg_grid.o2)) / float(lg_grid.d2) + 1e-9)
                    cx = math.floor(ang / float(lg_grid.d1) + 1e-9)
                    for kz in (cz, cz - 1):
                        for kx in (cx, cx - 1):
                            cands.add((max(0, kz), kx))
            return cands

        # receivers through both maps, radius clipped at 0
        for i in range(NX):
            c = log_cells(REC_IZ[i], REC_IX[i])
            assert not c or (int(t.rec_iz[s][i]), int(t.rec_ix[s][i])) in c, (s, i)
        # inverse gather of model samples away from the source
        v = t.inv[s]
        for i in np.nonzero(v >= 0)[0][::7]:
            c = log_cells(i // NX, i % NX)
            assert not c or (int(v[i] // NX), int(v[i] % NX)) in c, (s, i)
        assert (v >= 0).mean() > 0.5
        # forward gather: the model cell holds the exact preimage exp(u)/a + b
        # (one truncation at an exact integer may pick the cell below)
        u_out = lg_grid.coords().astype(np.complex128)
        zstar = np.exp(u_out) / a + zb[s]
        c = t.fwd[s]
        inside = ((zstar.real >= 0) & (zstar.real <= (NZ - 1) * DZ) & (zstar.imag >= 0) &
                  (zstar.imag <= (NX - 1) * DX))
        sel = inside & (c >= 0)
        assert sel.mean() > 0.3
        assert np.all(_cell_slack(zstar.real[sel] / DZ, c[sel] // NX))
        assert np.all(_cell_slack(zstar.imag[sel] / DX, c[sel] % NX))


def test_oracle_scale_carries_over_between_shots():
    """On a grid of centimetres the previous shot's scale exceeds every
    corner distance, so the next shot reuses 1/a (PropParam.h:91)."""
    d = 0.01
    t = CO.cmap_tables(1, 20, 20, 0.0, 0.0, d, d, [0, 10], [0, 10], [0] * 20, list(range(20)))
    far0 = math.hypot(19 * d, 19 * d)
    a0 = 1 / far0
    assert abs(t.jac[0, 0, 0] - far0) < 1e-6
    assert abs(t.jac[1, 0, 0] - max(a0, math.hypot(10 * d, 10 * d))) < 1e-4 * a0


def test_oracle_rotation_closed_forms():
    th = 0.35
    t = CO.cmap_tables(2, NZ, NX, 0.0, 0.0, DZ, DX, [5], [20], REC_IZ, REC_IX, angles=[th])
    rot = np.exp(-1j * th)
    corners = np.array([complex(z, x) for z in (0.0, (NZ - 1) * DZ) for x in (0.0, (NX - 1) * DX)]) * rot
    o1, o2 = min(0.0, corners.imag.min()), min(0.0, corners.real.min())
    d1 = (max(0.0, corners.imag.max()) - o1) / (NX - 1)
    d2 = (max(0.0, corners.real.max()) - o2) / (NZ - 1)
    assert np.allclose(t.axes[0], [o1, d1, o2, d2], rtol=1e-5, atol=1e-4)
    zr = np.array([complex(REC_IZ[i] * DZ, REC_IX[i] * DX) for i in range(NX)]) * rot
    fz, fx = (zr.real - o2) / d2, (zr.imag - o1) / d1
    tie = (np.abs(fz - np.round(fz)) < 1e-4) | (np.abs(fx - np.round(fx)) < 1e-4)
    assert np.all((t.rec_iz[0] == np.trunc(fz))[~tie]) and np.all((t.rec_ix[0] == np.trunc(fx))[~tie])
    zs = complex(5 * DZ, 20 * DX) * rot
    assert (t.src[0][0][0], t.src[0][1][0]) == (int((zs.real - o2) / d2), int((zs.imag - o1) / d1))
    # forward gather away from the clamped border: the cell of u e^{i th}
    g = t.maps[0][0].out
    zp = g.coords().astype(np.complex128) * np.exp(1j * th)
    cz, cx = zp.real / DZ, zp.imag / DX
    inner = (cz > 0.5) & (cz < NZ - 1.5) & (cx > 0.5) & (cx < NX - 1.5)
    tie = (np.abs(cz - np.round(cz)) < 1e-4) | (np.abs(cx - np.round(cx)) < 1e-4)
    sel = inner & ~tie
    assert sel.mean() > 0.2
    f = t.fwd[0]
    assert np.all(f[sel] // NX == np.floor(cz[sel])) and np.all(f[sel] % NX == np.floor(cx[sel]))


# ----------------------------------------------------- product vs oracle

@pytest.mark.parametrize("kind,geo", [
    (1, (40, 64, 0.0, 0.0, 10.0, 10.0)), (1, (33, 50, 5.0, -20.0, 7.5, 12.5)), (1, (300, 512, 0.0, 0.0, 10.0, 10.0)),
    (2, (40, 64, 0.0, 0.0, 10.0, 10.0)), (2, (33, 50, 5.0, -20.0, 7.5, 12.5)), (0, (20, 30, 0.0, 0.0, 5.0, 5.0)),
])
def test_cmap_plan_matches_oracle(kind, geo):
    nz, nx, oz, ox, dz, dx = geo
    siz, six = [0, 3, min(10, nz - 1)], [5, 30 % nx, 60 % nx]
    riz, rix = [0] * nx, list(range(nx))
    ang = [0.1, -0.3, 0.5]
    p = _plan(kind, nz, nx, oz, ox, dz, dx, siz, six, riz, rix, eps_log=0.01, angles=ang)
    t = CO.cmap_tables(kind, nz, nx, oz, ox, dz, dx, siz, six, riz, rix, eps_log=0.01, angles=ang)
    assert np.allclose(p.axes, t.axes, rtol=1e-6, atol=2e-6)
    assert np.allclose(p.jac, t.jac, rtol=1e-6)
    n = p.fwd.size
    # last-bit differences of numpy's complex exp/log against the C library
    # may flip a truncation at a cell boundary: at most 1e-4 of the entries
    assert (p.fwd != t.fwd).sum() <= max(2, n * 1e-4)
    assert (p.inv != t.inv).sum() <= max(2, n * 1e-4)
    assert (p.rec_iz != t.rec_iz).sum() + (p.rec_ix != t.rec_ix).sum() <= 1
    for (a, b), (c, d) in zip(p.src, t.src):
        assert np.array_equal(a, c) and np.array_equal(b, d)


def test_cmap_plan_rejects_bad_input():
    from wemfwi._lib import WemfwiError
    with pytest.raises(WemfwiError):
        _plan(3, NZ, NX, 0, 0, DZ, DX, [0], [5], [0, 0], [1, 2])
    with pytest.raises(WemfwiError):
        _plan(1, NZ, NX, 0, 0, DZ, DX, [0], [5], [0], [1])          # one receiver becomes a line
    with pytest.raises(WemfwiError):
        _plan(1, NZ, NX, 0, 0, DZ, DX, [NZ], [5], [0, 0], [1, 2])   # source outside the model
    with pytest.raises(ValueError):
        _plan(2, NZ, NX, 0, 0, DZ, DX, [0, 1], [5, 6], [0, 0], [1, 2], angles=[0.1])


# --------------------------------------------- mapped-grid oracle, identity map

def test_oracle_mapped_wem_with_identity_map_is_plain_wem():
    """With the identity map (unit Jacobian, identity gathers) the per-shot
    oracle must reproduce the plain WEM / Born oracle."""
    nz, nx = 8, 24
    rng = np.random.default_rng(3)
    slow = (1 / (1500 + 800 * rng.random((nz, nx))) ** 2).astype(np.complex64)
    den = (1 + 0.2 * rng.random((nz, nx))).astype(np.complex64)
    sv = O.Survey(nx=nx, nz=nz, dx=10.0, dz=10.0, freq=np.array([5.0, 7.0], np.float32), index=np.array([1, 2]),
                  wave_f=np.array([0, 1 + 0.5j, 0.7 - 0.2j, 0], complex), src_iz=[0, 1], src_ix=[4, 15],
                  rec_iz=[0] * 5, rec_ix=[2, 6, 10, 14, 20], nt_f=4, pad=4, nref=2, ngs=1, ngr=1)
    t = CO.cmap_tables(0, nz, nx, 0.0, 0.0, 10.0, 10.0, sv.src_iz, sv.src_ix, sv.rec_iz, sv.rec_ix)
    d0 = O.wem_forward(sv, slow, den)
    d1 = O.wem_forward_mapped(sv, slow, den, t.fwd, t.jac, t.axes)
    assert np.allclose(d0, d1, rtol=0, atol=1e-12 * np.abs(d0).max())
    dS = rng.standard_normal((nz, nx)) * 1e-9 + 0j
    dR = rng.standard_normal((nz, nx)) * 1e-2 + 0j
    b0 = O.born_forward(sv, slow, den, dS, dR)
    b1 = O.born_forward_mapped(sv, slow, den, dS, dR, t.fwd, t.jac, t.axes)
    assert np.allclose(b0, b1, rtol=0, atol=1e-12 * np.abs(b0).max())
    g0 = O.born_adjoint(sv, slow, den, b0)
    for how in ("inverse", "transpose"):
        g1 = O.born_adjoint_mapped(sv, slow, den, b0, t.fwd, t.inv, t.jac, t.axes, map_adjoint=how)
        for a, b in zip(g0, g1):
            assert np.allclose(a, b, rtol=0, atol=1e-12 * np.abs(a).max())


def _apply_csr(ptr, src, x):
    out = np.zeros(len(ptr) - 1, dtype=complex)
    for j in range(len(ptr) - 1):
        out[j] = x[src[ptr[j]:ptr[j + 1]]].sum()
    return out


@pytest.mark.parametrize("kind", [1, 2])
def test_map_gather_lists_match_oracle_maps(kind):
    """The device gather lists of the API (map_csr) against the oracle's
    per-shot map, zero pad, truncation, inverse map and transpose."""
    from wemfwi.api import map_csr
    nz, nx, pad = 12, 30, 3
    p = _plan(kind, nz, nx, 0.0, 0.0, 10.0, 10.0, [0, 4], [5, 20], [0] * 8, list(range(2, 26, 3)),
              eps_log=0.02, angles=[0.25, -0.1])
    rng = np.random.default_rng(1)
    m = rng.standard_normal((nz, nx)) + 1j * rng.standard_normal((nz, nx))
    fwd = _apply_csr(*map_csr(p, nz, nx, pad, "fwd"), m.ravel()).reshape(2, nz, nx + 2 * pad)
    for k in range(2):
        assert np.allclose(fwd[k], O._gather_pad(m, p.fwd[k], pad))
    g = rng.standard_normal(fwd.shape) + 1j * rng.standard_normal(fwd.shape)
    sl = slice(pad, pad + nx)
    want_inv = sum(np.where(p.inv[k] >= 0, g[k][:, sl].ravel()[np.maximum(p.inv[k], 0)], 0) for k in range(2))
    assert np.allclose(_apply_csr(*map_csr(p, nz, nx, pad, "inv"), g.ravel()), want_inv)
    # exact: <fwd m, g> = <m, exact g>
    ex = _apply_csr(*map_csr(p, nz, nx, pad, "exact"), g.ravel())
    assert abs(np.vdot(g.ravel(), fwd.ravel()) - np.vdot(ex, m.ravel())) < 1e-10 * np.abs(g).sum()
