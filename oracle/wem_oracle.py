"""FP64 CPU oracle of the one-way wave-equation (WEM) propagator hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product path never
routes through it.

It restates, in double precision and in plain NumPy, the source text of the
reference's propagator (paths are relative to the reference repository):

  ===========================  ===============================================
  function here                reference it follows
  ===========================  ===============================================
  kgrid                        propagator/src/Phshift_cmplx.cpp:12-24,
                               propagator/src/Scatter.cpp:24-31
  fft_u / ifft_u               operator/src/FFT1_c.cpp:73-103 (origin-shift
                               phase omitted: it cancels around every
                               k-diagonal operator, FFT1_c.cpp:82-83,96-97)
  phase_factor                 propagator/src/phase_shift.ispc:2-32
  taper_filter                 operator/src/Taper.h:11-23
  cos2_window                  propagator/src/Injection.h:53-65
  refsampler                   propagator/src/RefSampler.cpp:10-91
  onestep_fwd / onestep_adj    propagator/src/SSF.cpp:8-60, PSPI.cpp:8-53,
                               SplitStep.cpp:5-29, Selector.cpp:5-19,
                               OneStep.h:19-38
  down_* / up_*                propagator/src/OneWay.cpp:32-109
  scatter_fwd / scatter_adj    propagator/src/Scatter.cpp:35-111
  lindown_* / linup_*          propagator/src/LinOneWay.cpp:33-142
  ic_fwd / ic_adj              propagator/src/ic.cpp:5-45
  reflect_coef, dreflect_*     propagator/src/Reflect.h:12-48, dReflect.h:11-84
  inj_*                        propagator/src/Injection.cpp:8-80
  freq_grid, wavelet_spectrum  propagator/src/PropParam.h:45-85,
                               operator/src/FFT1_r.cpp:7-83
  device_indices               propagator/src/Device.h:39-63
  wem_forward                  propagator/src/wem_prop.cpp:9-78
  born_forward / born_adjoint  propagator/src/born_prop.cpp:11-235,
                               Born_refl.cpp:19-40
  bornrefl_forward / _adjoint  propagator/src/ext_born_refl.cpp:8-110
  extwem_forward               propagator/src/ext_wem_prop.cpp:8-59
  extborn_forward / _adjoint   propagator/src/ext_born_prop.cpp:8-194
  offset_ic_fwd / _adj         no reference counterpart (subsurface-offset
                               imaging condition, extends ic.cpp)
  offset_forward,              BornRefl's fwd/adj propagation
  offset_objective             (ext_born_refl.cpp:8-110) with the offset
                               imaging condition in place of ic.cpp: the
                               extended least-squares objective and gradient
  mapped_models,               propagator/src/wem_prop.cpp:20-45 and
  wem_forward_mapped,          born_prop.cpp:28-70,136-174,217-231 (conformal
  born_*_mapped                maps; tables from oracle/cmap_oracle.py)
  ===========================  ===============================================

Quirks of the reference are reproduced, not fixed: the k grid with
dk = 2*pi/(dx*(nk-1)) and the never-written k[nk/2]; the split-step phase
w*dz*(S-Sref)/sqrt(Sref); the complex damping S - i*eps*Re(S) of the phase
shift against the real shift S - eps*Re(S) of the scatterer; the stale
reference buffer of RefSampler; the operator order of Scatter::adjoint
("compat" adjoint mode); the dReflect coefficient products as written;
assignment (not addition) in Selector and Reflect.

Two deliberate restatements are mathematically identical to the reference:
  * Im(kz) is evaluated as -|b|/(2 Re kz) (or Re(kz) = |b|/(2|Im kz|) for
    a < 0) instead of -sqrt((c-a)/2); (c-a)(c+a) = b^2 makes both equal, and
    the product form avoids cancellation.
  * The Born adjoint sums every (shot, frequency) contribution in a fixed
    order instead of the unsynchronised accumulation of born_prop.cpp:217-232.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

C128 = np.complex128
PI_F32 = np.float32(4.0 * math.atan(1.0))

# Taylor coefficients of (1 - x)^(-1/2), Scatter.h:36
SCATTER_COEF = (1.0, 1.0 / 2.0, 3.0 / 8.0, 5.0 / 16.0, 35.0 / 128.0)


# --------------------------------------------------------------------------
# parameters
# --------------------------------------------------------------------------

@dataclass
class StepParams:
    """Configuration of one depth step (OneStep.h:19-38, OneWay.cpp:22-28).

    nx   : propagation width (model nx + 2*pad)
    dx,dz: grid spacing
    tap  : FFT zero padding per side (par "tap")
    pad  : taper width (par "pad")
    nref : number of reference slownesses (par "nref")
    eps  : damping (par "eps")
    prop : "ssf" or "pspi" (par "prop")
    ntaylor: Taylor order of the scatterer (par "ntaylor")
    """

    nx: int
    dx: float
    dz: float
    tap: int = 0
    pad: int = 0
    nref: int = 1
    eps: float = 0.04
    prop: str = "ssf"
    ntaylor: int = 1

    @property
    def nk(self) -> int:
        return self.nx + 2 * self.tap


@dataclass
class RefTables:
    """Output of RefSampler: per depth the reference slownesses and, per x,
    which reference selects it (-1: none)."""

    sref: np.ndarray          # complex64 [nz][nref]
    refid: np.ndarray         # int32 [nz][nx]
    overruns: int = 0         # depths where an unbounded while loop ran to nx


# --------------------------------------------------------------------------
# FFT, grids, phase shift
# --------------------------------------------------------------------------

def kgrid(nk: int, d: float) -> np.ndarray:
    """Wavenumber grid of the phase shift and the scatterer.

    Phshift_cmplx.cpp:12-24 (and Scatter.cpp:24-31): dk = 2*pi/(d*(nk-1));
    k[ik] = ik*dk for ik < nk/2; k[nk-ik] = -k[ik] for 1 <= ik < nk/2; the
    remaining entries (index nk/2, and nk/2+1 for odd nk) keep their zero
    initialisation.
    """
    dk = 2.0 * math.pi / (d * (nk - 1))
    k = np.zeros(nk, dtype=np.float64)
    h = nk // 2
    k[:h] = np.arange(h) * dk
    for ik in range(1, h):
        k[nk - ik] = -k[ik]
    return k


def fft_u(x: np.ndarray) -> np.ndarray:
    """Unitary forward FFT along the last axis (FFTW_FORWARD / sqrt(n))."""
    n = x.shape[-1]
    return np.fft.fft(x, axis=-1) / math.sqrt(n)


def ifft_u(x: np.ndarray) -> np.ndarray:
    """Unitary inverse FFT along the last axis (FFTW_BACKWARD / sqrt(n))."""
    n = x.shape[-1]
    return np.fft.ifft(x, axis=-1) * math.sqrt(n)


def kz_parts(w, s_ref: complex, eps: float, k: np.ndarray):
    """Return (re, im) of kz with kz = re + i*im, im <= 0.

    phase_shift.ispc:8-12: a = w^2*Re(S) - k^2, b = w^2*(Im(S) - eps*Re(S)),
    c = sqrt(a^2 + b^2), re = sqrt((c+a)/2), im = -sqrt((c-a)/2).
    """
    w = np.asarray(w, dtype=np.float64)[..., None]
    a = w * w * s_ref.real - k * k
    b = w * w * (s_ref.imag - eps * s_ref.real)
    c = np.hypot(a, b)
    big = np.sqrt((c + np.abs(a)) * 0.5)
    with np.errstate(divide="ignore", invalid="ignore"):
        small = np.where(big > 0, np.abs(b) / (2.0 * np.where(big > 0, big, 1.0)), 0.0)
    re = np.where(a >= 0, big, small)
    im = -np.where(a >= 0, small, big)
    return re, im


def phase_factor(w, dz: float, s_ref: complex, eps: float, k: np.ndarray) -> np.ndarray:
    """Forward phase-shift multiplier exp(-i*re*dz)*exp(im*dz) (ps_fwd,
    phase_shift.ispc:13-14).  The adjoint multiplier is its conjugate
    (ps_adj, phase_shift.ispc:29-30)."""
    re, im = kz_parts(w, complex(s_ref), eps, k)
    return np.exp(im * dz) * np.exp(-1j * re * dz)


def taper_filter(n: int, tap: int) -> np.ndarray:
    """Taper.h:11-23: f[i] = 0.5*(1 - cos(pi*i/(tap-1))) for i < tap,
    mirrored at the far end; tap == 0 is no taper.  tap == 1 divides by zero
    in the reference and is rejected here."""
    if tap == 1:
        raise ValueError("taper width pad=1 is undefined (0/0)")
    f = np.ones(n, dtype=np.float64)
    for i in range(tap):
        f[i] = 0.5 * (1.0 - math.cos(math.pi * i / (tap - 1)))
        f[n - 1 - i] = f[i]
    return f


def cos2_window(ng: int) -> np.ndarray:
    """Injection.h:53-65: cosine-squared spread of half-width ng."""
    if ng == 0:
        return np.ones(1)
    g = np.zeros(2 * ng + 1)
    size = 2 * ng + 1
    for i in range(size // 2):
        g[i] = (-math.cos(math.pi / ng * i) + 1.0) / 2.0
    for i in range(size // 2, size):
        g[i] = (math.cos(math.pi / ng * (i - size + ng + 1)) + 1.0) / 2.0
    return g


# --------------------------------------------------------------------------
# RefSampler (RefSampler.cpp:10-91), restated bit-faithfully in float32
# --------------------------------------------------------------------------

def refsampler(slow: np.ndarray, nref: int) -> RefTables:
    """Reference slownesses and nearest-reference partition per depth.

    uniform_sample (RefSampler.cpp:28-39): sref[iz][r] = S[iz][o + r*nx/nref]
    with o = nx/(2*nref); non-positive Re(sref) raises.
    getIndexes (RefSampler.cpp:41-84): sorts the nx real slownesses, sorts
    the *whole* nx-long reference buffer of which only nref entries are
    refreshed per depth (the rest are zeros from construction or values of
    earlier depths), and walks the sorted slownesses with midpoint tests.
    The while loops have no bound in the reference; here they stop at nx and
    the overrun is counted.
    """
    slow = np.asarray(slow)
    nz, nx = slow.shape
    if nref < 1 or nx < 2:
        raise ValueError("refsampler needs nref >= 1 and nx >= 2")
    sref = np.zeros((nz, nref), dtype=np.complex64)
    o = nx // (2 * nref)
    for iz in range(nz):
        for r in range(nref):
            off = o + r * nx // nref
            sref[iz, r] = np.complex64(slow[iz, off])
            if np.float32(sref[iz, r].real) <= 0:
                raise ValueError(
                    f"Ref velocity {float(sref[iz, r].real)} at the layer {iz}")

    refid = np.full((nz, nx), -1, dtype=np.int32)
    ref_buf = np.zeros(nx, dtype=np.float32)   # _slow_ref_slice, zero at construction
    overruns = 0
    for iz in range(nz):
        s = np.ascontiguousarray(slow[iz].real.astype(np.float32))
        for r in range(nref):
            ref_buf[r] = np.float32(sref[iz, r].real)
        is_sort = np.argsort(s, kind="stable")
        isref_sort = np.argsort(ref_buf[:nref], kind="stable")
        ss = np.sort(s)
        ref_buf.sort()                    # sorts all nx entries (stale buffer)
        R = ref_buf
        lists = {}
        off = 0
        while off < nx and (ss[off] - R[0]) < (R[1] - ss[off]):
            off += 1
        if off >= nx:
            overruns += 1
        lists[int(isref_sort[0])] = is_sort[0:off]
        off_beg = off
        for ir in range(1, nref - 1):
            while off < nx and (ss[off] - R[ir]) < (R[ir + 1] - ss[off]):
                off += 1
            if off >= nx:
                overruns += 1
            lists[int(isref_sort[ir])] = is_sort[off_beg:off]
            off_beg = off
        lists[int(isref_sort[nref - 1])] = is_sort[off:]
        for r, lst in lists.items():
            refid[iz, lst] = r
    return RefTables(sref=sref, refid=refid, overruns=overruns)


# --------------------------------------------------------------------------
# one depth step (SSF / PSPI)
# --------------------------------------------------------------------------

def _pad(x: np.ndarray, tap: int) -> np.ndarray:
    if tap == 0:
        return x.astype(C128, copy=True)
    z = np.zeros(x.shape[:-1] + (x.shape[-1] + 2 * tap,), dtype=C128)
    z[..., tap:tap + x.shape[-1]] = x
    return z


def _unpad(x: np.ndarray, tap: int, nx: int) -> np.ndarray:
    return x[..., tap:tap + nx]


def _ss_phase(w, dz: float, s_row: np.ndarray, s_ref: complex) -> np.ndarray:
    """SplitStep.cpp:13: exponent -i*w*dz*(S - Sref)/sqrt(Sref)."""
    w = np.asarray(w, dtype=np.float64)[..., None]
    return -1j * w * dz * (s_row.astype(C128) - s_ref) / np.sqrt(complex(s_ref))


def onestep_fwd(x: np.ndarray, iz: int, w, slow: np.ndarray, ref: RefTables,
                p: StepParams) -> np.ndarray:
    """SSF::forward (SSF.cpp:8-32) / PSPI::forward (PSPI.cpp:8-29).

    x: [..., nx] wavefield row(s); w: angular frequency (scalar or per row).
    out[x] = taper[x] * ss_r(x) * (Pad^H F^H P_r F Pad x)[x] for the
    reference r selecting x (Selector assigns; unselected x stay 0)."""
    nx = p.nx
    k = kgrid(p.nk, p.dx)
    taper = taper_filter(nx, p.pad)
    X = fft_u(_pad(x, p.tap))
    out = np.zeros(x.shape, dtype=C128)
    for r in range(p.nref):
        sel = ref.refid[iz] == r
        if not sel.any():
            continue
        sr = complex(ref.sref[iz, r])
        y = _unpad(ifft_u(X * phase_factor(w, p.dz, sr, p.eps, k)), p.tap, nx) * taper
        if p.prop != "pspi":
            y = y * np.exp(_ss_phase(w, p.dz, slow[iz], sr))
        out[..., sel] = y[..., sel]
    return out


def onestep_adj(d: np.ndarray, iz: int, w, slow: np.ndarray, ref: RefTables,
                p: StepParams) -> np.ndarray:
    """SSF::adjoint (SSF.cpp:34-59) / PSPI::adjoint (PSPI.cpp:31-53)."""
    nx = p.nx
    k = kgrid(p.nk, p.dx)
    taper = taper_filter(nx, p.pad)
    acc = np.zeros(d.shape[:-1] + (p.nk,), dtype=C128)
    for r in range(p.nref):
        sel = ref.refid[iz] == r
        if not sel.any():
            continue
        sr = complex(ref.sref[iz, r])
        z = np.where(sel, d, 0).astype(C128)
        if p.prop != "pspi":
            # SplitStep.cpp:26: exp(+i*w*dz*conj((S-Sref)/sqrt(Sref)))
            z = z * np.exp(np.conj(_ss_phase(w, p.dz, slow[iz], sr)))
        z = z * taper
        acc += fft_u(_pad(z, p.tap)) * np.conj(phase_factor(w, p.dz, sr, p.eps, k))
    return _unpad(ifft_u(acc), p.tap, nx)


# --------------------------------------------------------------------------
# Down / Up (OneWay.cpp:32-109)
# --------------------------------------------------------------------------

def _oneway(model, w, slow, ref, p, down: bool, adjoint: bool, inplace: bool,
            src_add: Optional[np.ndarray] = None):
    """Generic Down/Up recursion over the depth axis (axis -2).

    In place (the only way the reference's orchestrators call them):
      Down fwd: W[iz+1] = P_iz W[iz] + W[iz+1],           iz = 0..nz-2
      Down adj: W[iz-1] = P^H_{iz-1} W[iz] + W[iz-1],     iz = nz-1..1
      Up fwd:   W[iz-1] = P_iz W[iz] + W[iz-1],           iz = nz-1..1
      Up adj:   W[iz+1] = P^H_{iz+1} W[iz] + W[iz+1],     iz = 0..nz-2
    Out of place the same formulas read the *input* rows only (I + L).
    """
    m = np.array(model, dtype=C128, copy=True)
    nz = m.shape[-2]
    out = m if inplace else np.array(m, copy=True)
    fwd_dir = down != adjoint           # True: iz increases
    order = range(0, nz - 1) if fwd_dir else range(nz - 1, 0, -1)
    for iz in order:
        nxt = iz + 1 if fwd_dir else iz - 1
        src = out[..., iz, :] if inplace else m[..., iz, :]
        if adjoint:
            y = onestep_adj(src, nxt, w, slow, ref, p)
        else:
            y = onestep_fwd(src, iz, w, slow, ref, p)
        out[..., nxt, :] = y + m[..., nxt, :]
    return out


def down_fwd(m, w, slow, ref, p, inplace=True):
    return _oneway(m, w, slow, ref, p, down=True, adjoint=False, inplace=inplace)


def down_adj(d, w, slow, ref, p, inplace=True):
    return _oneway(d, w, slow, ref, p, down=True, adjoint=True, inplace=inplace)


def up_fwd(m, w, slow, ref, p, inplace=True):
    return _oneway(m, w, slow, ref, p, down=False, adjoint=False, inplace=inplace)


def up_adj(d, w, slow, ref, p, inplace=True):
    return _oneway(d, w, slow, ref, p, down=False, adjoint=True, inplace=inplace)


# --------------------------------------------------------------------------
# Scatter (Scatter.cpp) and LinDown / LinUp (LinOneWay.cpp)
# --------------------------------------------------------------------------

def _scatter_coef(slow_row: np.ndarray, eps: float, it: int) -> np.ndarray:
    """Scatter.cpp:82: coef[it] * (S - eps*Re(S))^(-1/2-it)."""
    s = slow_row.astype(C128)
    base = s - eps * s.real
    return SCATTER_COEF[it] * np.power(base, -0.5 - it)


def _scatter_kernel(x, w, k, tap, nx, it):
    """Pad^H F^H (k/w)^(2 it) F Pad x (Scatter.cpp:39-45,68-75)."""
    w = np.asarray(w, dtype=np.float64)[..., None]
    K = (k / w) ** (2 * it)
    return _unpad(ifft_u(fft_u(_pad(x, tap)) * K), tap, nx)


def scatter_fwd(x, iz, w, slow, p: StepParams):
    """Scatter::forward (Scatter.cpp:35-49):
    y = sum_it (-i*w*dz) * C_it(x) * (Pad^H F^H K_it F Pad x)."""
    k = kgrid(p.nk, p.dx)
    wv = np.asarray(w, dtype=np.float64)[..., None]
    y = np.zeros(np.shape(x), dtype=C128)
    for it in range(p.ntaylor + 1):
        c = _scatter_coef(slow[iz], p.eps, it)
        y = y + (-1j * wv * p.dz) * c * _scatter_kernel(x, w, k, p.tap, p.nx, it)
    return y


def scatter_adj(d, iz, w, slow, p: StepParams, mode: str = "compat"):
    """Scatter::adjoint (Scatter.cpp:52-66).

    compat: sum_it (+i*w*dz) * conj(C_it) * (M_it d)  -- the reference order
    exact:  sum_it (+i*w*dz) * M_it (conj(C_it) * d)  -- the true adjoint
    (they coincide for ntaylor = 0 or laterally constant slowness)."""
    k = kgrid(p.nk, p.dx)
    wv = np.asarray(w, dtype=np.float64)[..., None]
    y = np.zeros(np.shape(d), dtype=C128)
    for it in range(p.ntaylor + 1):
        c = np.conj(_scatter_coef(slow[iz], p.eps, it))
        if mode == "exact":
            y = y + (1j * wv * p.dz) * _scatter_kernel(c * d, w, k, p.tap, p.nx, it)
        else:
            y = y + (1j * wv * p.dz) * c * _scatter_kernel(d, w, k, p.tap, p.nx, it)
    return y


def lindown_fwd(m, bg, w, slow, ref, p, out=None, add=False):
    """LinDown::forward (LinOneWay.cpp:33-58): row 0 of the output is
    overwritten with zeros, then out[iz+1] += P_iz Sc_iz(bg[iz] * m[iz])."""
    nz = m.shape[-2]
    out = np.zeros(np.shape(m), dtype=C128) if (out is None or not add) else out.astype(C128)
    out[..., 0, :] = 0
    for iz in range(0, nz - 1):
        t = scatter_fwd(bg[..., iz, :] * m[..., iz, :], iz, w, slow, p)
        out[..., iz + 1, :] += onestep_fwd(t, iz, w, slow, ref, p)
    return out


def lindown_adj(d, bg, w, slow, ref, p, out=None, add=False, mode="compat"):
    """LinDown::adjoint (LinOneWay.cpp:60-87): row nz-1 of the model is
    overwritten with zeros, then m[iz-1] += conj(bg[iz-1]) *
    Sc^H_{iz-1}(P^H_{iz-1} d[iz])."""
    nz = d.shape[-2]
    out = np.zeros(np.shape(d), dtype=C128) if (out is None or not add) else out.astype(C128)
    out[..., nz - 1, :] = 0
    for iz in range(nz - 1, 0, -1):
        t = onestep_adj(d[..., iz, :], iz - 1, w, slow, ref, p)
        t = scatter_adj(t, iz - 1, w, slow, p, mode)
        out[..., iz - 1, :] += np.conj(bg[..., iz - 1, :]) * t
    return out


def linup_fwd(m, bg, w, slow, ref, p, out=None, add=False):
    """LinUp::forward (LinOneWay.cpp:89-115)."""
    nz = m.shape[-2]
    out = np.zeros(np.shape(m), dtype=C128) if (out is None or not add) else out.astype(C128)
    out[..., nz - 1, :] = 0
    for iz in range(nz - 1, 0, -1):
        t = scatter_fwd(bg[..., iz, :] * m[..., iz, :], iz, w, slow, p)
        out[..., iz - 1, :] += onestep_fwd(t, iz, w, slow, ref, p)
    return out


def linup_adj(d, bg, w, slow, ref, p, out=None, add=False, mode="compat"):
    """LinUp::adjoint (LinOneWay.cpp:117-142)."""
    nz = d.shape[-2]
    out = np.zeros(np.shape(d), dtype=C128) if (out is None or not add) else out.astype(C128)
    for iz in range(0, nz - 1):
        t = onestep_adj(d[..., iz, :], iz + 1, w, slow, ref, p)
        t = scatter_adj(t, iz + 1, w, slow, p, mode)
        out[..., iz + 1, :] += np.conj(bg[..., iz + 1, :]) * t
    return out


# --------------------------------------------------------------------------
# IC, Reflect, dReflect
# --------------------------------------------------------------------------

def ic_fwd(bg, m):
    """IC::forward (ic.cpp:5-16): d = bg * m."""
    return bg.astype(C128) * m


def ic_adj(bg, d):
    """IC::adjoint (ic.cpp:17-28): m = conj(bg) * d."""
    return np.conj(bg.astype(C128)) * d


def reflect_coef(slow, den):
    """Reflect.h:19-26: r[iz] = (rho[iz]sqrt(S[iz-1]) - rho[iz-1]sqrt(S[iz])) /
    (rho[iz]sqrt(S[iz-1]) + rho[iz-1]sqrt(S[iz])), r[0] = 0."""
    s = np.asarray(slow, dtype=C128)
    d = np.asarray(den, dtype=C128)
    r = np.zeros_like(s)
    a = d[1:] * np.sqrt(s[:-1])
    b = d[:-1] * np.sqrt(s[1:])
    r[1:] = (a - b) / (a + b)
    return r


def dreflect_tables(slow, den):
    """dReflect.h:11-27 as written: v = (rho2 sqrt(S1) + rho1 sqrt(S2))^-2,
    s1s2sq = rho1*rho2 / v, r1r2sq = 2*S1*S2 / v (1 = iz-1, 2 = iz)."""
    s = np.asarray(slow, dtype=C128)
    d = np.asarray(den, dtype=C128)
    s1s2 = np.array(s, copy=True)
    r1r2 = np.array(s, copy=True)
    v = (d[1:] * np.sqrt(s[:-1]) + d[:-1] * np.sqrt(s[1:])) ** -2
    s1s2[1:] = (d[:-1] * d[1:]) / v
    r1r2[1:] = 2.0 * (s[:-1] * s[1:]) / v
    return s1s2, r1r2


def dreflect_fwd(slow, den, dslow, dden):
    """dReflect::forward (dReflect.h:29-56), rows 1..nz-2."""
    s = np.asarray(slow, dtype=C128)
    dn = np.asarray(den, dtype=C128)
    s1s2, r1r2 = dreflect_tables(s, dn)
    nz = s.shape[0]
    out = np.zeros_like(s)
    iz = np.arange(1, nz - 1)
    c = np.sqrt(s[iz] / s[iz - 1])
    a = c * s1s2[iz]
    b = 1.0 / c * s1s2[iz]
    out[iz] += a * dslow[iz - 1] - b * dslow[iz]
    a2 = dn[iz - 1] * r1r2[iz]
    b2 = dn[iz] * r1r2[iz]
    out[iz] += a2 * dden[iz] - b2 * dden[iz - 1]
    return out


def dreflect_adj(slow, den, dr):
    """dReflect::adjoint (dReflect.h:58-84): returns (dslow, dden)."""
    s = np.asarray(slow, dtype=C128)
    dn = np.asarray(den, dtype=C128)
    s1s2, r1r2 = dreflect_tables(s, dn)
    nz = s.shape[0]
    gs = np.zeros_like(s)
    gd = np.zeros_like(s)
    for iz in range(1, nz - 1):
        c = np.sqrt(s[iz] / s[iz - 1])
        a = np.conj(c * s1s2[iz])
        b = np.conj(1.0 / c * s1s2[iz])
        gs[iz - 1] += a * dr[iz]
        gs[iz] += -b * dr[iz]
        a2 = np.conj(dn[iz - 1] * r1r2[iz])
        b2 = np.conj(dn[iz] * r1r2[iz])
        gd[iz] += a2 * dr[iz]
        gd[iz - 1] += -b2 * dr[iz]
    return gs, gd


# --------------------------------------------------------------------------
# Injection (Injection.cpp)
# --------------------------------------------------------------------------

def inj_points(iz: Sequence[int], ix: Sequence[int], ng: int, tap: int, nx: int):
    """Flattened (iz, x, weight, device) of the cos^2 spread with the in-bounds
    test of Injection.cpp:14-16."""
    g = cos2_window(ng)
    pts = []
    for i, (z, x) in enumerate(zip(iz, ix)):
        for ig in range(-ng, ng + 1):
            off = x + ig + tap
            if 0 <= off < nx:
                pts.append((z, off, g[ig + ng], i))
    return pts


def inj_wavelet_fwd(value: complex, iz, ix, ng, tap, nz, nx):
    """Injection::forward(complex1DReg -> complex2DReg) with add=false
    (Injection.cpp:8-19): a fresh wavefield with the spread source."""
    out = np.zeros((nz, nx), dtype=C128)
    for z, x, g, _ in inj_points(iz, ix, ng, tap, nx):
        out[z, x] += g * value
    return out


def inj_rec_fwd(trace_vals: np.ndarray, iz, ix, ng, tap, nz, nx):
    """Injection::forward(complex3DReg -> complex2DReg) with add=false
    (Injection.cpp:56-67): trace_vals[i] is data[shot][i][step]."""
    out = np.zeros((nz, nx), dtype=C128)
    for z, x, g, i in inj_points(iz, ix, ng, tap, nx):
        out[z, x] += g * trace_vals[i]
    return out


def inj_rec_adj(wfld: np.ndarray, iz, ix, ng, tap):
    """Injection::adjoint(complex3DReg <- complex2DReg) (Injection.cpp:69-80):
    returns the per-device sums to be added into data[shot][i][step]."""
    nx = wfld.shape[-1]
    vals = np.zeros(len(ix), dtype=C128)
    for z, x, g, i in inj_points(iz, ix, ng, tap, nx):
        vals[i] += g * wfld[z, x]
    return vals


# --------------------------------------------------------------------------
# frequency grid, wavelet, devices (PropParam.h, Device.h)
# --------------------------------------------------------------------------

def freq_grid(nt: int, dt: float, ot: float = 0.0, tpad: int = 0,
              fmin: float = 1.0, fmax: float = 0.0, nfreq: int = 0):
    """PropParam.h:49-78 in float32 arithmetic: returns (freq[nw] float32,
    index[nw] int)."""
    n = nt + tpad
    tmax = np.float32(np.float32(ot) + np.float32(dt) * np.float32(n - 1))
    fMIN = np.float32(fmin)
    fMAX = np.float32(fmax)
    if fMAX == 0:
        imin = 1
        imax = nfreq if nfreq else n // 2 + 1
    else:
        imin = int(np.float32(fMIN * tmax))
        imax = int(np.float32(fMAX * tmax + np.float32(0.5)))
    nw = imax - imin + 1
    freq = np.array([np.float32(fMIN + np.float32(i) / tmax) for i in range(nw)],
                    dtype=np.float32)
    index = np.arange(nw) + imin
    return freq, index


def wavelet_spectrum(wave: np.ndarray, tpad: int = 0) -> np.ndarray:
    """PropParam.h:49-54 + FFT1_r.cpp:74-83: unitary FFT of the zero-padded
    real wavelet (all n bins)."""
    w = np.concatenate([np.asarray(wave, dtype=np.float64), np.zeros(tpad)])
    return fft_u(w.astype(C128))


def device_indices(n: int, o_x: float, o_z: float, d_x: float, d_z: float,
                   ox: float, oz: float, dxv: float, dzv: float):
    """Device.h:50-62 in float32: ix = int((x-ox)/dx + .5f)."""
    ixs, izs = [], []
    f = np.float32
    for i in range(n):
        x = f(f(o_x) + f(i) * f(d_x))
        z = f(f(o_z) + f(i) * f(d_z))
        ix = int(f(f(x - f(ox)) / f(dxv) + f(0.5)))
        iz = int(f(f(z - f(oz)) / f(dzv) + f(0.5)))
        if ix < 0 or iz < 0:
            raise ValueError(f"Device {z},{x} is out of bounds!")
        ixs.append(ix)
        izs.append(iz)
    return izs, ixs


def edge_pad(a: np.ndarray, pad: int) -> np.ndarray:
    """Pad.h:77-96 with extend=true along x (last axis)."""
    if pad == 0:
        return np.array(a, dtype=C128, copy=True)
    return np.pad(np.asarray(a, dtype=C128), [(0, 0)] * (a.ndim - 1) + [(pad, pad)], mode="edge")


def zero_pad(a: np.ndarray, pad: int) -> np.ndarray:
    """Pad.h:77-96 with extend=false along x."""
    if pad == 0:
        return np.array(a, dtype=C128, copy=True)
    return np.pad(np.asarray(a, dtype=C128), [(0, 0)] * (a.ndim - 1) + [(pad, pad)])


# --------------------------------------------------------------------------
# operator-level restatements
# --------------------------------------------------------------------------

@dataclass
class Survey:
    """Everything PropParam/Acquisition would derive from the wavelet, the
    model hypercube and the parameter object."""

    nx: int                  # model nx (unpadded)
    nz: int
    dx: float
    dz: float
    freq: np.ndarray         # [nw] Hz
    index: np.ndarray        # [nw] data bin of each frequency
    wave_f: np.ndarray       # wavelet spectrum (all bins)
    src_iz: List[int]
    src_ix: List[int]
    rec_iz: List[int]
    rec_ix: List[int]
    nt_f: int                # data bins per trace
    tap: int = 0
    pad: int = 0
    nref: int = 1
    eps: float = 0.04
    prop: str = "ssf"
    ntaylor: int = 1
    ngs: int = 0
    ngr: int = 0
    ng: int = 0              # BornRefl source spread
    shot_src: Optional[list] = None   # per-shot (iz list, ix list) of mapped sources
    shot_rec: Optional[list] = None   # per-shot (iz list, ix list) of mapped receivers

    def sources(self, s: int):
        return self.shot_src[s] if self.shot_src else ([self.src_iz[s]], [self.src_ix[s]])

    def receivers(self, s: int):
        return self.shot_rec[s] if self.shot_rec else (self.rec_iz, self.rec_ix)

    def step_params(self, padded: bool = True) -> StepParams:
        nxp = self.nx + 2 * self.pad if padded else self.nx
        return StepParams(nx=nxp, dx=self.dx, dz=self.dz, tap=self.tap, pad=self.pad,
                          nref=self.nref, eps=self.eps, prop=self.prop,
                          ntaylor=self.ntaylor)

    @property
    def ns(self) -> int:
        return len(self.src_ix)

    @property
    def nr(self) -> int:
        return len(self.rec_ix)

    def omega(self, i: int) -> float:
        """Phshift.h:24-25 / SplitStep.h:15: w = 2*pi*freq."""
        return 2.0 * math.pi * float(self.freq[i])


def _padded_bg(sv: Survey, slow, den):
    sp = edge_pad(slow, sv.pad)
    dp = edge_pad(den, sv.pad)
    return sp, dp


def wem_forward(sv: Survey, slow, den) -> np.ndarray:
    """WEM::full_fwd (wem_prop.cpp:9-78) into complex data [ns][nr][nt_f]."""
    sp, dp = _padded_bg(sv, slow, den)
    p = sv.step_params()
    ref = refsampler(sp, sv.nref)
    r = reflect_coef(sp, dp)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for s in range(sv.ns):
        for i in range(len(sv.freq)):
            w = sv.omega(i)
            _wem_unit(sv, s, i, sp, ref, r, p, data)
    return data


def _wem_unit(sv, s, i, sp, ref, r, p, data):
    """Per-(shot, frequency) body of WEM::full_fwd (wem_prop.cpp:52-72)."""
    w = sv.omega(i)
    W = inj_wavelet_fwd(sv.wave_f[sv.index[i]], *sv.sources(s), sv.ngs, sv.pad, sv.nz, p.nx)
    W = down_fwd(W, w, sp, ref, p)
    data[s, :, sv.index[i]] += inj_rec_adj(W, *sv.receivers(s), sv.ngr, sv.pad)
    W = r * W
    W = up_fwd(W, w, sp, ref, p)
    data[s, :, sv.index[i]] += inj_rec_adj(W, *sv.receivers(s), sv.ngr, sv.pad)


def _born_unit_fwd(sv, s, i, sp, dp, ref, r, dS, dR, p, data):
    """Per-(shot, frequency) body of BornFull::full_fwd (born_prop.cpp:75-113)."""
    w = sv.omega(i)
    bi = sv.index[i]
    rec = (*sv.receivers(s), sv.ngr, sv.pad)
    bg = inj_wavelet_fwd(sv.wave_f[bi], *sv.sources(s), sv.ngs, sv.pad, sv.nz, p.nx)
    bg = down_fwd(bg, w, sp, ref, p)
    # forward scattering
    sc = lindown_fwd(dS, bg, w, sp, ref, p)
    sc = down_fwd(sc, w, sp, ref, p)
    data[s, :, bi] += inj_rec_adj(sc, *rec)
    # linearisation of reflectivity (Born_refl::forward, Born_refl.cpp:19-27)
    dr = dreflect_fwd(sp, dp, dS, dR)
    sc = up_fwd(ic_fwd(bg, dr), w, sp, ref, p)
    data[s, :, bi] += inj_rec_adj(sc, *rec)
    # down-scattering
    sc = lindown_fwd(dS, bg, w, sp, ref, p)
    sc = down_fwd(sc, w, sp, ref, p)
    sc = up_fwd(r * sc, w, sp, ref, p)
    data[s, :, bi] += inj_rec_adj(sc, *rec)
    # up background wavefield
    bg = up_fwd(r * bg, w, sp, ref, p)
    # up-scattering
    sc = linup_fwd(dS, bg, w, sp, ref, p)
    sc = up_fwd(sc, w, sp, ref, p)
    data[s, :, bi] += inj_rec_adj(sc, *rec)


def _born_unit_adj(sv, s, i, sp, dp, ref, r, data, p, mode):
    """Per-(shot, frequency) body of BornFull::full_adj (born_prop.cpp:178-216);
    returns this unit's (dS, dR) on the padded grid."""
    w = sv.omega(i)
    bi = sv.index[i]
    nz, nxp = sv.nz, p.nx
    tr = data[s, :, bi]
    rec = (*sv.receivers(s), sv.ngr, sv.pad, nz, nxp)
    gS = np.zeros((nz, nxp), dtype=C128)
    gR = np.zeros((nz, nxp), dtype=C128)
    bg = inj_wavelet_fwd(sv.wave_f[bi], *sv.sources(s), sv.ngs, sv.pad, nz, nxp)
    bg = down_fwd(bg, w, sp, ref, p)
    # Born_refl::adjoint (Born_refl.cpp:29-40)
    sc = up_adj(inj_rec_fwd(tr, *rec), w, sp, ref, p)
    a, b = dreflect_adj(sp, dp, ic_adj(bg, sc))
    gS += a
    gR += b
    # forward scattering
    sc = down_adj(inj_rec_fwd(tr, *rec), w, sp, ref, p)
    gS = lindown_adj(sc, bg, w, sp, ref, p, out=gS, add=True, mode=mode)
    # down-scattering
    sc = up_adj(inj_rec_fwd(tr, *rec), w, sp, ref, p)
    sc = down_adj(np.conj(r) * sc, w, sp, ref, p)
    gS = lindown_adj(sc, bg, w, sp, ref, p, out=gS, add=True, mode=mode)
    # up background wavefield
    bg = up_fwd(r * bg, w, sp, ref, p)
    # up-scattering
    sc = up_adj(inj_rec_fwd(tr, *rec), w, sp, ref, p)
    gS = linup_adj(sc, bg, w, sp, ref, p, out=gS, add=True, mode=mode)
    return gS, gR


def born_forward(sv: Survey, slow, den, dslow, dden) -> np.ndarray:
    """BornFull::full_fwd (born_prop.cpp:11-116): model perturbation is
    zero-padded (born_prop.cpp:52-58), background edge-padded."""
    sp, dp = _padded_bg(sv, slow, den)
    p = sv.step_params()
    ref = refsampler(sp, sv.nref)
    r = reflect_coef(sp, dp)
    dS = zero_pad(dslow, sv.pad)
    dR = zero_pad(dden, sv.pad)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for s in range(sv.ns):
        for i in range(len(sv.freq)):
            _born_unit_fwd(sv, s, i, sp, dp, ref, r, dS, dR, p, data)
    return data


def born_adjoint(sv: Survey, slow, den, data, mode: str = "compat"):
    """BornFull::full_adj (born_prop.cpp:119-235) with a deterministic sum
    over (shot, frequency) units, each unit accumulated in isolation."""
    sp, dp = _padded_bg(sv, slow, den)
    p = sv.step_params()
    ref = refsampler(sp, sv.nref)
    r = reflect_coef(sp, dp)
    gS = np.zeros((sv.nz, p.nx), dtype=C128)
    gR = np.zeros((sv.nz, p.nx), dtype=C128)
    for s in range(sv.ns):
        for i in range(len(sv.freq)):
            a, b = _born_unit_adj(sv, s, i, sp, dp, ref, r, data, p, mode)
            gS += a
            gR += b
    sl = slice(sv.pad, sv.pad + sv.nx)
    return gS[:, sl], gR[:, sl]


def bornrefl_forward(sv: Survey, slow, image) -> np.ndarray:
    """BornRefl fwd_propagate (ext_born_refl.cpp:8-56): unpadded grid,
    source spread ng without offset, receivers without spread; image is
    [nw][nz][nx]."""
    sp = np.asarray(slow, dtype=C128)
    p = sv.step_params(padded=False)
    ref = refsampler(sp, sv.nref)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for i in range(len(sv.freq)):
        w = sv.omega(i)
        bi = sv.index[i]
        for s in range(sv.ns):
            bg = inj_wavelet_fwd(sv.wave_f[bi], [sv.src_iz[s]], [sv.src_ix[s]], sv.ng, 0,
                                 sv.nz, sv.nx)
            bg = down_fwd(bg, w, sp, ref, p)
            sc = up_fwd(ic_fwd(bg, image[i]), w, sp, ref, p)
            data[s, :, bi] += inj_rec_adj(sc, sv.rec_iz, sv.rec_ix, 0, 0)
    return data


def bornrefl_adjoint(sv: Survey, slow, data) -> np.ndarray:
    """BornRefl adj_propagate (ext_born_refl.cpp:60-110): image[w] +=
    sum_s conj(D_s) * Up^H(R^H d_s)."""
    sp = np.asarray(slow, dtype=C128)
    p = sv.step_params(padded=False)
    ref = refsampler(sp, sv.nref)
    nw = len(sv.freq)
    image = np.zeros((nw, sv.nz, sv.nx), dtype=C128)
    for i in range(nw):
        w = sv.omega(i)
        bi = sv.index[i]
        for s in range(sv.ns):
            bg = inj_wavelet_fwd(sv.wave_f[bi], [sv.src_iz[s]], [sv.src_ix[s]], sv.ng, 0,
                                 sv.nz, sv.nx)
            bg = down_fwd(bg, w, sp, ref, p)
            sc = inj_rec_fwd(data[s, :, bi], sv.rec_iz, sv.rec_ix, 0, 0, sv.nz, sv.nx)
            sc = up_adj(sc, w, sp, ref, p)
            image[i] += ic_adj(bg, sc)
    return image


def bornrefl_wavefields(sv: Survey, slow, data):
    """Per (shot, frequency) source wavefield D and receiver wavefield
    Up^H(R^H d) of BornRefl::adj_propagate, for the offset imaging condition."""
    sp = np.asarray(slow, dtype=C128)
    p = sv.step_params(padded=False)
    ref = refsampler(sp, sv.nref)
    for i in range(len(sv.freq)):
        w = sv.omega(i)
        bi = sv.index[i]
        for s in range(sv.ns):
            bg = inj_wavelet_fwd(sv.wave_f[bi], [sv.src_iz[s]], [sv.src_ix[s]], sv.ng, 0,
                                 sv.nz, sv.nx)
            bg = down_fwd(bg, w, sp, ref, p)
            sc = inj_rec_fwd(data[s, :, bi], sv.rec_iz, sv.rec_ix, 0, 0, sv.nz, sv.nx)
            sc = up_adj(sc, w, sp, ref, p)
            yield i, s, bg, sc


def offset_ic_adj(bg: np.ndarray, sc: np.ndarray, nh: int) -> np.ndarray:
    """Subsurface-offset imaging condition (no reference counterpart):
    I[h][z][x] = conj(bg[z][x-h]) * sc[z][x+h], h = -(nh//2)..nh//2, zero
    where x-h or x+h leaves [0, nx).  h = 0 is the zero-lag IC of ic.cpp."""
    nz, nx = bg.shape
    out = np.zeros((nh, nz, nx), dtype=C128)
    hs = np.arange(nh) - nh // 2
    for j, h in enumerate(hs):
        x = np.arange(nx)
        ok = (x - h >= 0) & (x - h < nx) & (x + h >= 0) & (x + h < nx)
        xs = x[ok]
        out[j][:, xs] = np.conj(bg[:, xs - h]) * sc[:, xs + h]
    return out


def offset_ic_fwd(bg: np.ndarray, image: np.ndarray) -> np.ndarray:
    """Adjoint of offset_ic_adj with respect to sc: src[z][x+h] +=
    bg[z][x-h] * I[h][z][x]."""
    nh, nz, nx = image.shape
    out = np.zeros((nz, nx), dtype=C128)
    hs = np.arange(nh) - nh // 2
    for j, h in enumerate(hs):
        x = np.arange(nx)
        ok = (x - h >= 0) & (x - h < nx) & (x + h >= 0) & (x + h < nx)
        xs = x[ok]
        out[:, xs + h] += bg[:, xs - h] * image[j][:, xs]
    return out


def offset_forward(sv: Survey, slow, image) -> np.ndarray:
    """Data of the subsurface-offset extended image [nh][nz][nx] (the
    adjoint of the offset migration): per (shot, frequency) D = Down(source)
    as BornRefl::fwd_propagate (ext_born_refl.cpp:8-56), the scattered field
    offset_ic_fwd(D, image) in place of IC, Up, and the receivers."""
    sp = np.asarray(slow, dtype=C128)
    p = sv.step_params(padded=False)
    ref = refsampler(sp, sv.nref)
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for i in range(len(sv.freq)):
        w = sv.omega(i)
        bi = sv.index[i]
        for s in range(sv.ns):
            bg = inj_wavelet_fwd(sv.wave_f[bi], [sv.src_iz[s]], [sv.src_ix[s]], sv.ng, 0, sv.nz, sv.nx)
            bg = down_fwd(bg, w, sp, ref, p)
            sc = up_fwd(offset_ic_fwd(bg, image), w, sp, ref, p)
            data[s, :, bi] += inj_rec_adj(sc, sv.rec_iz, sv.rec_ix, 0, 0)
    return data


def offset_objective(sv: Survey, slow, image, dobs):
    """J = 1/2 ||L I - dobs||^2 over the survey's frequency bins and its
    gradient L^H (L I - dobs) [nh][nz][nx], L = offset_forward; the adjoint
    is the offset migration of bornrefl_wavefields with offset_ic_adj."""
    nh = image.shape[0]
    res = offset_forward(sv, slow, image) - np.asarray(dobs, dtype=C128)
    mask = np.zeros(sv.nt_f, bool)
    mask[np.asarray(sv.index)] = True
    res[:, :, ~mask] = 0.0
    J = 0.5 * float(np.sum(np.abs(res) ** 2))
    grad = np.zeros((nh, sv.nz, sv.nx), dtype=C128)
    for _, _, bg, sc in bornrefl_wavefields(sv, slow, res):
        grad += offset_ic_adj(bg, sc, nh)
    return J, grad


def extwem_forward(sv: Survey, slow3, den3) -> np.ndarray:
    """extWEM::full_fwd (ext_wem_prop.cpp:8-59): one [S, rho] slice per
    frequency, edge-padded; otherwise the WEM recursion."""
    p = sv.step_params()
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for i in range(len(sv.freq)):
        sp, dp = _padded_bg(sv, slow3[i], den3[i])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(sp, dp)
        w = sv.omega(i)
        for s in range(sv.ns):
            W = inj_wavelet_fwd(sv.wave_f[sv.index[i]], [sv.src_iz[s]], [sv.src_ix[s]],
                                sv.ngs, sv.pad, sv.nz, p.nx)
            W = down_fwd(W, w, sp, ref, p)
            data[s, :, sv.index[i]] += inj_rec_adj(W, sv.rec_iz, sv.rec_ix, sv.ngr, sv.pad)
            W = up_fwd(r * W, w, sp, ref, p)
            data[s, :, sv.index[i]] += inj_rec_adj(W, sv.rec_iz, sv.rec_ix, sv.ngr, sv.pad)
    return data


def extborn_forward(sv: Survey, slow3, den3, dslow3, dden3) -> np.ndarray:
    """extBornFull::full_fwd (ext_born_prop.cpp:8-97)."""
    p = sv.step_params()
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for i in range(len(sv.freq)):
        sp, dp = _padded_bg(sv, slow3[i], den3[i])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(sp, dp)
        dS = zero_pad(dslow3[i], sv.pad)
        dR = zero_pad(dden3[i], sv.pad)
        for s in range(sv.ns):
            _born_unit_fwd(sv, s, i, sp, dp, ref, r, dS, dR, p, data)
    return data


def extborn_adjoint(sv: Survey, slow3, den3, data, mode: str = "compat"):
    """extBornFull::full_adj (ext_born_prop.cpp:99-194).

    The serial shot loop shares one model slice per frequency and every
    LinDown::adjoint overwrites its last depth row with zeros
    (LinOneWay.cpp:64-65), so in "compat" mode row nz-1 of the slowness slice
    keeps only the last shot's LinUp contribution; "exact" mode sums all."""
    p = sv.step_params()
    nw = len(sv.freq)
    gS = np.zeros((nw, sv.nz, p.nx), dtype=C128)
    gR = np.zeros((nw, sv.nz, p.nx), dtype=C128)
    for i in range(nw):
        sp, dp = _padded_bg(sv, slow3[i], den3[i])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(sp, dp)
        for s in range(sv.ns):
            a, b = _born_unit_adj(sv, s, i, sp, dp, ref, r, data, p, mode)
            if mode == "compat":
                gS[i, -1] = 0
            gS[i] += a
            gR[i] += b
    sl = slice(sv.pad, sv.pad + sv.nx)
    return gS[:, :, sl], gR[:, :, sl]


# --------------------------------------------------------------------------
# conformal-map grids (one model per shot)
# --------------------------------------------------------------------------

def mapped_models(slow, den, fwd, jac, pad: int):
    """wem_prop.cpp:24-40 for every shot: the model gathered onto the shot's
    grid (fwd[s], -1 = 0), edge-padded; Reflect's slowness is that padded map,
    the propagation slowness the same scaled in complex64 by jac[s][0] then
    jac[s][1], twice, over the first nz*nx samples of the padded array.
    Returns [(scaled slowness, padded density, Reflect slowness)] in C128."""
    s0 = np.asarray(slow, dtype=np.complex64).ravel()
    d0 = np.asarray(den, dtype=np.complex64).ravel()
    nz, nx = np.asarray(slow).shape
    n = nz * nx
    out = []
    for k in range(fwd.shape[0]):
        f = fwd[k]
        g = np.where(f >= 0, s0[np.maximum(f, 0)], 0).astype(np.complex64).reshape(nz, nx)
        gd = np.where(f >= 0, d0[np.maximum(f, 0)], 0).astype(np.complex64).reshape(nz, nx)
        sp = np.pad(g, [(0, 0), (pad, pad)], mode="edge")
        dp = np.pad(gd, [(0, 0), (pad, pad)], mode="edge")
        refl = sp.copy()
        flat = sp.reshape(-1)
        for _ in range(2):
            for j in (jac[k, 0], jac[k, 1]):
                flat[:n] = (flat[:n].real * j + 1j * (flat[:n].imag * j)).astype(np.complex64)
        out.append((sp.astype(C128), dp.astype(C128), refl.astype(C128)))
    return out


def _shot_params(sv: Survey, axes_s) -> StepParams:
    return StepParams(nx=sv.nx + 2 * sv.pad, dx=float(axes_s[1]), dz=float(axes_s[3]), tap=sv.tap, pad=sv.pad,
                      nref=sv.nref, eps=sv.eps, prop=sv.prop, ntaylor=sv.ntaylor)


def wem_forward_mapped(sv: Survey, slow, den, fwd, jac, axes) -> np.ndarray:
    """WEM::full_fwd on the per-shot grids (sv carries the mapped devices)."""
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for s, (sp, dp, rs) in enumerate(mapped_models(slow, den, fwd, jac, sv.pad)):
        p = _shot_params(sv, axes[s])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(rs, dp)
        for i in range(len(sv.freq)):
            _wem_unit(sv, s, i, sp, ref, r, p, data)
    return data


def _gather_pad(m, f, pad):
    """cmap->forward of a perturbation then Pad (born_prop.cpp:50-58)."""
    m0 = np.asarray(m, dtype=C128).ravel()
    nz, nx = np.asarray(m).shape
    g = np.where(f >= 0, m0[np.maximum(f, 0)], 0).reshape(nz, nx)
    return zero_pad(g, pad)


def born_forward_mapped(sv: Survey, slow, den, dslow, dden, fwd, jac, axes) -> np.ndarray:
    """BornFull::full_fwd on the per-shot grids (born_prop.cpp:28-113)."""
    data = np.zeros((sv.ns, sv.nr, sv.nt_f), dtype=C128)
    for s, (sp, dp, rs) in enumerate(mapped_models(slow, den, fwd, jac, sv.pad)):
        p = _shot_params(sv, axes[s])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(rs, dp)
        dS = _gather_pad(dslow, fwd[s], sv.pad)
        dR = _gather_pad(dden, fwd[s], sv.pad)
        for i in range(len(sv.freq)):
            _born_unit_fwd(sv, s, i, sp, dp, ref, r, dS, dR, p, data)
    return data


def born_adjoint_mapped(sv: Survey, slow, den, data, fwd, inv, jac, axes, mode: str = "compat",
                        map_adjoint: str = "inverse"):
    """BornFull::full_adj on the per-shot grids (born_prop.cpp:136-231): each
    shot's (dS, dR), summed over frequencies, is truncated (Pad::adjoint) and
    carried back by the inverse map (map_adjoint='inverse', born_prop.cpp:220,227)
    or by the transpose of the forward gather ('transpose'), then summed over
    shots."""
    nz, nx = sv.nz, sv.nx
    gS = np.zeros(nz * nx, dtype=C128)
    gR = np.zeros(nz * nx, dtype=C128)
    sl = slice(sv.pad, sv.pad + nx)
    for s, (sp, dp, rs) in enumerate(mapped_models(slow, den, fwd, jac, sv.pad)):
        p = _shot_params(sv, axes[s])
        ref = refsampler(sp, sv.nref)
        r = reflect_coef(rs, dp)
        aS = np.zeros((nz, p.nx), dtype=C128)
        aR = np.zeros((nz, p.nx), dtype=C128)
        for i in range(len(sv.freq)):
            a, b = _born_unit_adj(sv, s, i, sp, dp, ref, r, data, p, mode)
            aS += a
            aR += b
        for acc, g in ((gS, aS[:, sl].ravel()), (gR, aR[:, sl].ravel())):
            if map_adjoint == "inverse":
                v = inv[s]
                acc += np.where(v >= 0, g[np.maximum(v, 0)], 0)
            else:
                f = fwd[s]
                np.add.at(acc, f[f >= 0], g[f >= 0])
    return gS.reshape(nz, nx), gR.reshape(nz, nx)


def cdot(a: np.ndarray, b: np.ndarray) -> float:
    """complexHyper::dot (complexHyper.cc:73-84): Re sum a * conj(b)."""
    return float(np.real(np.vdot(np.asarray(b).ravel(), np.asarray(a).ravel())))
