"""Long recursions keep the norm of propagating waves.

With eps = 0 and a row made only of propagating FFT modes (|k| well below
w sqrt(S)), a homogeneous depth step multiplies each mode by a unit-modulus
phase exp(-i kz dz) (phase_shift.ispc:8-14 with b = 0): Down forward and
adjoint are unitary on such rows, so after 500 steps (the depth of the
deepest benchmark workload) the row's L2 norm must equal the source's to
1e-5.  What FP32 leaves is a random walk of the arithmetic and the
per-mode rounding of the phase table, which repeats at every depth of a
homogeneous medium (a few 1e-6 after 500 steps); a modulus bias of the
FFT's constants compounds on every mode instead (float-rounded radix-16,
radix-8 and pi/8-family twiddles: about -5e-8 per step, -2e-5 to -3.5e-5
after 500) and fails the bound.  Every radix plan of the power-of-two
lengths (FftPlan<N, K>, forced with WEMFWI_FFT_PLAN) and the runtime plans
of other lengths (radix 3 and 5 butterflies) are covered, SSF and PSPI."""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ = 501
TOL = 1e-5

# nx (= FFT length), number of compiled plans (1: runtime plan or one plan)
CASES = [(128, 1), (256, 1), (512, 1), (1024, 3), (2048, 2), (500, 1), (600, 1), (1000, 1)]


@pytest.fixture
def forced_plan():
    old = os.environ.get("WEMFWI_FFT_PLAN")
    yield
    if old is None:
        os.environ.pop("WEMFWI_FFT_PLAN", None)
    else:
        os.environ["WEMFWI_FFT_PLAN"] = old


def _plane_waves(nx, dx, w, s, rng):
    """Rows of FFT modes j with |k_j| < 0.7 w sqrt(S) on the k grid of
    Phshift_cmplx.cpp:12-24 (dk = 2 pi / (dx (nx - 1)))."""
    dk = 2 * np.pi / (dx * (nx - 1))
    jmax = int(0.7 * w * np.sqrt(s) / dk)
    assert 4 <= jmax < nx // 2
    x = np.arange(nx)
    row = np.zeros(nx, complex)
    for j in range(-jmax, jmax + 1):
        c = rng.standard_normal() + 1j * rng.standard_normal()
        row += c * np.exp(2j * np.pi * j * x / nx)
    return row / np.linalg.norm(row)


@pytest.mark.parametrize("prop", ["ssf", "pspi"])
@pytest.mark.parametrize("nx,nplans", CASES)
def test_down_keeps_the_norm_of_propagating_rows(nx, nplans, prop, forced_plan):
    import torch
    from wemfwi import Model
    dx = dz = 10.0
    s = 1.0 / 2000.0 ** 2
    slow = np.full((NZ, nx), s, dtype=np.complex64)
    m = Model(slow, dx, dz, nref=1, prop=prop, eps=0.0)
    freqs = np.array([12.0, 29.0], dtype=np.float32)
    rng = np.random.default_rng(nx)
    rows = [_plane_waves(nx, dx, 2 * np.pi * float(f), s, rng) for f in freqs]
    for k in range(nplans):
        if nplans > 1:
            os.environ["WEMFWI_FFT_PLAN"] = str(k)
        for adjoint in (False, True):
            src = np.zeros((2, NZ, nx), dtype=np.complex64)
            first, last = (NZ - 1, 0) if adjoint else (0, NZ - 1)
            for i in range(2):
                src[i, first] = rows[i]
            t = torch.from_numpy(src).cuda()
            print(f"nx={nx} {prop} plan {k} adjoint={adjoint}: {NZ - 1} steps", flush=True)
            m.oneway(False, adjoint, freqs, t)
            got = t[:, last].cpu().numpy().astype(complex)
            for i in range(2):
                drift = np.linalg.norm(got[i]) / np.linalg.norm(src[i, first].astype(complex)) - 1
                print(f"   {freqs[i]:.0f} Hz: norm drift {drift:+.2e}", flush=True)
                assert abs(drift) <= TOL, (nx, prop, k, adjoint, i, drift)
    m.close()
