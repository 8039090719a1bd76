"""Host-side acquisition parameters through the C ABI (no GPU): the
frequency grid / wavelet spectrum of PropParam::calcPropParam and the device
indices of Device::createRegGeom against the oracle, including the
reference's own parameter set (propagator/test/test_wem_ops.py:15-47)."""

import numpy as np
import pytest

from oracle import wem_oracle as O


def _ricker(nt, dt, f0=5.0, t0=0.5):
    t = np.arange(nt) * dt - t0
    a = (np.pi * f0 * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


@pytest.mark.parametrize("nt,dt,ot,tpad,fmin,fmax,nfreq", [
    (1001, 0.004, 0.0, 0, 1.0, 0.0, 10),      # the reference test's parameters (fmax=0, nfreq=10)
    (876, 0.004, 0.0, 0, 2.0, 20.0, 0),       # 64 frequencies 2-20 Hz, tmax = 3.5 s
    (500, 0.002, 0.0, 100, 3.0, 45.0, 0),     # tpad pads the trace axis
    (301, 0.008, 0.1, 0, 1.0, 0.0, 0),        # nfreq = 0: (nt+tpad)/2 + 1 bins, shifted origin
])
def test_prop_param_matches_oracle(nt, dt, ot, tpad, fmin, fmax, nfreq):
    from wemfwi.params import prop_param
    wave = _ricker(nt, dt)
    freq, index, wave_f = prop_param(wave, dt, ot=ot, tpad=tpad, fmin=fmin, fmax=fmax, nfreq=nfreq)
    f_o, i_o = O.freq_grid(nt, dt, ot=ot, tpad=tpad, fmin=fmin, fmax=fmax, nfreq=nfreq)
    assert np.array_equal(freq, f_o)          # float32 arithmetic, bit for bit
    assert np.array_equal(index, i_o)
    want = O.wavelet_spectrum(wave, tpad)
    assert np.linalg.norm(wave_f - want) < 1e-6 * np.linalg.norm(want)


def test_prop_param_reference_test_grid_and_64_frequency_grid():
    from wemfwi.params import prop_param
    freq, index, _ = prop_param(np.zeros(1001, np.float32), 0.004, fmin=1, fmax=0, nfreq=10)
    assert list(index) == list(range(1, 11)) and freq[0] == 1.0
    freq, index, _ = prop_param(np.zeros(876, np.float32), 0.004, fmin=2, fmax=20)
    assert len(freq) == 64 and index[0] == 7 and index[-1] == 70


def test_prop_param_rejects_small_capacity_and_empty_range():
    import ctypes as C
    from wemfwi._lib import WemfwiError, check, lib
    nw = C.c_int()
    f = (C.c_float * 2)()
    ix = (C.c_int * 2)()
    rc = lib.wemfwi_prop_param(100, 0.0, 0.004, None, 0, 1.0, 0.0, 10, 2, C.byref(nw), f, ix, None)
    assert rc == 1 and nw.value == 10
    with pytest.raises(WemfwiError):
        check(lib.wemfwi_prop_param(100, 0.0, 0.004, None, 0, 30.0, 5.0, 0, 0, C.byref(nw), None, None, None))


def test_device_indices_match_oracle_and_reference_errors():
    from wemfwi._lib import WemfwiError
    from wemfwi.params import device_indices
    # the reference test's acquisition on a 10 m grid (test_wem_ops.py:15-33)
    par = {"ns": 10, "osx": 1000, "dsx": 10, "osz": 50, "nr": 101, "orx": 0, "drx": 10, "orz": 0}
    for name in ("s", "r"):
        iz, ix = device_indices(par, name, 0.0, 0.0, 10.0, 10.0)
        n = par["n" + name]
        wz, wx = O.device_indices(n, par.get(f"o{name}x", 0), par.get(f"o{name}z", 0),
                                  par.get(f"d{name}x", 0), par.get(f"d{name}z", 0), 0.0, 0.0, 10.0, 10.0)
        assert list(iz) == wz and list(ix) == wx
    # round half up on an off-grid spacing
    iz, ix = device_indices({"ns": 4, "osx": 7.5, "dsx": 12.5, "osz": 0.0}, "s", 0.0, 0.0, 5.0, 5.0)
    assert list(ix) == [2, 4, 7, 9]
    with pytest.raises(WemfwiError):   # no devices
        device_indices({"ns": 0}, "s", 0.0, 0.0, 10.0, 10.0)
    with pytest.raises(WemfwiError):   # spacing zero
        device_indices({"ns": 3, "osx": 10.0}, "s", 0.0, 0.0, 10.0, 10.0)
    with pytest.raises(WemfwiError):   # starts outside the model
        device_indices({"ns": 3, "osx": -10.0, "dsx": 10.0}, "s", 0.0, 0.0, 10.0, 10.0)
