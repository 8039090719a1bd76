"""Acquisition parameters of the reference's operators, computed by the
library's host code: the frequency grid + wavelet spectrum of
PropParam::calcPropParam (propagator/src/PropParam.h:45-85) and the regular
device geometry of Device::createRegGeom (propagator/src/Device.h:39-63).

Parameter names are the reference's paramObj keys (fmin, fmax, nfreq, tpad,
ns/osx/osz/dsx/dsz, nr/orx/orz/drx/drz).
"""

from __future__ import annotations

import ctypes as C
from typing import Mapping, Tuple

import numpy as np

from ._lib import check, lib


def prop_param(wave, dt: float, ot: float = 0.0, tpad: int = 0, fmin: float = 1.0,
               fmax: float = 0.0, nfreq: int = 0) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(freq [nw] float32 Hz, index [nw] int32 data bins, wave_f [nt+tpad] complex64)
    exactly as PropParam::calcPropParam derives them from a float wavelet."""
    w = np.ascontiguousarray(np.asarray(wave, dtype=np.float32))
    nt = int(w.size)
    nw = C.c_int()
    check(lib.wemfwi_prop_param(nt, float(ot), float(dt), None, int(tpad), float(fmin), float(fmax),
                                int(nfreq), 0, C.byref(nw), None, None, None))
    n = nw.value
    freq = np.zeros(n, dtype=np.float32)
    index = np.zeros(n, dtype=np.int32)
    wave_f = np.zeros(nt + int(tpad), dtype=np.complex64)
    check(lib.wemfwi_prop_param(nt, float(ot), float(dt), w.ctypes.data_as(C.POINTER(C.c_float)),
                                int(tpad), float(fmin), float(fmax), int(nfreq), n, C.byref(nw),
                                freq.ctypes.data_as(C.POINTER(C.c_float)),
                                index.ctypes.data_as(C.POINTER(C.c_int)),
                                wave_f.ctypes.data_as(C.c_void_p)))
    return freq, index, wave_f


def device_indices(par: Mapping, name: str, ox: float, oz: float, dx: float,
                   dz: float) -> Tuple[np.ndarray, np.ndarray]:
    """(iz, ix) of the devices named 's' (sources) or 'r' (receivers) from the
    parameters n<name>, o<name>x, o<name>z, d<name>x, d<name>z on a model grid
    with origin (ox, oz) and spacing (dx, dz)."""
    if name not in ("s", "r"):
        raise ValueError("Device must be specified to be r or s!")
    n = int(par.get("n" + name, 0))
    iz = np.zeros(max(n, 0), dtype=np.int32)
    ix = np.zeros(max(n, 0), dtype=np.int32)
    check(lib.wemfwi_device_indices(n, float(par.get(f"o{name}x", 0.0)), float(par.get(f"o{name}z", 0.0)),
                                    float(par.get(f"d{name}x", 0.0)), float(par.get(f"d{name}z", 0.0)),
                                    float(ox), float(oz), float(dx), float(dz),
                                    iz.ctypes.data_as(C.POINTER(C.c_int)),
                                    ix.ctypes.data_as(C.POINTER(C.c_int))))
    return iz, ix
