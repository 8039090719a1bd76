"""ctypes loader of oracle/libcpu_wem.so, the FP32 C restatement used as the
CPU baseline -- test infrastructure only."""

import ctypes as C
import os

import numpy as np

from . import wem_oracle as O

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libcpu_wem.so")


def _lib():
    if not os.path.exists(_LIB):
        raise ImportError(f"{_LIB} not built (make -C oracle)")
    lib = C.CDLL(_LIB)
    f = lib.cpu_oneway
    f.restype = C.c_int
    f.argtypes = [C.c_int] * 7 + [C.c_float] * 3 + [C.c_void_p] * 5 + [C.c_int] * 3
    return lib


def oneway(W: np.ndarray, freq_hz, slow: np.ndarray, ref: "O.RefTables", p: "O.StepParams",
           up: bool, adjoint: bool, nthreads: int = 0) -> np.ndarray:
    """In-place Down/Up over W[nrows][nz][nx] (complex64), returns W."""
    lib = _lib()
    W = np.ascontiguousarray(W, dtype=np.complex64)
    nrows, nz, nx = W.shape
    f = np.ascontiguousarray(np.broadcast_to(np.asarray(freq_hz, dtype=np.float32), (nrows,)))
    s = np.ascontiguousarray(slow, dtype=np.complex64)
    sr = np.ascontiguousarray(ref.sref, dtype=np.complex64)
    rid = np.ascontiguousarray(ref.refid, dtype=np.int32)
    rc = lib.cpu_oneway(nrows, nz, nx, p.tap, p.pad, p.nref, 1 if p.prop == "pspi" else 0,
                        p.dx, p.dz, p.eps, f.ctypes.data, s.ctypes.data, sr.ctypes.data,
                        rid.ctypes.data, W.ctypes.data, int(up), int(adjoint), nthreads)
    if rc != 0:
        raise ValueError("cpu_oneway: FFT length must be a power of two")
    return W
