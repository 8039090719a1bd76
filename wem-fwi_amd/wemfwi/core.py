"""Device-level wrappers over the C ABI: background models and the batched
Down/Up and single-step entry points on torch tensors resident on the GPU.

torch only provides device memory and the stream here; every operation is a
call into libwemfwi.so.
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from ._lib import ModelDesc, check, lib


def _stream_ptr(stream: Optional[torch.cuda.Stream]) -> C.c_void_p:
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _c64_host(a) -> np.ndarray:
    if isinstance(a, torch.Tensor):
        a = a.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex64))


def _dev_c64(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or t.dtype != torch.complex64 or not t.is_cuda:
        raise TypeError(f"{name} must be a complex64 CUDA tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def refsampler(slow, nref: int):
    """RefSampler through the C ABI (host): returns (sref[nz][nref] complex64,
    refid[nz][nx] int32)."""
    s = _c64_host(slow)
    nz, nx = s.shape
    sref = np.zeros((nz, nref), dtype=np.complex64)
    refid = np.zeros((nz, nx), dtype=np.int32)
    check(lib.wemfwi_refsampler(nz, nx, s.ctypes.data, nref, sref.ctypes.data,
                                refid.ctypes.data_as(C.POINTER(C.c_int))))
    return sref, refid


class Model:
    """Background slowness model of the one-way operators on the propagation
    grid (the state OneWay/OneStep/RefSampler derive from 'slow' and 'par')."""

    def __init__(self, slow, dx: float, dz: float, *, tap: int = 0, pad: int = 0,
                 nref: int = 1, prop: str = "ssf", eps: float = 0.04, ntaylor: int = 1,
                 den=None, slow_refl=None):
        """slow_refl: slowness of the Reflect coefficients when it differs
        from the propagation slowness (the conformal-map grids)."""
        s = _c64_host(slow)
        nz, nx = s.shape
        d = None if den is None else _c64_host(den)
        self.desc = ModelDesc(nz=nz, nx=nx, dx=dx, dz=dz, tap=tap, taper=pad, nref=nref,
                              pspi=1 if prop == "pspi" else 0, eps=eps, ntaylor=ntaylor)
        h = C.c_void_p()
        if slow_refl is None:
            check(lib.wemfwi_model_create(C.byref(self.desc), s.ctypes.data,
                                          None if d is None else d.ctypes.data, C.byref(h)))
        else:
            r = _c64_host(slow_refl)
            if r.shape != s.shape:
                raise ValueError("slow_refl must have the shape of slow")
            check(lib.wemfwi_model_create_refl(C.byref(self.desc), s.ctypes.data,
                                               None if d is None else d.ctypes.data, r.ctypes.data,
                                               C.byref(h)))
        self._h = h
        self.nz, self.nx = nz, nx

    def close(self):
        if getattr(self, "_h", None):
            lib.wemfwi_model_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- operators

    def oneway(self, up: bool, adjoint: bool, freq_hz: Sequence[float], inp: torch.Tensor,
               out: Optional[torch.Tensor] = None, add: bool = True,
               stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        """Down (up=False) / Up (up=True) over a batch [nrows][nz][nx]; out=None
        or out is inp runs the in-place recursion."""
        _dev_c64(inp, "inp")
        if out is None:
            out = inp
        _dev_c64(out, "out")
        nrows = inp.shape[0] if inp.dim() == 3 else 1
        if tuple(inp.shape[-2:]) != (self.nz, self.nx) or out.shape != inp.shape:
            raise ValueError("oneway: wavefield shape does not match the model")
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(freq_hz, dtype=np.float32), (nrows,)))
        check(lib.wemfwi_oneway(self._h, int(up), int(adjoint), nrows,
                                f.ctypes.data_as(C.POINTER(C.c_float)),
                                C.c_void_p(inp.data_ptr()), C.c_void_p(out.data_ptr()),
                                int(bool(add)), _stream_ptr(stream)))
        return out

    def oneway_recursive(self, up: bool, adjoint: bool, freq_hz: Sequence[float],
                         src: torch.Tensor, out: torch.Tensor,
                         stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        """out = (I - L)^-1 src: the in-place recursion on a copy of src."""
        _dev_c64(src, "src")
        _dev_c64(out, "out")
        nrows = src.shape[0] if src.dim() == 3 else 1
        if tuple(src.shape[-2:]) != (self.nz, self.nx) or out.shape != src.shape:
            raise ValueError("oneway_recursive: wavefield shape does not match the model")
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(freq_hz, dtype=np.float32), (nrows,)))
        check(lib.wemfwi_oneway_recursive(self._h, int(up), int(adjoint), nrows,
                                          f.ctypes.data_as(C.POINTER(C.c_float)),
                                          C.c_void_p(src.data_ptr()), C.c_void_p(out.data_ptr()),
                                          _stream_ptr(stream)))
        return out

    def set_freqs(self, freq_hz: Sequence[float], stream: Optional[torch.cuda.Stream] = None):
        """Builds the phase tables of a frequency set (Phshift::setFreq for all
        frequencies at once)."""
        f = np.ascontiguousarray(np.asarray(freq_hz, dtype=np.float32).ravel())
        check(lib.wemfwi_model_set_freqs(self._h, len(f), f.ctypes.data_as(C.POINTER(C.c_float)),
                                         _stream_ptr(stream)))

    def onestep(self, adjoint: bool, freq_hz, depth, inp: torch.Tensor, out: torch.Tensor,
                add: bool = False, stream: Optional[torch.cuda.Stream] = None) -> torch.Tensor:
        """SSF/PSPI step on rows [nrows][nx] at depth[i], freq_hz[i]."""
        _dev_c64(inp, "inp")
        _dev_c64(out, "out")
        nrows = inp.shape[0] if inp.dim() == 2 else 1
        if inp.shape[-1] != self.nx or out.shape != inp.shape:
            raise ValueError("onestep: row shape does not match the model")
        f = np.ascontiguousarray(np.broadcast_to(np.asarray(freq_hz, dtype=np.float32), (nrows,)))
        dd = np.ascontiguousarray(np.broadcast_to(np.asarray(depth, dtype=np.int32), (nrows,)))
        check(lib.wemfwi_onestep(self._h, int(adjoint), nrows,
                                 f.ctypes.data_as(C.POINTER(C.c_float)),
                                 dd.ctypes.data_as(C.POINTER(C.c_int)),
                                 C.c_void_p(inp.data_ptr()), C.c_void_p(out.data_ptr()),
                                 int(bool(add)), _stream_ptr(stream)))
        return out


def _batch_args(models, freq_hz, src, out):
    n = len(models)
    if not (len(freq_hz) == len(src) == len(out) == n) or n == 0:
        raise ValueError("batch: one model, frequency, source and output per row")
    for m, a, b in zip(models, src, out):
        _dev_c64(a, "src")
        _dev_c64(b, "out")
        if tuple(a.shape[-2:]) != (m.nz, m.nx) or a.numel() != m.nz * m.nx or b.shape != a.shape:
            raise ValueError("batch: buffer shape does not match its model")
    f = np.ascontiguousarray(np.asarray(freq_hz, dtype=np.float32))
    mh = (C.c_void_p * n)(*[m._h.value for m in models])
    ih = (C.c_void_p * n)(*[a.data_ptr() for a in src])
    oh = (C.c_void_p * n)(*[b.data_ptr() for b in out])
    return n, f, mh, ih, oh


def oneway_batch(up: bool, adjoint: bool, models: Sequence[Model], freq_hz: Sequence[float],
                 src: Sequence[torch.Tensor], out: Sequence[torch.Tensor],
                 stream: Optional[torch.cuda.Stream] = None):
    """Recursive Down/Up of rows with their own models and frequencies in one
    call: out[i] = (I - L_i)^-1 src[i], each [nz_i][nx_i] on its model's grid
    (src[i] may be out[i])."""
    n, f, mh, ih, oh = _batch_args(models, freq_hz, src, out)
    check(lib.wemfwi_oneway_batch(int(up), int(adjoint), n, mh, f.ctypes.data_as(C.POINTER(C.c_float)), ih, oh,
                                  _stream_ptr(stream)))
    return out


class BatchPlan:
    """oneway_batch as a plan: rows fixed once, each run launches only.  The
    models and tensors must stay alive (and keep their frequency sets) while
    the plan is used."""

    def __init__(self, models: Sequence[Model], freq_hz: Sequence[float], src: Sequence[torch.Tensor],
                 out: Sequence[torch.Tensor], stream: Optional[torch.cuda.Stream] = None):
        n, f, mh, ih, oh = _batch_args(models, freq_hz, src, out)
        self._keep = (list(models), list(src), list(out))
        h = C.c_void_p()
        check(lib.wemfwi_batch_create(n, mh, f.ctypes.data_as(C.POINTER(C.c_float)), ih, oh, _stream_ptr(stream),
                                      C.byref(h)))
        self._h = h

    def run(self, up: bool, adjoint: bool, stream: Optional[torch.cuda.Stream] = None):
        check(lib.wemfwi_batch_run(self._h, int(up), int(adjoint), _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib.wemfwi_batch_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


class _BatchOp(C.Structure):
    _fields_ = [("plan", C.c_void_p), ("up", C.c_int), ("adjoint", C.c_int), ("tables", C.c_int)]


class BatchGraph:
    """A sequence of batch-plan sweeps captured once into a hipGraph and
    replayed by launch(): ops are (plan, up, adjoint, tables) with tables =
    True recomputing the plan models' phase tables first.  Creation runs the
    sequence once (the sweep autotuner decides then).  The plans must stay
    alive while the graph is used."""

    def __init__(self, ops, stream: Optional[torch.cuda.Stream] = None):
        ops = list(ops)
        self._keep = [op[0] for op in ops]
        arr = (_BatchOp * len(ops))(*[_BatchOp(p._h, int(up), int(adj), int(tab)) for p, up, adj, tab in ops])
        h = C.c_void_p()
        check(lib.wemfwi_batch_graph_create(C.cast(arr, C.c_void_p), len(ops), _stream_ptr(stream), C.byref(h)))
        self._h = h

    def launch(self, stream: Optional[torch.cuda.Stream] = None):
        check(lib.wemfwi_graph_launch(self._h, _stream_ptr(stream)))

    def close(self):
        if getattr(self, "_h", None):
            lib.wemfwi_graph_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass


def cache_bytes() -> int:
    """Bytes of freed per-call device buffers the library keeps for reuse on
    the current device (wemfwi_cache_bytes)."""
    return int(lib.wemfwi_cache_bytes())


def empty_cache() -> None:
    """Return the library's cached device buffers to the driver
    (wemfwi_cache_release), e.g. before torch needs the memory."""
    check(lib.wemfwi_cache_release())


def timing(enable: bool) -> None:
    """Enables (and clears) or disables HIP-event kernel timing in the library."""
    check(lib.wemfwi_timing(int(enable)))


def autotune_report() -> list:
    """The sweep autotuner's decisions so far: per (FFT length, full row,
    split step, adjoint, device) the radix plan chosen and the measured mean
    launch time of every plan (ms), entries with "kernel": "oneway_kernel";
    per batch group shape the launch order chosen and the best time of every
    candidate order, entries with "kernel": "batch_order"."""
    import json
    n = lib.wemfwi_autotune_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib.wemfwi_autotune_report(buf, n + 1)
    return json.loads(buf.value.decode())


def kernel_time(name: str):
    """(total milliseconds, launch count) of a library kernel since timing(True)."""
    ms = C.c_double()
    n = C.c_long()
    check(lib.wemfwi_timing_get(name.encode(), C.byref(ms), C.byref(n)))
    return ms.value, n.value
