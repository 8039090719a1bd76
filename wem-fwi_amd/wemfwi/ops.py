"""Shot x frequency operators through the C ABI: WEM, BornFull, BornRefl, the
frequency-extended variants and the subsurface-offset image.

Arrays on the propagation grid follow the reference's layout: [nz][nx] with x
fastest; data are complex [ns][nr][nt] with the frequency bins of the survey
filled (Injection.cpp:69-80).
"""

from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from ._lib import check, lib
from .core import Model, _dev_c64, _stream_ptr


class SurveyDesc(C.Structure):
    _fields_ = [
        ("ns", C.c_int), ("src_iz", C.POINTER(C.c_int)), ("src_ix", C.POINTER(C.c_int)),
        ("nr", C.c_int), ("rec_iz", C.POINTER(C.c_int)), ("rec_ix", C.POINTER(C.c_int)),
        ("nw", C.c_int), ("freq_hz", C.POINTER(C.c_float)), ("bin", C.POINTER(C.c_int)),
        ("wavelet", C.c_void_p), ("nt", C.c_int), ("ngs", C.c_int), ("ngr", C.c_int),
        ("offset", C.c_int),
    ]


def _ip(a):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
    return a, a.ctypes.data_as(C.POINTER(C.c_int))


class Survey:
    """Acquisition + frequency grid (PropParam/Device/Injection state)."""

    def __init__(self, src_iz, src_ix, rec_iz, rec_ix, freq_hz, bins, wavelet, nt: int,
                 ngs: int = 0, ngr: int = 0, offset: int = 0):
        self._keep = []
        siz, p_siz = _ip(src_iz)
        six, p_six = _ip(src_ix)
        riz, p_riz = _ip(rec_iz)
        rix, p_rix = _ip(rec_ix)
        f = np.ascontiguousarray(np.asarray(freq_hz, dtype=np.float32))
        b, p_b = _ip(bins)
        wv = np.ascontiguousarray(np.asarray(wavelet, dtype=np.complex64))
        self._keep += [siz, six, riz, rix, f, b, wv]
        d = SurveyDesc(ns=len(siz), src_iz=p_siz, src_ix=p_six, nr=len(riz), rec_iz=p_riz,
                       rec_ix=p_rix, nw=len(f), freq_hz=f.ctypes.data_as(C.POINTER(C.c_float)),
                       bin=p_b, wavelet=wv.ctypes.data, nt=int(nt), ngs=int(ngs), ngr=int(ngr),
                       offset=int(offset))
        h = C.c_void_p()
        check(lib.wemfwi_survey_create(C.byref(d), C.byref(h)))
        self._h = h
        self.ns, self.nr, self.nw, self.nt = len(siz), len(riz), len(f), int(nt)
        self.checkpoint = 0

    def close(self):
        if getattr(self, "_h", None):
            lib.wemfwi_survey_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    @property
    def data_shape(self):
        return (self.ns, self.nr, self.nt)

    def set_checkpoint(self, every: int):
        """Born adjoints with this survey keep the background and receiver
        wavefield rows every `every` depth steps and recompute the rest
        segment by segment (0: keep whole fields).  The result is bitwise
        the same; the per-row device memory drops from 4 nz nx samples to
        about (4 every + 2 nz / every) nx."""
        every = int(every)
        if every < 0:
            raise ValueError("checkpoint interval must be >= 0")
        check(lib.wemfwi_survey_set_checkpoint(self._h, every))
        self.checkpoint = every

    def set_shot_devices(self, shot: int, src_iz, src_ix, rec_iz, rec_ix):
        """Gives shot `shot` its own source points and receiver positions
        (the mapped devices of a conformal-map grid)."""
        siz, p_siz = _ip(src_iz)
        six, p_six = _ip(src_ix)
        riz, p_riz = _ip(rec_iz)
        rix, p_rix = _ip(rec_ix)
        if len(siz) != len(six) or len(riz) != self.nr or len(rix) != self.nr:
            raise ValueError("set_shot_devices: device list sizes")
        check(lib.wemfwi_survey_set_shot_devices(self._h, int(shot), len(siz), p_siz, p_six, p_riz, p_rix))


def _ptr(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def _models(models: Sequence[Model]):
    arr = (C.c_void_p * len(models))(*[m._h.value for m in models])
    return arr


def wem_forward(model: Model, survey: Survey, data: torch.Tensor, add: bool = False, stream=None):
    _dev_c64(data, "data")
    check(lib.wemfwi_wem_forward(model._h, survey._h, _ptr(data), int(add), _stream_ptr(stream)))
    return data


def extwem_forward(models: Sequence[Model], survey: Survey, data: torch.Tensor, add: bool = False,
                   stream=None):
    _dev_c64(data, "data")
    check(lib.wemfwi_extwem_forward(_models(models), survey._h, _ptr(data), int(add),
                                    _stream_ptr(stream)))
    return data


def born_forward(model: Model, survey: Survey, dslow: torch.Tensor, dden: torch.Tensor,
                 data: torch.Tensor, add: bool = False, stream=None):
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    check(lib.wemfwi_born_forward(model._h, survey._h, _ptr(dslow), _ptr(dden), _ptr(data),
                                  int(add), _stream_ptr(stream)))
    return data


def born_adjoint(model: Model, survey: Survey, dslow: torch.Tensor, dden: torch.Tensor,
                 data: torch.Tensor, add: bool = False, exact: bool = False, stream=None):
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    check(lib.wemfwi_born_adjoint(model._h, survey._h, _ptr(dslow), _ptr(dden), _ptr(data),
                                  int(add), int(exact), _stream_ptr(stream)))
    return dslow, dden


class GradientPlan:
    """The FWI gradient of one model as a replayable graph
    (wemfwi_gradient_create): res = WEM(model) - obs, J = 0.5 ||res||^2 and
    (gS, gR) = Born(model)^H res (exact adjoint by default), every
    (shot, frequency) row batched.  Creating the plan computes the gradient
    once; run() replays the captured graph and returns J; gS, gR and res
    hold the results.  A new model needs a new plan."""

    def __init__(self, model: Model, survey: Survey, obs: torch.Tensor, exact: bool = True, stream=None):
        _dev_c64(obs, "obs")
        self.obs = obs
        self.res = torch.empty_like(obs)
        self.gS = torch.empty((model.nz, model.nx), dtype=torch.complex64, device=obs.device)
        self.gR = torch.empty_like(self.gS)
        h = C.c_void_p()
        check(lib.wemfwi_gradient_create(model._h, survey._h, _ptr(obs), _ptr(self.res), _ptr(self.gS),
                                         _ptr(self.gR), int(bool(exact)), _stream_ptr(stream), C.byref(h)))
        self._h = h
        self._keep = (model, survey)

    @property
    def misfit(self) -> float:
        j = C.c_double()
        check(lib.wemfwi_gradient_misfit(self._h, C.byref(j)))
        return j.value

    def run(self, stream=None) -> float:
        j = C.c_double()
        check(lib.wemfwi_gradient_launch(self._h, _stream_ptr(stream), C.byref(j)))
        return j.value

    def close(self):
        if getattr(self, "_h", None):
            lib.wemfwi_gradient_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass


def extborn_forward(models: Sequence[Model], survey: Survey, dslow, dden, data, add=False,
                    stream=None):
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    check(lib.wemfwi_extborn_forward(_models(models), survey._h, _ptr(dslow), _ptr(dden),
                                     _ptr(data), int(add), _stream_ptr(stream)))
    return data


def extborn_adjoint(models: Sequence[Model], survey: Survey, dslow, dden, data, add=False,
                    exact=False, stream=None):
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    check(lib.wemfwi_extborn_adjoint(_models(models), survey._h, _ptr(dslow), _ptr(dden),
                                     _ptr(data), int(add), int(exact), _stream_ptr(stream)))
    return dslow, dden


def bornrefl_forward(model: Model, survey: Survey, image, data, add=False, stream=None):
    _dev_c64(image, "image")
    _dev_c64(data, "data")
    check(lib.wemfwi_bornrefl_forward(model._h, survey._h, _ptr(image), _ptr(data), int(add),
                                      _stream_ptr(stream)))
    return data


def bornrefl_adjoint(model: Model, survey: Survey, image, data, add=False, stream=None):
    _dev_c64(image, "image")
    _dev_c64(data, "data")
    check(lib.wemfwi_bornrefl_adjoint(model._h, survey._h, _ptr(image), _ptr(data), int(add),
                                      _stream_ptr(stream)))
    return image


def offset_image_forward(model: Model, survey: Survey, nh: int, image, data, add=False,
                         stream=None):
    _dev_c64(image, "image")
    _dev_c64(data, "data")
    check(lib.wemfwi_offset_image_forward(model._h, survey._h, int(nh), _ptr(image), _ptr(data),
                                          int(add), _stream_ptr(stream)))
    return data


def offset_image_adjoint(model: Model, survey: Survey, nh: int, image, data, add=False,
                         stream=None):
    _dev_c64(image, "image")
    _dev_c64(data, "data")
    check(lib.wemfwi_offset_image_adjoint(model._h, survey._h, int(nh), _ptr(image), _ptr(data),
                                          int(add), _stream_ptr(stream)))
    return image


class ImageShotDesc(C.Structure):
    _fields_ = [
        ("model", C.c_void_p), ("x0", C.c_int), ("nw", C.c_int), ("freq_hz", C.POINTER(C.c_float)),
        ("wavelet", C.c_void_p), ("src_iz", C.c_int), ("src_ix", C.c_int), ("nr", C.c_int),
        ("rec_iz", C.POINTER(C.c_int)), ("rec_ix", C.POINTER(C.c_int)), ("data", C.c_void_p),
    ]


class ImageShot:
    """One shot of an uneven migration: its background `model` on its
    aperture (nz x nx_s), placed at image column x0; its frequencies and
    source spectrum; its source point and receivers on the aperture grid; its
    data, a CUDA complex64 tensor [nr][nw]."""

    def __init__(self, model: Model, x0: int, freq_hz, wavelet, src_iz: int, src_ix: int, rec_iz, rec_ix,
                 data: torch.Tensor):
        self.model = model
        self.x0 = int(x0)
        self.freq = np.ascontiguousarray(np.asarray(freq_hz, dtype=np.float32))
        self.wavelet = np.ascontiguousarray(np.asarray(wavelet, dtype=np.complex64))
        self.src = (int(src_iz), int(src_ix))
        self.rec_iz, _ = _ip(rec_iz)
        self.rec_ix, _ = _ip(rec_ix)
        if len(self.wavelet) != len(self.freq) or len(self.rec_iz) != len(self.rec_ix):
            raise ValueError("ImageShot: frequency / wavelet or receiver list sizes differ")
        _dev_c64(data, "data")
        if tuple(data.shape) != (len(self.rec_iz), len(self.freq)):
            raise ValueError("ImageShot: data must be [nr][nw]")
        self.data = data

    def desc(self) -> ImageShotDesc:
        return ImageShotDesc(model=self.model._h, x0=self.x0, nw=len(self.freq),
                             freq_hz=self.freq.ctypes.data_as(C.POINTER(C.c_float)),
                             wavelet=self.wavelet.ctypes.data, src_iz=self.src[0], src_ix=self.src[1],
                             nr=len(self.rec_iz), rec_iz=self.rec_iz.ctypes.data_as(C.POINTER(C.c_int)),
                             rec_ix=self.rec_ix.ctypes.data_as(C.POINTER(C.c_int)), data=_ptr(self.data))


def uneven_offset_image_adjoint(shots: Sequence[ImageShot], nh: int, image: torch.Tensor, ngs: int = 0,
                                add: bool = False, stream=None) -> torch.Tensor:
    """Subsurface-offset migration of shots with their own apertures and
    frequency bands into image [nh][nz][nx_image] (CUDA complex64): all
    (shot, frequency) rows swept in batched launches, the image summed in
    FP64 shot by shot."""
    _dev_c64(image, "image")
    if image.dim() != 3 or image.shape[0] != nh:
        raise ValueError("image must be [nh][nz][nx_image]")
    arr = (ImageShotDesc * max(1, len(shots)))(*[sh.desc() for sh in shots])
    check(lib.wemfwi_uneven_offset_image_adjoint(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(nh), int(image.shape[1]),
                                                 int(image.shape[2]), _ptr(image), int(add),
                                                 _stream_ptr(stream)))
    return image


def uneven_offset_gradient(shots: Sequence[ImageShot], nh: int, model: torch.Tensor, grad: torch.Tensor,
                           ngs: int = 0, add: bool = False, stream=None) -> float:
    """Extended least-squares objective J = 1/2 sum_s ||L_s I - d_s||^2 of
    shots with their own apertures and bands (each ImageShot's data are its
    observed data) and its gradient sum_s L_s^H (L_s I - d_s) into grad;
    model I and grad are [nh][nz][nx_image] (CUDA complex64).  L_s is the
    adjoint of the shot's offset migration (uneven_offset_image_adjoint).
    Returns J (FP64 sum)."""
    _dev_c64(model, "model")
    _dev_c64(grad, "grad")
    if model.dim() != 3 or model.shape[0] != nh or tuple(grad.shape) != tuple(model.shape):
        raise ValueError("model and grad must both be [nh][nz][nx_image]")
    arr = (ImageShotDesc * max(1, len(shots)))(*[sh.desc() for sh in shots])
    j = C.c_double(0.0)
    check(lib.wemfwi_uneven_offset_gradient(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(nh),
                                            int(model.shape[1]), int(model.shape[2]), _ptr(model), _ptr(grad),
                                            C.byref(j), int(add), _stream_ptr(stream)))
    return float(j.value)


class BornShot(ImageShot):
    """One shot of the uneven Born operators: its background `model` on its
    aperture (nz x nx_s), placed at column x0 of the perturbation grid; its
    frequencies and source spectrum; its source point and receivers on the
    aperture grid; its data, a CUDA complex64 tensor [nr][nw] (written by
    uneven_born_forward, read by uneven_born_adjoint)."""


def _born_shots(shots: Sequence[BornShot]):
    return (ImageShotDesc * max(1, len(shots)))(*[sh.desc() for sh in shots])


def uneven_born_forward(shots: Sequence[BornShot], dslow: torch.Tensor, dden: torch.Tensor, ngs: int = 0,
                        ngr: int = 0, add: bool = False, stream=None):
    """BornFull::forward of shots with their own apertures and bands
    (BornFull.cpp:64-117 per shot): every shot's data [nr][nw] receives the
    scattered field of the perturbations dslow, dden [nz][nx_model] (CUDA
    complex64) windowed to its aperture; all rows swept in batched
    launches."""
    _dev_c64(dslow, "dslow")
    _dev_c64(dden, "dden")
    if dslow.dim() != 2 or tuple(dden.shape) != tuple(dslow.shape):
        raise ValueError("dslow and dden must both be [nz][nx_model]")
    arr = _born_shots(shots)
    check(lib.wemfwi_uneven_born_forward(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(ngr),
                                         int(dslow.shape[0]), int(dslow.shape[1]), _ptr(dslow), _ptr(dden),
                                         int(add), _stream_ptr(stream)))
    return [sh.data for sh in shots]


def uneven_born_adjoint(shots: Sequence[BornShot], dslow: torch.Tensor, dden: torch.Tensor, ngs: int = 0,
                        ngr: int = 0, add: bool = False, exact: bool = False, stream=None):
    """BornFull::adjoint of the same shots: the gradient slices of every
    shot's data, summed over shots in shot order in FP64 into dslow, dden
    [nz][nx_model]; exact selects the exact Scatter adjoint (as born_adjoint)."""
    _dev_c64(dslow, "dslow")
    _dev_c64(dden, "dden")
    if dslow.dim() != 2 or tuple(dden.shape) != tuple(dslow.shape):
        raise ValueError("dslow and dden must both be [nz][nx_model]")
    arr = _born_shots(shots)
    check(lib.wemfwi_uneven_born_adjoint(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(ngr),
                                         int(dslow.shape[0]), int(dslow.shape[1]), _ptr(dslow), _ptr(dden),
                                         int(add), int(exact), _stream_ptr(stream)))
    return dslow, dden


class ExtBornShotDesc(C.Structure):
    _fields_ = [
        ("models", C.c_void_p), ("x0", C.c_int), ("nw", C.c_int), ("w0", C.c_int), ("freq_hz", C.POINTER(C.c_float)),
        ("wavelet", C.c_void_p), ("src_iz", C.c_int), ("src_ix", C.c_int), ("nr", C.c_int),
        ("rec_iz", C.POINTER(C.c_int)), ("rec_ix", C.POINTER(C.c_int)), ("data", C.c_void_p),
    ]


class ExtBornShot:
    """One shot of the uneven extBorn operators: one background model per
    frequency (models[i], nz x nx_s, the shot's aperture) placed at column x0
    of the perturbation grid; frequency i meets perturbation slice w0 + i;
    its frequencies, source spectrum, source point, receivers (aperture grid)
    and data, a CUDA complex64 tensor [nr][nw]."""

    def __init__(self, models: Sequence[Model], x0: int, w0: int, freq_hz, wavelet, src_iz: int, src_ix: int,
                 rec_iz, rec_ix, data: torch.Tensor):
        self.models = list(models)
        self.x0, self.w0 = int(x0), int(w0)
        self.freq = np.ascontiguousarray(np.asarray(freq_hz, dtype=np.float32))
        self.wavelet = np.ascontiguousarray(np.asarray(wavelet, dtype=np.complex64))
        self.src = (int(src_iz), int(src_ix))
        self.rec_iz, _ = _ip(rec_iz)
        self.rec_ix, _ = _ip(rec_ix)
        if len(self.models) != len(self.freq) or len(self.wavelet) != len(self.freq):
            raise ValueError("ExtBornShot: one model and one wavelet sample per frequency")
        if len(self.rec_iz) != len(self.rec_ix):
            raise ValueError("ExtBornShot: receiver list sizes differ")
        _dev_c64(data, "data")
        if tuple(data.shape) != (len(self.rec_iz), len(self.freq)):
            raise ValueError("ExtBornShot: data must be [nr][nw]")
        self.data = data
        self._mh = _models(self.models)

    def desc(self) -> ExtBornShotDesc:
        return ExtBornShotDesc(models=C.cast(self._mh, C.c_void_p), x0=self.x0, nw=len(self.freq), w0=self.w0,
                               freq_hz=self.freq.ctypes.data_as(C.POINTER(C.c_float)),
                               wavelet=self.wavelet.ctypes.data, src_iz=self.src[0], src_ix=self.src[1],
                               nr=len(self.rec_iz), rec_iz=self.rec_iz.ctypes.data_as(C.POINTER(C.c_int)),
                               rec_ix=self.rec_ix.ctypes.data_as(C.POINTER(C.c_int)), data=_ptr(self.data))


def _ext_slices(dslow: torch.Tensor, dden: torch.Tensor):
    _dev_c64(dslow, "dslow")
    _dev_c64(dden, "dden")
    if dslow.dim() != 3 or tuple(dden.shape) != tuple(dslow.shape):
        raise ValueError("dslow and dden must both be [nslice][nz][nx_model]")


def uneven_extborn_forward(shots: Sequence[ExtBornShot], dslow: torch.Tensor, dden: torch.Tensor, ngs: int = 0,
                           ngr: int = 0, add: bool = False, stream=None):
    """extBornFull::forward of shots with their own apertures and bands
    (ext_born_prop.cpp:8-97 per shot): every shot's data [nr][nw] receives
    the scattered field of its frequencies' perturbation slices w0 + i of
    dslow, dden [nslice][nz][nx_model] (CUDA complex64) windowed to its
    aperture; all rows swept in batched launches."""
    _ext_slices(dslow, dden)
    arr = (ExtBornShotDesc * max(1, len(shots)))(*[sh.desc() for sh in shots])
    check(lib.wemfwi_uneven_extborn_forward(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(ngr),
                                            int(dslow.shape[1]), int(dslow.shape[2]), int(dslow.shape[0]),
                                            _ptr(dslow), _ptr(dden), int(add), _stream_ptr(stream)))
    return [sh.data for sh in shots]


def uneven_extborn_adjoint(shots: Sequence[ExtBornShot], dslow: torch.Tensor, dden: torch.Tensor, ngs: int = 0,
                           ngr: int = 0, add: bool = False, exact: bool = False, stream=None):
    """extBornFull::adjoint of the same shots into the slices dslow, dden
    [nslice][nz][nx_model], summed over shots in shot order in FP64; exact =
    False keeps the reference's last-row rule of its serial shot loop."""
    _ext_slices(dslow, dden)
    arr = (ExtBornShotDesc * max(1, len(shots)))(*[sh.desc() for sh in shots])
    check(lib.wemfwi_uneven_extborn_adjoint(C.cast(arr, C.c_void_p), len(shots), int(ngs), int(ngr),
                                            int(dslow.shape[1]), int(dslow.shape[2]), int(dslow.shape[0]),
                                            _ptr(dslow), _ptr(dden), int(add), int(exact), _stream_ptr(stream)))
    return dslow, dden


def shotwem_forward(models: Sequence[Model], survey: Survey, data: torch.Tensor, add: bool = False,
                    stream=None):
    """WEM with one background model per shot (conformal-map grids)."""
    _dev_c64(data, "data")
    if len(models) != survey.ns:
        raise ValueError("shotwem_forward: one model per shot")
    check(lib.wemfwi_shotwem_forward(_models(models), survey._h, _ptr(data), int(add), _stream_ptr(stream)))
    return data


def shotborn_forward(models: Sequence[Model], survey: Survey, dslow, dden, data, add=False, stream=None):
    """BornFull with one background model and perturbation slice per shot:
    dslow, dden [ns][nz][nx]."""
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    if len(models) != survey.ns:
        raise ValueError("shotborn_forward: one model per shot")
    check(lib.wemfwi_shotborn_forward(_models(models), survey._h, _ptr(dslow), _ptr(dden), _ptr(data),
                                      int(add), _stream_ptr(stream)))
    return data


def shotborn_adjoint(models: Sequence[Model], survey: Survey, dslow, dden, data, add=False, exact=False,
                     stream=None):
    """Adjoint of shotborn_forward into the per-shot slices dslow, dden."""
    for n, t in (("dslow", dslow), ("dden", dden), ("data", data)):
        _dev_c64(t, n)
    if len(models) != survey.ns:
        raise ValueError("shotborn_adjoint: one model per shot")
    check(lib.wemfwi_shotborn_adjoint(_models(models), survey._h, _ptr(dslow), _ptr(dden), _ptr(data),
                                      int(add), int(exact), _stream_ptr(stream)))
    return dslow, dden


class GatherPlan:
    """out[j] = sum of in[src[c]] for c in [ptr[j], ptr[j+1]) on the device:
    the model <-> grid gathers of the conformal maps."""

    def __init__(self, ptr: np.ndarray, src: np.ndarray, device="cuda"):
        ptr = np.ascontiguousarray(np.asarray(ptr, dtype=np.int64))
        src = np.ascontiguousarray(np.asarray(src, dtype=np.int64))
        if ptr.ndim != 1 or ptr.size < 1 or ptr[0] != 0 or ptr[-1] != src.size or np.any(np.diff(ptr) < 0):
            raise ValueError("GatherPlan: invalid pointer array")
        self.nout = ptr.size - 1
        self.nin_min = int(src.max()) + 1 if src.size else 0
        self.ptr = torch.from_numpy(ptr).to(device)
        self.src = torch.from_numpy(src if src.size else np.zeros(1, np.int64)).to(device)

    @staticmethod
    def from_index(idx: np.ndarray, offsets: Optional[np.ndarray] = None) -> "GatherPlan":
        """Gather out[j] = in[idx[j] + offsets[j]] (idx < 0: zero)."""
        idx = np.asarray(idx, dtype=np.int64).ravel()
        ok = idx >= 0
        src = idx[ok] + (0 if offsets is None else np.asarray(offsets, dtype=np.int64).ravel()[ok])
        ptr = np.concatenate([[0], np.cumsum(ok)])
        return GatherPlan(ptr, src)

    def __call__(self, inp: torch.Tensor, out: torch.Tensor, add: bool = False, stream=None) -> torch.Tensor:
        _dev_c64(inp, "inp")
        _dev_c64(out, "out")
        if out.numel() != self.nout or inp.numel() < self.nin_min:
            raise ValueError("GatherPlan: input / output sizes do not match the plan")
        check(lib.wemfwi_gather_sum(self.nout, _ptr(self.ptr), _ptr(self.src), _ptr(inp), _ptr(out), int(add),
                                    _stream_ptr(stream)))
        return out
