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
