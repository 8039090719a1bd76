"""ctypes binding of libwemfwi.so (the C ABI of include/wemfwi.h).

The library is built in-tree (wem-fwi_amd/build.py).  Importing this module
never falls back to another implementation: a missing or unloadable library
raises ImportError, and every failing call raises WemfwiError.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libwemfwi.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "wemfwi.h")

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_UNSUPPORTED = 3
ERR_REFERENCE = 4
ERR_ALLOC = 5


class WemfwiError(RuntimeError):
    """Error returned through the C ABI (message from wemfwi_last_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"wemfwi error {code}: {msg}")
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [
        ("nz", C.c_int), ("nx", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
        ("tap", C.c_int), ("taper", C.c_int), ("nref", C.c_int), ("pspi", C.c_int),
        ("eps", C.c_float), ("ntaylor", C.c_int),
    ]


_vp = C.c_void_p
_i = C.c_int
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)

# name -> (restype, argtypes); every symbol declared in include/wemfwi.h
SIGNATURES = {
    "wemfwi_last_error": (C.c_char_p, []),
    "wemfwi_version": (_i, []),
    "wemfwi_device_count": (_i, []),
    "wemfwi_cache_bytes": (C.c_longlong, []),
    "wemfwi_cache_release": (_i, []),
    "wemfwi_prop_param": (_i, [_i, C.c_float, C.c_float, _fp, _i, C.c_float, C.c_float, _i, _i, _ip,
                               _fp, _ip, _vp]),
    "wemfwi_trace_to_freq": (_i, [_i, _i, _i, _vp, _i, _ip, _i, _vp, _i, _vp]),
    "wemfwi_freq_to_trace": (_i, [_i, _i, _i, _vp, _i, _i, _ip, _vp, _i, _vp]),
    "wemfwi_device_indices": (_i, [_i, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_float, _ip, _ip]),
    "wemfwi_device_coords": (_i, [_i, _fp, _fp, C.c_float, C.c_float, C.c_float, C.c_float, _ip, _ip]),
    "wemfwi_refsampler": (_i, [_i, _i, _vp, _i, _vp, _ip]),
    "wemfwi_model_create": (_i, [C.POINTER(ModelDesc), _vp, _vp, C.POINTER(_vp)]),
    "wemfwi_model_create_refl": (_i, [C.POINTER(ModelDesc), _vp, _vp, _vp, C.POINTER(_vp)]),
    "wemfwi_model_destroy": (_i, [_vp]),
    "wemfwi_oneway": (_i, [_vp, _i, _i, _i, _fp, _vp, _vp, _i, _vp]),
    "wemfwi_onestep": (_i, [_vp, _i, _i, _fp, _ip, _vp, _vp, _i, _vp]),
    "wemfwi_oneway_recursive": (_i, [_vp, _i, _i, _i, _fp, _vp, _vp, _vp]),
    "wemfwi_oneway_noalloc": (_i, [_vp, _i, _i, _i, _fp, _vp, _vp, _i, _vp]),
    "wemfwi_gradient_create": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i, _vp, C.POINTER(_vp)]),
    "wemfwi_gradient_launch": (_i, [_vp, _vp, C.POINTER(C.c_double)]),
    "wemfwi_gradient_misfit": (_i, [_vp, C.POINTER(C.c_double)]),
    "wemfwi_gradient_destroy": (_i, [_vp]),
    "wemfwi_oneway_noalloc_max_rows": (C.c_longlong, [_vp]),
    "wemfwi_oneway_batch": (_i, [_i, _i, _i, _vp, _fp, _vp, _vp, _vp]),
    "wemfwi_batch_create": (_i, [_i, _vp, _fp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "wemfwi_batch_run": (_i, [_vp, _i, _i, _vp]),
    "wemfwi_batch_destroy": (_i, [_vp]),
    "wemfwi_batch_graph_create": (_i, [_vp, _i, _vp, C.POINTER(_vp)]),
    "wemfwi_graph_launch": (_i, [_vp, _vp]),
    "wemfwi_graph_destroy": (_i, [_vp]),
    "wemfwi_model_set_freqs": (_i, [_vp, _i, _fp, _vp]),
    "wemfwi_survey_create": (_i, [_vp, C.POINTER(_vp)]),
    "wemfwi_survey_destroy": (_i, [_vp]),
    "wemfwi_survey_set_checkpoint": (_i, [_vp, _i]),
    "wemfwi_phshift": (_i, [_i, C.c_float, C.c_float, C.c_float, _i, _i, _fp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_splitstep": (_i, [_vp, _i, _i, _fp, _ip, _vp, _i, _ip, _vp, _vp, _i, _vp]),
    "wemfwi_autotune_report": (_i, [C.c_char_p, _i]),
    "wemfwi_uneven_offset_image_adjoint": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _i, _vp]),
    "wemfwi_uneven_offset_gradient": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, C.POINTER(C.c_double), _i, _vp]),
    "wemfwi_uneven_born_forward": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _i, _vp]),
    "wemfwi_uneven_born_adjoint": (_i, [_vp, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp]),
    "wemfwi_uneven_extborn_forward": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _i, _vp]),
    "wemfwi_uneven_extborn_adjoint": (_i, [_vp, _i, _i, _i, _i, _i, _i, _vp, _vp, _i, _i, _vp]),
    "wemfwi_wem_forward": (_i, [_vp, _vp, _vp, _i, _vp]),
    "wemfwi_born_forward": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_born_adjoint": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "wemfwi_extwem_forward": (_i, [_vp, _vp, _vp, _i, _vp]),
    "wemfwi_extborn_forward": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_extborn_adjoint": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "wemfwi_cmap_plan": (_i, [_i, _i, _i, C.c_float, C.c_float, C.c_float, C.c_float, _i, _ip, _ip, _i, _ip,
                              _ip, C.c_float, _fp, _fp, _ip, _ip, _fp, _ip, _ip, _ip, _ip, _ip]),
    "wemfwi_survey_set_shot_devices": (_i, [_vp, _i, _i, _ip, _ip, _ip, _ip]),
    "wemfwi_shotwem_forward": (_i, [_vp, _vp, _vp, _i, _vp]),
    "wemfwi_shotborn_forward": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_shotborn_adjoint": (_i, [_vp, _vp, _vp, _vp, _vp, _i, _i, _vp]),
    "wemfwi_gather_sum": (_i, [C.c_longlong, _vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_bornrefl_forward": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_bornrefl_adjoint": (_i, [_vp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_offset_image_forward": (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp]),
    "wemfwi_offset_image_adjoint": (_i, [_vp, _vp, _i, _vp, _vp, _i, _vp]),
    "wemfwi_scatter": (_i, [_vp, _i, _i, _i, _fp, _ip, _vp, _vp, _i, _vp]),
    "wemfwi_lin_oneway": (_i, [_vp, _i, _i, _i, _i, _fp, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_reflect": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp]),
    "wemfwi_dreflect": (_i, [_vp, _i, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_ic": (_i, [_i, C.c_longlong, _vp, _vp, _vp, _i, _vp]),
    "wemfwi_injection": (_i, [_i, _i, _i, _ip, _ip, _i, _i, _i, _i, _vp, _vp, _i, _vp]),
    "wemfwi_filter_matrix": (_i, [_i, _i, C.POINTER(C.c_double), _i, _i, C.POINTER(C.c_double)]),
    "wemfwi_freq_filter": (_i, [_i, C.c_longlong, _vp, C.c_float, C.c_float, C.c_float, _vp, _vp, _i, _vp]),
    "wemfwi_soft_clip": (_i, [C.c_longlong, _i, _i, C.c_float, C.c_float, C.c_float, _i, _vp, _vp, _vp, _i,
                              _vp]),
    "wemfwi_interp_filter": (_i, [_i, _i, C.c_float, _i, C.c_float, C.c_float, C.c_float, C.c_float, _fp]),
    "wemfwi_axis_apply": (_i, [_vp, _i, _i, C.c_longlong, C.c_longlong, _i, _vp, _vp, _i, _vp]),
    "wemfwi_timing": (_i, [_i]),
    "wemfwi_timing_get": (_i, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libwemfwi.so not built at {LIB_PATH} (run wem-fwi_amd/build.py)")
    # torch ships its own libamdhip64.so.7 and libhsa-runtime64.so.1 (same
    # sonames as /opt/rocm's).  Loaded first, they satisfy this library's
    # dependencies and the process has one HIP runtime; loading this library
    # first would bring in /opt/rocm's copies beside torch's, and the second
    # HSA runtime reports no device.
    import torch  # noqa: F401
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as e:  # pragma: no cover - environment specific
        raise ImportError(f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def check(rc: int) -> None:
    if rc != OK:
        msg = lib.wemfwi_last_error()
        raise WemfwiError(rc, msg.decode() if msg else "")
