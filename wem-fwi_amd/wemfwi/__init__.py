"""wemfwi -- MI355X-native one-way wave-equation propagator (HIP, gfx950).

The hot path (SSF/PSPI depth stepping, Born operators, imaging condition) runs
in libwemfwi.so; this package is the Python side of its C ABI.
"""

from ._lib import LIB_PATH, WemfwiError, lib  # noqa: F401
from .core import Model, autotune_report, cache_bytes, empty_cache, kernel_time, refsampler, timing  # noqa: F401
from .params import device_indices, prop_param  # noqa: F401

__all__ = ["Model", "refsampler", "WemfwiError", "LIB_PATH", "prop_param", "device_indices",
           "api", "fwi", "ops"]
