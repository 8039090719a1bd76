"""C ABI checks that need no GPU: the library loads, exports every entry
point of include/wemfwi.h, and its host-side RefSampler equals the oracle's
restatement of RefSampler.cpp:10-91 bit for bit."""

import re

import numpy as np
import pytest

from oracle import wem_oracle as O


def _header_symbols():
    from wemfwi._lib import HEADER_PATH
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"\b(wemfwi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from wemfwi._lib import SIGNATURES, lib
    syms = _header_symbols()
    assert len(syms) >= 8
    for s in syms:
        assert hasattr(lib, s), f"{s} not exported"
        assert s in SIGNATURES, f"{s} has no Python signature"
    assert lib.wemfwi_version() >= 100


def _models():
    rng = np.random.default_rng(21)
    out = []
    # smooth random (history-dependent regime after the buffer fills)
    v = 1500 + 3000 * rng.random((70, 24))
    out.append((1 / v ** 2).astype(np.complex64))
    # layered with lateral perturbation, nz > nx/nref
    z = np.linspace(0, 1, 120)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.1 * np.sin(np.arange(40) / 5.0))[None, :]
    out.append((1 / v ** 2 - 1e-10j).astype(np.complex64))
    # exact ties (constant layers)
    v = np.where(np.arange(50)[:, None] < 25, 1000.0, 2000.0) * np.ones((1, 30))
    out.append((1 / v ** 2).astype(np.complex64))
    return out


@pytest.mark.parametrize("nref", [1, 2, 3, 4, 5])
def test_refsampler_abi_equals_oracle(nref):
    from wemfwi import refsampler
    for s in _models():
        sref, refid = refsampler(s, nref)
        t = O.refsampler(s, nref)
        np.testing.assert_array_equal(sref, t.sref)
        np.testing.assert_array_equal(refid, t.refid)


def test_refsampler_abi_rejects_nonpositive_reference():
    from wemfwi import WemfwiError, refsampler
    s = np.ones((4, 10), dtype=np.complex64)
    s[2, 5] = -1
    with pytest.raises(WemfwiError) as e:
        refsampler(s, 1)
    assert e.value.code == 4 and "layer 2" in str(e.value)
