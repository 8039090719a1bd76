"""wemfwi_oneway_noalloc against wemfwi_oneway through the C ABI: the same
Down/Up sweeps (PSPI with three references on full 1024-lane rows, which
park spectra in the scratch rows at multi-reference depths; SSF on padded
600-sample rows of the runtime FFT plan), bitwise equal on two calls in a
row, the second with fewer rows than the first; and the argument errors
(too many rows for the scratch array, a frequency without a phase table)."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _model(nz, nx, nref, prop, tap, seed):
    from wemfwi import Model
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    return Model((1 / v ** 2).astype(np.complex64), 10.0, 10.0, tap=tap, nref=nref, prop=prop, eps=0.04)


def _call(fn, m, up, adjoint, freqs, t, add=1):
    import torch
    from wemfwi._lib import check
    f = np.ascontiguousarray(np.asarray(freqs, dtype=np.float32))
    check(fn(m._h, int(up), int(adjoint), t.shape[0], f.ctypes.data_as(C.POINTER(C.c_float)),
             C.c_void_p(t.data_ptr()), C.c_void_p(t.data_ptr()), add,
             C.c_void_p(torch.cuda.current_stream().cuda_stream)))


@pytest.mark.parametrize("nz,nx,nref,prop,tap", [(40, 1024, 3, "pspi", 0), (30, 590, 2, "ssf", 5)])
def test_noalloc_matches_oneway(nz, nx, nref, prop, tap):
    import torch
    from wemfwi._lib import lib
    m = _model(nz, nx, nref, prop, tap, seed=nx)
    rng = np.random.default_rng(7)
    freqs = np.linspace(3.0, 28.0, 24).astype(np.float32)
    m.set_freqs(freqs)
    cap = lib.wemfwi_oneway_noalloc_max_rows(m._h)
    print(f"nx={nx} {prop} nref={nref}: at most {cap} rows", flush=True)
    assert cap >= 24
    for nrows in (24, 9):     # second call: fewer rows
        w = (rng.standard_normal((nrows, nz, nx)) + 1j * rng.standard_normal((nrows, nz, nx))).astype(np.complex64)
        f = freqs[:nrows]
        for up, adjoint in ((False, False), (False, True), (True, False), (True, True)):
            a = torch.from_numpy(w).cuda()
            b = a.clone()
            print(f"  {nrows} rows up={up} adjoint={adjoint}", flush=True)
            _call(lib.wemfwi_oneway, m, up, adjoint, f, a)
            _call(lib.wemfwi_oneway_noalloc, m, up, adjoint, f, b)
            torch.cuda.synchronize()
            assert torch.isfinite(a.abs()).all() and float(a.abs().max()) > 0
            assert torch.equal(a, b), float((a - b).abs().max())
    m.close()


def test_noalloc_argument_errors():
    import torch
    from wemfwi._lib import ERR_ARG, WemfwiError, lib
    m = _model(8, 2048, 1, "ssf", 0, seed=1)
    freqs = np.array([5.0], np.float32)
    m.set_freqs(freqs)
    cap = lib.wemfwi_oneway_noalloc_max_rows(m._h)
    assert 0 < cap < 8192
    # one row more than the scratch array holds: rejected before any work
    t = torch.zeros((cap + 1, 8, 2048), dtype=torch.complex64, device="cuda")
    f = np.full(cap + 1, 5.0, np.float32)
    rc = lib.wemfwi_oneway_noalloc(m._h, 0, 0, cap + 1, f.ctypes.data_as(C.POINTER(C.c_float)),
                                   C.c_void_p(t.data_ptr()), C.c_void_p(t.data_ptr()), 1, None)
    assert rc == ERR_ARG
    assert b"exceed" in lib.wemfwi_last_error()
    # a frequency without a phase table: the non-allocating call does not
    # build one
    with pytest.raises(WemfwiError) as e:
        _call(lib.wemfwi_oneway_noalloc, m, False, False, [7.0], t[:1])
    assert e.value.code == ERR_ARG
    m.close()
