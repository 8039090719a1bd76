"""The FP32 C restatement (CPU baseline) against the FP64 oracle."""

import numpy as np
import pytest

from oracle import wem_oracle as O


def test_cpu_baseline_matches_oracle():
    from oracle import cpu_wem
    rng = np.random.default_rng(9)
    nz, nx = 30, 56
    z = np.linspace(0, 1, nz)[:, None]
    s = (1 / ((1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))) ** 2).astype(np.complex64)
    for prop, nref, tap, pad in (("ssf", 1, 4, 6), ("pspi", 3, 4, 0), ("ssf", 2, 36, 8)):
        p = O.StepParams(nx=nx, dx=10.0, dz=10.0, tap=tap, pad=pad, nref=nref, eps=0.04, prop=prop)
        ref = O.refsampler(s, nref)
        W = np.zeros((2, nz, nx), dtype=np.complex64)
        W[:, 0, 20:26] = 1.0
        W[:, nz // 2, 30] = 0.5j
        freqs = np.array([4.0, 18.0], dtype=np.float32)
        for up in (False, True):
            for adj in (False, True):
                got = cpu_wem.oneway(W.copy(), freqs, s, ref, p, up, adj, nthreads=2)
                fn = {(False, False): O.down_fwd, (False, True): O.down_adj,
                      (True, False): O.up_fwd, (True, True): O.up_adj}[(up, adj)]
                for i in range(2):
                    want = fn(W[i].astype(complex), 2 * np.pi * float(freqs[i]), s, ref, p)
                    err = np.linalg.norm(got[i] - want) / np.linalg.norm(want)
                    # single precision with the reference's cancellation-prone
                    # Im(kz) formula: agreement to ~1e-4
                    assert err < 2e-4, (prop, up, adj, err)
