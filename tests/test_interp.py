"""CPU checks of the model interpolation filters (FWI preconditioning): the
oracle pinned by what the kernels must satisfy -- Lanczos interpolates (1 at
its own node, 0 at the other coarse nodes), the CR spline is a partition of
unity that reproduces linear functions away from the edges, the B-spline is
a partition of unity -- and the library's host filters equal the oracle's."""

import numpy as np
import pytest

from oracle import interp_oracle as I


def test_oracle_lanczos_interpolates_at_coarse_nodes():
    F = I.lanczos_filter(41, 1.0, 11, 4.0, 3)
    for j in range(11):
        row = F[4 * j]
        assert abs(row[j] - 1) < 1e-12 and np.all(np.abs(np.delete(row, j)) < 1e-12)
    assert np.all(np.abs(F.sum(axis=1) - 1) < 0.02)


def test_oracle_splines_partition_unity_and_linear_reproduction():
    dm, nm = 5.0, 9
    F = I.spline_filter(41, 1.0, nm, dm, 0.5, 0.0)          # CR
    assert np.allclose(F.sum(axis=1), 1, atol=1e-12)
    lin = 2.0 + 0.3 * np.arange(nm) * dm
    fine = F @ lin
    inner = slice(2 * 5, 41 - 2 * 5)
    assert np.allclose(fine[inner], 2.0 + 0.3 * np.arange(41)[inner], atol=1e-10)
    assert np.allclose(I.spline_filter(41, 1.0, nm, dm, 0.0, 1.0).sum(axis=1), 1, atol=1e-12)   # B


@pytest.mark.parametrize("kind", [0, 1])
def test_library_filters_equal_oracle(kind):
    from wemfwi.interp import interp_filter
    for (nd, dd, nm, dm, a, b, tap) in [(41, 1.0, 11, 4.0, 3, 0, 0.0), (57, 0.7, 9, 5.1, 2, 0, 0.2),
                                        (30, 10.0, 7, 48.3, 0.5, 0.0, 0.1), (30, 10.0, 7, 48.3, 1 / 3, 1 / 3, 0.0)]:
        got = interp_filter(kind, nd, dd, nm, dm, a, b, tap)
        want = I.lanczos_filter(nd, dd, nm, dm, a, tap) if kind == 0 else I.spline_filter(nd, dd, nm, dm, a, b, tap)
        # float32 evaluation as in the reference: the cubic's terms reach
        # 24 a + 8 b before cancelling, so its rounding is ~1e-7 of that
        scale = max(1.0, np.abs(want).max(), 24 * abs(a) + 8 * abs(b))
        assert np.abs(got - want).max() < 2e-6 * scale


def test_oracle_separable_adjoint():
    rng = np.random.default_rng(3)
    fs = [I.lanczos_filter(13, 1.0, 5, 3.0, 2), I.spline_filter(11, 1.0, 4, 3.3, 0.5, 0.0),
          I.lanczos_filter(7, 1.0, 3, 2.0, 1)]
    m = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    d = rng.standard_normal((7, 11, 13)) + 1j * rng.standard_normal((7, 11, 13))
    a = np.vdot(d, I.apply(fs, m))
    b = np.vdot(I.adjoint(fs, d), m)
    assert abs(a - b) < 1e-10 * abs(a)
    # against the explicit triple sum of Interpolation3D.h
    want = np.einsum("ia,jb,kc,cba->kji", fs[0], fs[1], fs[2], m)
    assert np.allclose(I.apply(fs, m), want)
