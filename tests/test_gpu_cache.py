"""The library's cache of per-call device buffers (wemfwi_cache_bytes /
wemfwi_cache_release): buffers freed by a call are kept and reused, reused
blocks (holding the previous call's values) give bitwise the same results as
fresh ones, and releasing the cache returns every block."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NZ, NXI = 16, 200


def _shots(torch):
    from wemfwi import Model
    from wemfwi.ops import ImageShot
    rng = np.random.default_rng(2)
    z = np.linspace(0, 1, NZ)[:, None]
    world = (1 / ((1500 + 1500 * z) * (1 + 0.05 * rng.standard_normal((NZ, NXI)))) ** 2).astype(np.complex64)
    shots = []
    for x0, nx, nw in ((0, 128, 3), (60, 140, 4)):
        m = Model(np.ascontiguousarray(world[:, x0:x0 + nx]), 10.0, 10.0, nref=2, prop="pspi", eps=0.04)
        rix = list(range(0, nx, 3))
        d = (rng.standard_normal((len(rix), nw)) + 1j * rng.standard_normal((len(rix), nw))).astype(np.complex64)
        shots.append(ImageShot(m, x0, 3.0 + 2.0 * np.arange(nw), np.ones(nw), 0, nx // 2, [0] * len(rix), rix,
                               torch.from_numpy(d).cuda()))
    return shots


def test_cached_blocks_are_reused_and_released():
    import torch
    import wemfwi
    from wemfwi.ops import uneven_offset_gradient
    shots = _shots(torch)
    nh = 5
    I = (torch.randn((nh, NZ, NXI), dtype=torch.complex64, device="cuda", generator=None) * 0.1)
    g1 = torch.zeros((nh, NZ, NXI), dtype=torch.complex64, device="cuda")
    wemfwi.empty_cache()
    assert wemfwi.cache_bytes() == 0
    J1 = uneven_offset_gradient(shots, nh, I, g1)
    held = wemfwi.cache_bytes()
    print(f"cached after one call: {held} bytes", flush=True)
    assert held > 0
    for _ in range(3):                       # reused blocks, stale contents
        g2 = torch.zeros_like(g1)
        J2 = uneven_offset_gradient(shots, nh, I, g2)
        assert J2 == J1 and torch.equal(g1, g2)
        assert wemfwi.cache_bytes() == held  # the same blocks went back
    wemfwi.empty_cache()
    assert wemfwi.cache_bytes() == 0
    g3 = torch.zeros_like(g1)
    assert uneven_offset_gradient(shots, nh, I, g3) == J1 and torch.equal(g1, g3)
