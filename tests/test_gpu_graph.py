"""Batch sweeps captured into a hipGraph (wemfwi_batch_graph_create /
wemfwi_graph_launch): every replay, with eager work of the program and
synchronisations between replays, equals the eager run of the same plans
bitwise; rows of two widths (two kernel groups on side streams) and the
in-place phase-table recomputation are part of the graph; a rebuilt phase
table makes a launch fail instead of reading a stale table."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _slow(nz, nx, seed):
    rng = np.random.default_rng(seed)
    z = np.linspace(0, 1, nz)[:, None]
    v = (1500 + 2000 * z) * (1 + 0.05 * rng.standard_normal((nz, nx)))
    return (1 / v ** 2).astype(np.complex64)


def _setup():
    import torch
    from wemfwi import Model
    from wemfwi.core import BatchPlan
    nz = 30
    ma = Model(_slow(nz, 256, 1), 10.0, 10.0, nref=2, prop="pspi", eps=0.04)
    mb = Model(_slow(nz, 200, 2), 10.0, 10.0, tap=28, pad=10, nref=1, prop="ssf", eps=0.04)
    rng = np.random.default_rng(3)
    models, freqs, src, out = [], [], [], []
    for i in range(12):
        m, nx = (ma, 256) if i % 3 else (mb, 200)
        s = torch.zeros((nz, nx), dtype=torch.complex64, device="cuda")
        s[0, nx // 2] = complex(rng.standard_normal(), rng.standard_normal())
        s[nz // 2, nx // 3] = 0.5
        models.append(m)
        freqs.append(float(np.float32(3.0 + 1.5 * i)))
        src.append(s)
        out.append(torch.empty_like(s))
    fwd = BatchPlan(models, freqs, src, out)
    adj = BatchPlan(models, freqs, out, out)
    return ma, mb, models, freqs, src, out, fwd, adj


def test_graph_replays_equal_eager_run():
    import torch
    from wemfwi.core import BatchGraph
    ma, mb, models, freqs, src, out, fwd, adj = _setup()
    fwd.run(False, False)
    adj.run(False, True)
    torch.cuda.synchronize()
    ref = [o.clone() for o in out]
    g = BatchGraph([(fwd, False, False, True), (adj, False, True, False)])
    scratch = torch.zeros((4, 30, 256), dtype=torch.complex64, device="cuda")
    for rep in range(4):
        for o in out:
            o.fill_(7.0 + rep)                 # eager work of the program between replays
        # library work on one of the models, with frequencies of its table
        ma.oneway(False, False, np.array([freqs[1], freqs[2], freqs[4], freqs[5]], np.float32), scratch, scratch)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        print(f"graph replay {rep}", flush=True)
        for i, o in enumerate(out):
            assert torch.equal(o, ref[i]), (rep, i)
    g.close()


def test_graph_refuses_a_rebuilt_phase_table():
    import torch
    from wemfwi._lib import WemfwiError
    from wemfwi.core import BatchGraph
    ma, mb, models, freqs, src, out, fwd, adj = _setup()
    g = BatchGraph([(fwd, False, False, False)])
    g.launch()
    ma.set_freqs(np.array([40.0], np.float32))   # another frequency set: the plan's table is gone
    torch.cuda.synchronize()
    with pytest.raises(WemfwiError) as e:
        g.launch()
    assert e.value.code == 1
