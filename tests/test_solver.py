"""CPU checks of the FWI driver solvers (wemfwi.solver: the restated
python-solver LBFGS / NLCG / LCG with bounds) on problems with known
minimisers, and of shot sharding: the contiguous shot split and the one
all-reduce, run with two gloo ranks."""

import math
import os

import numpy as np
import pytest
import torch

from wemfwi import solver as S


def _quadratic(n=12, seed=0, complex_=True):
    g = torch.Generator().manual_seed(seed)
    dt = torch.complex128 if complex_ else torch.float64
    A = torch.randn(n + 4, n, generator=g, dtype=torch.float64)
    if complex_:
        A = torch.complex(A, torch.randn(n + 4, n, generator=g, dtype=torch.float64))
    b = torch.randn(n + 4, generator=g, dtype=torch.float64).to(dt)
    xs = torch.linalg.lstsq(A, b.unsqueeze(1)).solution.squeeze(1)

    def fg(x):
        r = A @ x[0] - b
        return 0.5 * float((r.conj() * r).real.sum()), [A.conj().T @ r]
    return fg, xs, A, b


@pytest.mark.parametrize("complex_", [False, True])
def test_lbfgs_reaches_least_squares_minimum(complex_):
    fg, xs, _, _ = _quadratic(complex_=complex_)
    x, h = S.LBFGS(40, m_steps=6).run(S.Problem(fg), [torch.zeros_like(xs)])
    assert torch.linalg.norm(x[0] - xs) < 1e-6 * torch.linalg.norm(xs)
    assert all(b <= a + 1e-12 for a, b in zip(h.objective, h.objective[1:]))


def test_lbfgs_rosenbrock():
    def fg(x):
        a, b = x[0][0], x[0][1]
        f = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        g = torch.stack([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])
        return float(f), [g]
    x, h = S.LBFGS(100).run(S.Problem(fg), [torch.tensor([-1.2, 1.0], dtype=torch.float64)])
    assert torch.allclose(x[0], torch.tensor([1.0, 1.0], dtype=torch.float64), atol=1e-5)


@pytest.mark.parametrize("beta", ["FR", "PRP", "HS", "HZ"])
def test_nlcg_reaches_least_squares_minimum(beta):
    fg, xs, _, _ = _quadratic(n=8, seed=2)
    x, h = S.NLCG(60, beta_type=beta).run(S.Problem(fg), [torch.zeros_like(xs)])
    assert torch.linalg.norm(x[0] - xs) < 1e-5 * torch.linalg.norm(xs)


def test_bounds_project_the_real_part():
    # minimiser of 1/2 |x - c|^2 with c outside the box sits on the bound
    c = torch.tensor([3.0 + 1j, -2.0 - 1j, 0.5 + 0.2j], dtype=torch.complex128)

    def fg(x):
        r = x[0] - c
        return 0.5 * float((r.conj() * r).real.sum()), [r]
    x, _ = S.LBFGS(20).run(S.Problem(fg, -1.0, 1.0), [torch.zeros_like(c)])
    assert torch.allclose(x[0].real, torch.tensor([1.0, -1.0, 0.5], dtype=torch.float64), atol=1e-8)
    assert torch.allclose(x[0].imag, c.imag, atol=1e-8)


def test_cgls_solves_normal_equations():
    _, xs, A, b = _quadratic(n=10, seed=5)
    x = S.cgls(lambda v: [A @ v[0]], lambda r: [A.conj().T @ r[0]], [torch.zeros_like(xs)], [b], 30)
    assert torch.linalg.norm(x[0] - xs) < 1e-8 * torch.linalg.norm(xs)


def test_shot_ranges_partition_shots():
    for ns in (1, 5, 32, 33):
        for world in (1, 2, 3, 8):
            rs = [S.shot_range(ns, r, world) for r in range(world)]
            assert rs[0][0] == 0 and rs[-1][1] == ns
            assert all(a[1] == b[0] for a, b in zip(rs, rs[1:]))
            sizes = [b - a for a, b in rs]
            assert max(sizes) - min(sizes) <= 1


def _rank(rank, world, port, out):
    import torch.distributed as dist
    os.environ.update(MASTER_ADDR="127.0.0.1", MASTER_PORT=str(port))
    dist.init_process_group("gloo", rank=rank, world_size=world)
    # per-shot objective 1/2 |x - c_s|^2 over this rank's shots
    ns = 7
    c = torch.arange(ns * 3, dtype=torch.float64).reshape(ns, 3) * (1 + 0.5j)
    x = torch.full((3,), 0.25 - 0.5j, dtype=torch.complex128)
    s0, s1 = S.shot_range(ns, rank, world)
    f = sum(0.5 * float(((x - c[s]).abs() ** 2).sum()) for s in range(s0, s1))
    g = [sum((x - c[s] for s in range(s0, s1)), torch.zeros(3, dtype=torch.complex128))]
    F, G = S.allreduce_value_and_grad(f, g)
    if rank == 0:
        torch.save({"f": F, "g": G[0]}, out)
    dist.destroy_process_group()


def test_sharded_objective_reduces_to_the_full_sum(tmp_path):
    import socket
    import torch.multiprocessing as mp
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    out = str(tmp_path / "r.pt")
    mp.spawn(_rank, args=(2, port, out), nprocs=2, join=True)
    res = torch.load(out, weights_only=True)
    ns = 7
    c = torch.arange(ns * 3, dtype=torch.float64).reshape(ns, 3) * (1 + 0.5j)
    x = torch.full((3,), 0.25 - 0.5j, dtype=torch.complex128)
    assert abs(res["f"] - 0.5 * float(((x - c) .abs() ** 2).sum())) < 1e-9
    assert torch.allclose(res["g"], (x - c).sum(0))
