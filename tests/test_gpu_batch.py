"""Batched stepping of uneven simulations (flow, mesh, degree, scheme, FP64 or
mixed): every member's populations are bitwise those of the same scheme stepped
alone, with and without captured graphs, before and after the batch autotuner
picks its shapes."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402

MEMBERS = [
    ("cavity", 24, 1, A.LEELIN, A.FP64, 400.0),
    ("tgv", 10, 2, A.BARDOW, A.MIXED, 100.0),
    ("couette", 16, 1, A.BARDOW, A.FP64, 200.0),
    ("cavity", 12, 2, A.LEELIN, A.MIXED, 1000.0),
    ("tgv", 20, 1, A.LEELIN, A.FP64, 50.0),
    ("cavity", 6, 3, A.BARDOW, A.FP64, 100.0),
    ("cavity", 40, 1, A.LEELIN, A.FP64, 3200.0),
    ("couette", 9, 2, A.LEELIN, A.MIXED, 500.0),
]


def make(case, N, p, kind, prec, Re, mach=0.1):
    tc = A.test_cases.make(case, mach=mach, modes=(1, 1), stretch=1.2)
    tc.set_reynolds(Re)
    d = A.Discretization(tc, N, p)
    s = (A.Bardow if kind == A.BARDOW else A.LeeLin)(d, mass=A.MASS_LUMPED, precision=prec)
    s.set_time_step(d.time_step(0.4), tc.relaxation_time())
    s.set_initial_populations(tc)
    return s


@pytest.mark.parametrize("graph_steps", [0, 8])
def test_batch_matches_alone(graph_steps):
    batched = [make(*m) for m in MEMBERS]
    alone = [make(*m) for m in MEMBERS]
    b = A.Batch(batched, graph_steps=graph_steps)
    assert b.info()["groups"] == 5  # (degree, scheme, precision) pairs of MEMBERS
    b.step(21)
    b.synchronize()
    for s in alone:
        s.step(21)
    for m, s1, s2 in zip(MEMBERS, batched, alone):
        assert np.array_equal(s1.f, s2.f), m
    # shapes from the autotuner, then more steps (odd counts cross the graph's parity)
    b.autotune()
    b.step(19)
    b.synchronize()
    for s in alone:
        s.step(19)
    for m, s1, s2 in zip(MEMBERS, batched, alone):
        assert np.array_equal(s1.f, s2.f), m
    b.close()


@pytest.mark.parametrize("variant,grid", [(A.TILE_WIDE, 0), (A.TILE_SPLIT4, 64), (A.TILE_TALL, 8)])
def test_batch_shapes(variant, grid):
    batched = [make(*m) for m in MEMBERS[:5]]
    alone = [make(*m) for m in MEMBERS[:5]]
    b = A.Batch(batched)
    b.set_shape(variant, grid)
    b.step(10)
    b.synchronize()
    for s in alone:
        s.step(10)
    for m, s1, s2 in zip(MEMBERS, batched, alone):
        assert np.array_equal(s1.f, s2.f), m
    b.close()


def test_batch_rejects_consistent_mass():
    tc = A.test_cases.make("tgv", mach=0.1)
    tc.set_reynolds(100.0)
    d = A.Discretization(tc, 8, 1)
    s = A.Bardow(d, mass=A.MASS_CG)
    s.set_time_step(d.time_step(0.25), tc.relaxation_time())
    with pytest.raises(A.LbfemError):
        A.Batch([s])
