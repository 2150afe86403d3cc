"""Batched stepping of uneven simulations (flow, mesh, degree, scheme, FP64 or
mixed, lumped or consistent mass): every member's populations (and CG iteration
counts) are bitwise those of the same scheme stepped alone, with and without
captured graphs, before and after the batch autotuner picks its shapes."""
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


# consistent-mass members: (flow, N, p, scheme, streaming, mass, Re)
CG_MEMBERS = [
    ("cavity", 24, 1, A.BARDOW, A.TG2, A.MASS_CG, 400.0),
    ("tgv", 12, 2, A.LEELIN, A.TG2, A.MASS_CG, 100.0),
    ("couette", 16, 1, A.BARDOW, A.TG3, A.MASS_CG, 200.0),
    ("tgv", 16, 1, A.BARDOW, A.TG3_SPLIT, A.MASS_CG, 100.0),
    ("cavity", 8, 3, A.LEELIN, A.TG2, A.MASS_CG, 1000.0),
    ("cavity", 40, 1, A.BARDOW, A.TG2, A.MASS_CG, 1000.0),
    ("tgv", 20, 1, A.BARDOW, A.TG2, A.MASS_LUMPED, 100.0),
    ("couette", 9, 2, A.LEELIN, A.TG2, A.MASS_CG, 500.0),
]


def make_cg(case, N, p, kind, streaming, mass, Re, cfl=None):
    tc = A.test_cases.make(case, mach=0.1, modes=(1, 1), stretch=1.2)
    tc.set_reynolds(Re)
    d = A.Discretization(tc, N, p)
    s = (A.Bardow if kind == A.BARDOW else A.LeeLin)(d, streaming=streaming, mass=mass, cg_tolerance=1e-10)
    s.set_time_step(d.time_step(cfl or (0.9 if streaming != A.TG2 else 0.3)), tc.relaxation_time())
    s.set_initial_populations(tc)
    return s


def bitwise(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def test_batch_consistent_mass_matches_alone():
    """Consistent-mass members (Bardow and Lee-Lin, TG2, TG3, split TG3, Q1-Q3, walls and periodic) stepped
    in one batch with a lumped member: populations bitwise those stepped alone, and every solve's CG
    iterations equal, step by step."""
    batched = [make_cg(*m) for m in CG_MEMBERS]
    alone = [make_cg(*m) for m in CG_MEMBERS]
    b = A.Batch(batched)
    assert b.info()["groups"] == 6  # five CG (degree, scheme, streaming) groups and one lumped group
    for step in range(6):
        b.step(1)
        b.synchronize()
        for m, s1, s2 in zip(CG_MEMBERS, batched, alone):
            s2.step(1)
            assert np.array_equal(s1.last_cg_iterations(), s2.last_cg_iterations()), (step, m)
    b.step(9)
    b.synchronize()
    for s in alone:
        s.step(9)
    for m, s1, s2 in zip(CG_MEMBERS, batched, alone):
        assert s1.cg_iterations == s2.cg_iterations, m
        assert bitwise(s1.f, s2.f), m
    assert (b.member_status() == A.OK).all()
    b.close()


def test_batch_consistent_mass_join_and_leave():
    """CG members join a running batch at different steps (their buffers at different parities) and leave
    it; each one is bitwise the scheme stepped alone for the same steps."""
    specs = CG_MEMBERS[:4]
    batched = [make_cg(*m) for m in specs]
    # unequal parities at the start: one member has stepped alone first
    batched[1].step(1)
    steps = [0, 1, 0, 0]
    b = A.Batch([batched[0]])
    inside = {0}
    for who, what in [(None, 3), (1, "add"), (None, 4), (2, "add"), (0, "remove"), (None, 5), (3, "add"),
                      (None, 2), (1, "remove"), (None, 3)]:
        if who is None:
            b.step(what)
            for k in inside:
                steps[k] += what
        elif what == "add":
            b.add(batched[who])
            inside.add(who)
        else:
            b.remove(batched[who])
            inside.discard(who)
    b.close()
    for k, m in enumerate(specs):
        s = make_cg(*m)
        s.step(steps[k])
        assert bitwise(batched[k].f, s.f), (m, steps[k])
        assert batched[k].cg_iterations == s.cg_iterations, m


def test_batch_noconvergence_names_the_member():
    ok = make_cg("cavity", 12, 1, A.BARDOW, A.TG2, A.MASS_CG, 400.0)
    tc = A.test_cases.make("tgv", mach=0.1)
    tc.set_reynolds(100.0)
    d = A.Discretization(tc, 12, 1)
    capped = A.Bardow(d, mass=A.MASS_CG, cg_tolerance=1e-30, cg_max_iterations=3)
    capped.set_time_step(d.time_step(0.25), tc.relaxation_time())
    capped.set_initial_populations(tc)
    b = A.Batch([ok, capped])
    with pytest.raises(A.NoConvergence):
        b.step(1)
    assert list(b.member_status()) == [A.OK, A.ENOCONV]
    b.close()


def test_batch_rejects_mixed_consistent_mass():
    tc = A.test_cases.make("tgv", mach=0.1)
    tc.set_reynolds(100.0)
    d = A.Discretization(tc, 8, 1)
    s = A.Bardow(d, mass=A.MASS_CG, precision=A.MIXED)
    s.set_time_step(d.time_step(0.25), tc.relaxation_time())
    with pytest.raises(A.LbfemError):
        A.Batch([s])
