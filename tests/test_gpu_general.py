"""Parity of the general-quadrilateral path (the vortex's distorted mesh,
include/lbfem/test_cases.h:53-65, docs/method.md:216-224) with the CPU oracle:
geometry, the cell-wise advection and mass operators, the mass solves and
whole steps of every scheme, mass and streaming; and the published errors of
docs/results.md:222-233 reproduced on the GPU. FP64; <= 1e-12 relative per
operator application and step, <= 1e-10 after many steps, CG iterations
within +-1 per solve.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # collected on CPU hosts, skipped there
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402
import oracle as O  # noqa: E402


def pair(N, p, eps, Re=100.0, mach=0.1, modes=(1, 4)):
    tc = A.test_cases.TaylorGreenVortex(mach, *modes, distort=eps)
    tc.set_reynolds(Re)
    gd = A.Discretization(tc, N, p)
    od = O.Disc(N, p, O.TGV, mach=mach, modes=modes, reynolds=Re, distort=eps)
    return tc, gd, od


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("p", [1, 2, 3])
def test_general_geometry_matches_oracle(p):
    _, gd, od = pair(9, p, 0.04)
    assert np.abs(gd.node - od.nodes()).max() < 1e-15
    assert (gd.constrained() == od.constrained()).all()
    assert np.allclose(gd.lumped_mass(), od.lumped_mass(), rtol=1e-14, atol=0)
    assert np.allclose(gd.mass_diagonal(), od.mass_diagonal(), rtol=1e-14, atol=0)
    assert len(gd.walls[0]) == 0


@pytest.mark.parametrize("p,N", [(1, 40), (1, 7), (2, 17), (3, 11)])
@pytest.mark.parametrize("kind", [A.BARDOW, A.LEELIN])
def test_general_advection_operator(p, N, kind):
    """N covers one partial tile, several tiles and ragged edge tiles of the cell kernel."""
    _, gd, od = pair(N, p, 0.04)
    rng = np.random.default_rng(4 + p)
    f = O.W[:, None] * (1 + 0.05 * rng.standard_normal((9, od.n)))
    feq = od.compute_equilibrium(f)
    dt, lam = 0.01, 0.003
    ll = kind == A.LEELIN
    want = od.advection(f, dt, lam, feq=feq if ll else None, leelin=ll, remove_mass_flux=not ll)
    rhs = gd.device_vector(8)
    A.advection_apply(gd, kind, rhs, gd.to_device(f, 9), gd.to_device(feq, 9), dt, lam,
                      remove_mass_flux=not ll)
    torch.cuda.synchronize()
    got = gd.to_host(rhs, 8)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
    # a split sweep (axis directions only): the populations at rest in it get 0
    e = O.D2Q9_E.copy()
    e[:, 1] = 0.0
    want = od.advection(f, dt, lam, e=e, remove_mass_flux=True)
    A.advection_apply(gd, A.BARDOW, rhs, gd.to_device(f, 9), None, dt, lam, e=e)
    torch.cuda.synchronize()
    got = gd.to_host(rhs, 8)
    assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("p,N", [(1, 33), (2, 10), (3, 7)])
def test_general_mass_operator_and_solves(p, N):
    _, gd, od = pair(N, p, 0.03)
    rng = np.random.default_rng(12)
    r = rng.standard_normal(od.n)
    r[od.constrained()] = 0
    xd, yd = gd.to_device(r, 1), gd.device_vector(1)
    for c, e in ((0.0, (0.0, 0.0)), (0.02, (1.0, -1.0)), (0.01, (0.0, 1.0))):
        A.streaming_vmult(gd, yd, xd, c, e)
        torch.cuda.synchronize()
        want = od.vmult(r, c, e)
        assert np.abs(gd.to_host(yd, 1)[0] - want).max() < 1e-14 * np.abs(want).max()
    for mass in (A.MASS_LUMPED, A.MASS_RICHARDSON, A.MASS_CG):
        xo, its_o, _ = od.mass_solve(r, mass=mass, cg_tol=1e-12, richardson=3, tg3=0.01, e=(1.0, 1.0))
        xg = gd.device_vector(1)
        its = A.mass_solve(gd, xg, gd.to_device(r, 1), mass=mass, cg_tolerance=1e-12, richardson=3,
                           tg3_coefficient=0.01, e=(1.0, 1.0))
        got = gd.to_host(xg, 1)[0]
        assert np.abs(got - xo).max() < 1e-11 * np.abs(xo).max(), mass
        assert abs(its - its_o) <= 1


def run_both(N, p, eps, scheme, mass, streaming, steps, cfl=0.25, cg_tol=1e-12, richardson=2):
    _, gd, od = pair(N, p, eps)
    dt, lam = od.time_step(cfl), od.relaxation_time
    cls = A.Bardow if scheme == A.BARDOW else A.LeeLin
    gs = cls(gd, streaming=streaming, mass=mass, cg_tolerance=cg_tol, richardson=richardson)
    gs.set_time_step(dt, lam)
    gs.set_initial_populations()
    osch = O.Scheme(od, O.BARDOW if scheme == A.BARDOW else O.LEELIN, streaming, mass, richardson=richardson,
                    cg_tol=cg_tol)
    osch.set_time_step(dt, lam)
    osch.set_initial_populations()
    assert rel(gs.f, osch.populations()) < 1e-14, "initial populations"
    errs = []
    for _ in range(steps):
        gs.step(1)
        osch.step(1)
        errs.append(rel(gs.f, osch.populations()))
    return gs, osch, errs


@pytest.mark.parametrize("scheme,mass,streaming,p", [
    (A.BARDOW, A.MASS_LUMPED, A.TG2, 1), (A.BARDOW, A.MASS_LUMPED, A.TG2, 2), (A.BARDOW, A.MASS_LUMPED, A.TG2, 3),
    (A.LEELIN, A.MASS_LUMPED, A.TG2, 1), (A.LEELIN, A.MASS_LUMPED, A.TG2, 3),
    (A.BARDOW, A.MASS_CG, A.TG2, 1), (A.BARDOW, A.MASS_CG, A.TG2, 2), (A.LEELIN, A.MASS_CG, A.TG2, 1),
    (A.BARDOW, A.MASS_CG, A.TG3, 1), (A.BARDOW, A.MASS_CG, A.TG3_SPLIT, 1), (A.BARDOW, A.MASS_CG, A.TG3_SPLIT, 2),
    (A.BARDOW, A.MASS_RICHARDSON, A.TG2, 1),
])
def test_general_steps(scheme, mass, streaming, p):
    N = {1: 16, 2: 9, 3: 6}[p]
    steps = 4
    gs, osch, errs = run_both(N, p, 0.04, scheme, mass, streaming, steps,
                              cfl=0.5 if streaming != A.TG2 else 0.25)
    assert errs[0] < 1e-12, errs
    assert max(errs) < 1e-11, errs
    if mass == A.MASS_CG:
        sweeps = 2 if streaming == A.TG3_SPLIT else 1
        assert abs(gs.cg_iterations - osch.cg_iterations) <= 8 * sweeps * steps, (gs.cg_iterations,
                                                                                 osch.cg_iterations)


def test_general_many_steps():
    """<= 1e-10 after 300 steps (Bardow, consistent mass, TG2, CFL 0.25 on the +-25 % mesh)."""
    gs, osch, errs = run_both(24, 1, 0.04, A.BARDOW, A.MASS_CG, A.TG2, 0)
    gs.step(300)
    osch.step(300)
    assert rel(gs.f, osch.populations()) < 1e-10


@pytest.mark.parametrize("streaming,cfl,eps,published", [
    (A.TG2, 0.25, 0.04, 2.07e-2),
    (A.TG3_SPLIT, 1.0, 0.02, 1.23e-2),
    (A.TG3, 1.0, 0.04, 3.38e-2),
])
def test_general_published_distortion_errors(streaming, cfl, eps, published):
    """docs/results.md:226-233 on the GPU path: 48^2 Q1, modes (1, 4), Re 100, Ma 0.05, t = 0.25 t_ref,
    Bardow, consistent mass (CG 1e-8), n = ceil(T / dt) (cgdbe.cc:237-241); three published digits."""
    tc, gd, od = pair(48, 1, eps, mach=0.05)
    T = 0.25 * tc.reference_time()
    n = math.ceil(T / gd.time_step(cfl))
    gs = A.Bardow(gd, streaming=streaming, mass=A.MASS_CG, cg_tolerance=1e-8)
    gs.set_time_step(T / n, tc.relaxation_time())
    gs.set_initial_populations()
    gs.step(n)
    sums = od.diagnostics(gs.f, T)
    err = math.sqrt(sums[0] / sums[1])
    assert abs(err - published) <= 0.5 * 10 ** (math.floor(math.log10(published)) - 2) + 1e-12, err


def test_general_rejections():
    _, gd, _ = pair(8, 1, 0.02)
    with pytest.raises(A.LbfemError):
        A.Bardow(gd, mass=A.MASS_LUMPED, precision=A.MIXED)
    s = A.Bardow(gd, mass=A.MASS_LUMPED)
    s.set_time_step(gd.time_step(0.25), 0.01)
    with pytest.raises(A.LbfemError):
        A.Batch([s])
