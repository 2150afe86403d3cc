"""The continuous-batching server: requests that arrive while others run join the
running batch (lbfem_batch_add) and leave it when they finish, reach a steady
state or diverge (lbfem_batch_remove); every result is bitwise the one the same
request gets stepped alone under the same retirement rules."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402
import oracle as O  # noqa: E402
from lbfem_amd import serving, workloads  # noqa: E402


def bitwise(a, b):
    return a.shape == b.shape and np.array_equal(np.ascontiguousarray(a).view(np.uint64),
                                                 np.ascontiguousarray(b).view(np.uint64))


def check_alone(done, reqs, precision, window):
    by_id = {d["id"]: d for d in done}
    for r in reqs:
        s, steps, status = serving.run_alone(A, r, precision, window=window)
        d = by_id[r["id"]]
        assert d["status"] == status, (r, d["status"], status)
        if status == "diverged":  # found at the first check after it happened, which depends on the batch
            continue
        assert d["steps"] == steps, (r, d["steps"], steps)
        assert bitwise(d["f"], s.f), r


@pytest.mark.parametrize("precision", [A.MIXED, A.FP64])
def test_served_results_equal_alone(precision):
    reqs = workloads.request_stream(7, 10, rate=400.0, sizes=(16, 24, 40), steps=(200, 900), re=(100.0, 1000.0),
                                    steady=0.5, tol=1e-3)
    srv = serving.Server(A, precision=precision, max_slots=4, tick_steps=150, window=100, keep_results=True)
    done, summary = srv.run(reqs)
    assert summary["requests"] == len(reqs)
    assert summary["ticks"] > 1
    # the slot limit made later requests wait for earlier ones to finish, and the
    # batch took in requests while others were part way through their steps
    assert any(b["admit"] >= a["done"] for a in done for b in done)
    assert any(a["admit"] < b["admit"] < a["done"] for a in done for b in done)
    assert {"steps", "steady", "not steady"} & set(summary["statuses"])
    check_alone(done, reqs, precision, 100)


def test_steady_and_diverged_retirement():
    """A Couette start-up reaches the steady tolerance and retires before its cap; a fine cavity at small
    dt / lambda diverges and retires with that status; a fixed-step request runs its steps."""
    base = {"arrival": 0.0, "U": 0.08, "mach": 0.08 * 3 ** 0.5, "p": 1}
    reqs = [dict(base, id=0, case="couette", N=16, Re=20.0, steps=60000, mode="steady", tol=1e-9),
            dict(base, id=1, case="cavity", N=512, Re=60.0, steps=6000, mode="steps"),
            dict(base, id=2, case="cavity", N=32, Re=400.0, steps=3000, mode="steps")]
    srv = serving.Server(A, precision=A.FP64, keep_results=True)
    done, summary = srv.run(reqs)
    st = {d["id"]: d for d in done}
    assert st[0]["status"] == "steady" and st[0]["steps"] < 60000, st[0]
    assert st[1]["status"] == "diverged", st[1]
    assert st[2]["status"] == "steps" and st[2]["steps"] == 3000
    check_alone(done, reqs, A.FP64, 1000)


def test_batch_join_and_leave_bitwise():
    """Members join a running batch at different steps and leave it; each one's populations are bitwise
    those of the same scheme stepped alone for the same steps."""
    specs = [("cavity", 24, 1, A.MASS_LUMPED, A.FP64), ("tgv", 10, 2, A.MASS_LUMPED, A.MIXED),
             ("couette", 16, 1, A.MASS_LUMPED, A.FP64), ("cavity", 12, 2, A.MASS_LUMPED, A.MIXED)]

    def make(case, N, p, mass, prec):
        tc = A.test_cases.make(case, mach=0.1, stretch=1.2)
        tc.set_reynolds(300.0)
        d = A.Discretization(tc, N, p)
        s = A.Bardow(d, mass=mass, precision=prec)
        s.set_time_step(d.time_step(0.4), tc.relaxation_time())
        s.set_initial_populations(tc)
        return s

    for grid in (0, 64):
        b = A.Batch([], graph_steps=8)
        b.set_shape(-1, grid)
        batched = [make(*x) for x in specs]
        steps = [0] * len(specs)
        plan = [(0, "add"), (1, "add"), (None, 21), (2, "add"), (None, 17), (0, "remove"), (3, "add"), (None, 30),
                (1, "remove"), (None, 9), (2, "remove"), (3, "remove")]
        inside = set()
        for who, what in plan:
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
        for k, x in enumerate(specs):
            alone = make(*x)
            alone.step(steps[k])
            assert steps[k] > 0
            assert bitwise(batched[k].f, alone.f), (grid, x, steps[k])


@pytest.mark.parametrize("precision", [A.FP64, A.MIXED])
def test_velocity_change_matches_the_oracle_diagnostics(precision):
    """The device monitor's mean change is cgdbe's steady-state rate numerator over its area
    (apps/cgdbe/cgdbe.cc:385-393), as the oracle's diagnostics compute it from the same populations; the
    largest change is checked against the same nodal velocities; the batch form equals the scheme form."""
    tc = A.test_cases.make("cavity", mach=0.1, stretch=1.2)
    tc.set_reynolds(400.0)
    d = A.Discretization(tc, 20, 2)
    od = O.Disc(20, 2, O.CAVITY, mach=0.1, stretch=1.2, reynolds=400.0)
    s = A.Bardow(d, mass=A.MASS_LUMPED, precision=precision)
    s.set_time_step(od.time_step(0.4), od.relaxation_time)
    s.set_initial_populations()
    s.step(5)
    mx0, mean0, bad0 = s.velocity_change()
    assert np.isinf(mx0) and bad0 == 0
    f0 = s.f
    s.step(7)
    mx, mean, bad = s.velocity_change()
    f1 = s.f
    _, u0, v0 = A.moments(f0)
    _, u1, v1 = A.moments(f1)
    sums = od.diagnostics(f1, 0.0, u_prev=np.stack([u0, v0], 1))
    assert bad == 0
    assert abs(mean - sums[5] / sums[4]) <= 1e-12 * abs(mean)
    indep = ~d.constrained()
    want = np.hypot(u1 - u0, v1 - v0)[indep].max()
    assert abs(mx - want) <= 1e-12 * want
    # the batch form, on a copy stepped the same way
    s2 = A.Bardow(d, mass=A.MASS_LUMPED, precision=precision)
    s2.set_time_step(od.time_step(0.4), od.relaxation_time)
    s2.set_initial_populations()
    s2.step(5)
    b = A.Batch([s2])
    b.velocity_change()
    b.step(7)
    bm, bmean, bbad = b.velocity_change()
    b.close()
    assert bm[0] == mx and bmean[0] == mean and bbad[0] == 0


def test_admission_budget_limits_concurrency():
    reqs = workloads.request_stream(11, 6, rate=1000.0, sizes=(32,), steps=(100, 300), q2=0.0, steady=0.0)
    one = serving.request_bytes(reqs[0])
    srv = serving.Server(A, budget_bytes=1.5 * one, max_slots=96, tick_steps=100)
    done, summary = srv.run(reqs)
    assert summary["requests"] == 6
    # with room for one request at a time, each admission follows the previous completion
    d = sorted(done, key=lambda x: x["admit"])
    for a, b in zip(d, d[1:]):
        assert b["admit"] >= a["done"] - 1e-9
