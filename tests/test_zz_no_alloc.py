"""lbfem_mass_solve_noalloc (the mass solve on the library's fixed per-GPU work space, no
allocation) against lbfem_mass_solve, through the C ABI: the same increment and the same CG
iteration count, bit for bit, on two calls in a row with the second on a smaller mesh (the
shared work space keeps the first call's contents), for CG with and without the TG3 term,
Richardson and the lumped mass; a mesh beyond the work space's size is refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402


def _disc(case, N, p):
    tc = A.test_cases.make(case, mach=0.1, stretch=1.2)
    tc.set_reynolds(400.0)
    return A.Discretization(tc, N, p)


def _bits(t):
    return t.detach().cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("mass,tg3", [(A.MASS_CG, 0.0), (A.MASS_CG, 2e-4), (A.MASS_RICHARDSON, 0.0),
                                      (A.MASS_LUMPED, 0.0)])
def test_noalloc_matches_mass_solve(mass, tg3):
    rng = np.random.default_rng(3)
    # two calls in a row, the second on a smaller mesh (and another degree and flow)
    for case, N, p in (("cavity", 200, 1), ("tgv", 40, 2)):
        d = _disc(case, N, p)
        r = rng.standard_normal(d.n_nodes)
        r[d.constrained()] = 0.0
        x0 = 0.1 * rng.standard_normal(d.n_nodes)  # a CG starting vector, as the schemes pass one
        x0[d.constrained()] = 0.0
        rd = d.to_device(r, 1)
        xa, xb = d.to_device(x0, 1), d.to_device(x0, 1)
        kw = dict(mass=mass, richardson=3, cg_tolerance=1e-10, e=(1.0, -1.0), tg3_coefficient=tg3)
        its_a = A.mass_solve(d, xa, rd, **kw)
        its_b = A.mass_solve_noalloc(d, xb, rd, **kw)
        torch.cuda.synchronize()
        assert its_a == its_b, (case, N, p, its_a, its_b)
        assert np.array_equal(_bits(xa), _bits(xb)), (case, N, p)
        if mass == A.MASS_CG:
            assert its_a > 0
        assert np.isfinite(d.to_host(xb, 1)).all()


def test_noalloc_refuses_a_mesh_beyond_its_work_space():
    d = _disc("tgv", 1100, 1)  # 1101^2 = 1 212 201 nodes > 1 179 648
    r = d.device_vector(1)
    x = d.device_vector(1)
    with pytest.raises(A.LbfemError) as err:
        A.mass_solve_noalloc(d, x, r, mass=A.MASS_CG)
    assert err.value.code == A.EINVAL
    assert "fixed work space" in str(err.value)
