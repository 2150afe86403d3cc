"""The tile kernel's variants (wide tiles, populations split over 2 or 4
threads, small tiles) give bitwise identical populations, and the autotuner
times them without changing the scheme's state. Bitwise: every variant computes
each output value by the same operation sequence."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402


def scheme(case, N, p, kind, mass, precision=A.FP64, streaming=A.TG2, Re=100.0):
    tc = A.test_cases.make(case, mach=0.1, modes=(1, 1), stretch=1.2)
    tc.set_reynolds(Re)
    d = A.Discretization(tc, N, p)
    s = (A.Bardow if kind == A.BARDOW else A.LeeLin)(d, streaming=streaming, mass=mass, cg_tolerance=1e-12,
                                                      precision=precision)
    s.set_time_step(d.time_step(0.25), tc.relaxation_time())
    s.set_initial_populations(tc)
    return d, s


CASES = [
    ("tgv", 20, 1, A.BARDOW, A.MASS_LUMPED, A.FP64),
    ("cavity", 37, 1, A.LEELIN, A.MASS_LUMPED, A.FP64),
    ("cavity", 13, 2, A.BARDOW, A.MASS_LUMPED, A.FP64),
    ("couette", 9, 3, A.LEELIN, A.MASS_LUMPED, A.FP64),
    ("cavity", 37, 1, A.LEELIN, A.MASS_LUMPED, A.MIXED),
    ("tgv", 11, 2, A.BARDOW, A.MASS_LUMPED, A.MIXED),
    ("cavity", 21, 1, A.LEELIN, A.MASS_CG, A.FP64),
    ("cavity", 10, 2, A.BARDOW, A.MASS_CG, A.FP64),
    # periodic meshes with Lee-Lin (the slave lines' predictor-corrector), both precisions
    ("tgv", 20, 1, A.LEELIN, A.MASS_LUMPED, A.FP64),
    ("tgv", 20, 1, A.LEELIN, A.MASS_LUMPED, A.MIXED),
    ("couette", 70, 1, A.LEELIN, A.MASS_LUMPED, A.FP64),
    ("couette", 70, 1, A.BARDOW, A.MASS_LUMPED, A.MIXED),
    # Q2 in FP32, where the row-march kernel runs (strips of 60 columns): walls on four sides, a
    # last strip of one column (61 nodes), periodic lines, and the advection mode under CG
    ("cavity", 37, 2, A.LEELIN, A.MASS_LUMPED, A.MIXED),
    ("cavity", 30, 2, A.BARDOW, A.MASS_LUMPED, A.MIXED),
    ("tgv", 31, 2, A.BARDOW, A.MASS_LUMPED, A.MIXED),
    ("tgv", 31, 2, A.LEELIN, A.MASS_LUMPED, A.MIXED),
    ("couette", 40, 2, A.BARDOW, A.MASS_LUMPED, A.MIXED),
    ("cavity", 30, 2, A.BARDOW, A.MASS_CG, A.MIXED),
]


@pytest.mark.parametrize("case,N,p,kind,mass,prec", CASES)
def test_variants_bitwise_identical(case, N, p, kind, mass, prec):
    ref = None
    for v in range(A.TILE_VARIANTS):
        _, s = scheme(case, N, p, kind, mass, prec)
        s.tile_variant = v
        assert s.tile_variant == v
        s.step(3)
        f = s.f
        if ref is None:
            ref = f
        else:
            assert np.array_equal(f, ref), (v, np.abs(f - ref).max())


def test_autotune_keeps_state_and_results():
    d, s = scheme("cavity", 40, 1, A.LEELIN, A.MASS_LUMPED)
    _, s2 = scheme("cavity", 40, 1, A.LEELIN, A.MASS_LUMPED)
    s.step(2)
    s2.step(2)
    before = s.f
    v, t = s.autotune(force=True)
    assert 0 <= v < A.TILE_VARIANTS and s.tile_variant == v
    assert (t > 0).all(), t
    assert t[v] == t.min()
    assert np.array_equal(s.f, before)
    s.step(2)
    s2.step(2)
    assert np.array_equal(s.f, s2.f)
    # the choice is cached for the same kernel and mesh
    _, s3 = scheme("cavity", 40, 1, A.LEELIN, A.MASS_LUMPED)
    v3, t3 = s3.autotune()
    assert v3 == v and (t3 < 0).all()


def test_variant_out_of_range_rejected():
    _, s = scheme("tgv", 8, 1, A.BARDOW, A.MASS_LUMPED)
    with pytest.raises(A.LbfemError):
        s.tile_variant = A.TILE_VARIANTS
