"""CPU checks of the product boundary: the C-ABI library loads and exports
every symbol include/lbfem_hip.h declares; no compute calls (no GPU here)."""
import ctypes
import os
import re

import pytest

import lbfem_amd as A

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "lbfem_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbfem_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_reference_operator_api():
    names = declared()
    for n in ("lbfem_advection_apply", "lbfem_mass_solve", "lbfem_collide_bgk", "lbfem_compute_equilibrium",
              "lbfem_predictor_corrector", "lbfem_scheme_step", "lbfem_scheme_set_time_step",
              "lbfem_scheme_set_initial_populations", "lbfem_disc_create"):
        assert n in names


@pytest.mark.skipif(not os.path.exists(A.LIB_PATH), reason="liblbfem_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(A.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the binding's signature table covers the header too
    A.lib()


@pytest.mark.skipif(not os.path.exists(A.LIB_PATH), reason="liblbfem_hip.so not built")
def test_error_reporting_without_device_work():
    assert A.lib().lbfem_disc_create(None, None) == A.EINVAL
    assert "null" in A.last_error()


@pytest.mark.skipif(not os.path.exists(A.LIB_PATH), reason="liblbfem_hip.so not built")
def test_distortion_is_the_periodic_vortex_mesh_only():
    """--distort maps the vortex's periodic box (apps/cgdbe/cgdbe.cc:77, test_cases.h:53-65): a walled or
    stretched box with a distortion is rejected before any device work."""
    import ctypes as C
    h = C.c_void_p()
    for n_periodic, stretch in ((0, 0.0), (1, 0.0), (2, 1.2)):
        desc = A.MeshDesc(8, 1, n_periodic, 1.0, stretch, 0.02)
        assert A.lib().lbfem_disc_create(C.byref(desc), C.byref(h)) == A.EINVAL
        assert "distortion" in A.last_error()
    desc = A.MeshDesc(8, 1, 2, 1.0, 0.0, float("nan"))
    assert A.lib().lbfem_disc_create(C.byref(desc), C.byref(h)) == A.EINVAL
