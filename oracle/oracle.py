"""ctypes wrapper of the CPU oracle (liblbfem_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path. Imported by tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg; never by the product
package. Builds nothing itself: `make -C oracle` (or __graft_entry__.build())
produces the library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LBFEM_ORACLE_OMP=1 loads the OpenMP build (bitwise the serial results; the multi-core CPU baseline)
LIB_PATH = os.path.join(HERE, "liblbfem_oracle_omp.so" if os.environ.get("LBFEM_ORACLE_OMP") == "1"
                        else "liblbfem_oracle.so")

TGV, COUETTE, CAVITY = 0, 1, 2
LEELIN, BARDOW = 0, 1
TG2, TG3, TG3_SPLIT = 0, 1, 2
MASS_CG, MASS_LUMPED, MASS_RICHARDSON = 0, 1, 2
OK, EINVAL, ENOCONV = 0, 1, 2

Q = 9
EX = np.array([0, 1, 1, 0, -1, -1, -1, 0, 1], dtype=np.float64)
EY = np.array([0, 0, 1, 1, 1, 0, -1, -1, -1], dtype=np.float64)
W = np.array([4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36])
D2Q9_E = np.stack([EX, EY], axis=1).copy()

_lib = None
_P = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: build it with `make -C {HERE}`")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        sig = {
            "ora_disc_create": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double]),
            "ora_disc_create2": (vp, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_int]),
            "ora_disc_general_vertices": (C.c_int, [vp, _P, _P]),
            "ora_disc_destroy": (None, [vp]),
            "ora_disc_n_nodes": (C.c_int, [vp]),
            "ora_disc_n_per_dir": (C.c_int, [vp]),
            "ora_disc_n_periodic": (C.c_int, [vp]),
            "ora_disc_h_min": (C.c_double, [vp]),
            "ora_disc_time_step": (C.c_double, [vp, C.c_double]),
            "ora_disc_nodes": (None, [vp, _P]),
            "ora_disc_vertices": (None, [vp, _P, _P]),
            "ora_disc_lumped_mass": (None, [vp, _P]),
            "ora_disc_mass_diagonal": (None, [vp, _P]),
            "ora_disc_constrained": (None, [vp, _IP]),
            "ora_disc_n_walls": (C.c_int, [vp]),
            "ora_disc_walls": (None, [vp, _IP, _P]),
            "ora_case_set_reynolds": (None, [vp, C.c_double]),
            "ora_case_relaxation_time": (C.c_double, [vp]),
            "ora_case_reference_time": (C.c_double, [vp]),
            "ora_case_U0": (C.c_double, [vp]),
            "ora_case_energy_scale": (C.c_double, [vp]),
            "ora_case_initial_populations": (None, [vp, C.c_double, _P]),
            "ora_case_exact": (None, [vp, C.c_double, _P]),
            "ora_gll_nodes": (None, [C.c_int, _P]),
            "ora_element_matrices": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _P, _P, _P]),
            "ora_collide_bgk": (None, [vp, _P, C.c_double]),
            "ora_compute_equilibrium": (None, [vp, _P, _P]),
            "ora_predictor_corrector": (None, [vp, _P, _P, _P, C.c_double]),
            "ora_advection_apply": (None, [vp, _P, _P, _P, _P, C.c_double, C.c_double, C.c_int, C.c_int]),
            "ora_streaming_vmult": (None, [vp, _P, _P, C.c_double, C.c_double, C.c_double]),
            "ora_mass_solve": (C.c_int, [vp, _P, _P, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                         C.c_double, C.c_double, _IP]),
            "ora_scheme_create": (vp, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int]),
            "ora_scheme_destroy": (None, [vp]),
            "ora_scheme_set_time_step": (None, [vp, C.c_double, C.c_double]),
            "ora_scheme_set_initial_populations": (None, [vp]),
            "ora_scheme_set_populations": (None, [vp, _P]),
            "ora_scheme_get_populations": (None, [vp, _P]),
            "ora_scheme_step": (C.c_int, [vp, C.c_int]),
            "ora_scheme_cg_iterations": (C.c_longlong, [vp]),
            "ora_scheme_last_cg_iterations": (None, [vp, _IP]),
            "ora_scheme_get_increments": (None, [vp, C.c_int, _P]),
            "ora_diagnostics": (None, [vp, _P, C.c_double, _P, _P]),
            "ora_stream_function": (C.c_int, [vp, _P, _P, _P, C.POINTER(C.c_int)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def gll_nodes(p):
    x = np.zeros(p + 1)
    lib().ora_gll_nodes(p, _p(x))
    return x


def element_matrices(p, h, e):
    nd = (p + 1) ** 2
    M, Cm, K = (np.zeros((nd, nd)) for _ in range(3))
    lib().ora_element_matrices(p, h, float(e[0]), float(e[1]), _p(M), _p(Cm), _p(K))
    return M, Cm, K


class Disc:
    """Discretization + TestCase of the oracle (discretization.h, test_cases.h)."""

    def __init__(self, n_cells, degree=1, case=TGV, mach=0.1, modes=(1, 4), stretch=0.0, reynolds=None,
                 distort=0.0, general=False):
        self.h = lib().ora_disc_create2(n_cells, degree, case, mach, modes[0], modes[1], stretch, distort,
                                        int(general))
        if not self.h:
            raise ValueError("invalid oracle discretization")
        self.N, self.p, self.case = n_cells, degree, case
        self.n = lib().ora_disc_n_nodes(self.h)
        self.np1 = lib().ora_disc_n_per_dir(self.h)
        if reynolds is not None:
            self.set_reynolds(reynolds)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_disc_destroy(self.h)
            self.h = None

    # --- geometry
    @property
    def h_min(self):
        return lib().ora_disc_h_min(self.h)

    @property
    def n_periodic(self):
        return lib().ora_disc_n_periodic(self.h)

    def time_step(self, cfl):
        return lib().ora_disc_time_step(self.h, cfl)

    def nodes(self):
        xy = np.zeros((self.n, 2))
        lib().ora_disc_nodes(self.h, _p(xy))
        return xy

    def vertices(self):
        x = np.zeros(self.N + 1)
        y = np.zeros(self.N + 1)
        lib().ora_disc_vertices(self.h, _p(x), _p(y))
        return x, y

    def general_vertices(self):
        """(N+1)^2 vertex coordinates (x fastest) of a general mesh, or None for a tensor mesh."""
        x = np.zeros((self.N + 1) ** 2)
        y = np.zeros((self.N + 1) ** 2)
        return (x, y) if lib().ora_disc_general_vertices(self.h, _p(x), _p(y)) else None

    def lumped_mass(self):
        v = np.zeros(self.n)
        lib().ora_disc_lumped_mass(self.h, _p(v))
        return v

    def mass_diagonal(self):
        v = np.zeros(self.n)
        lib().ora_disc_mass_diagonal(self.h, _p(v))
        return v

    def constrained(self):
        v = np.zeros(self.n, dtype=np.int32)
        lib().ora_disc_constrained(self.h, v.ctypes.data_as(_IP))
        return v.astype(bool)

    def walls(self):
        k = lib().ora_disc_n_walls(self.h)
        nodes = np.zeros(k, dtype=np.int32)
        vel = np.zeros((k, 2))
        if k:
            lib().ora_disc_walls(self.h, nodes.ctypes.data_as(_IP), _p(vel))
        return nodes, vel

    # --- test case
    def set_reynolds(self, Re):
        lib().ora_case_set_reynolds(self.h, Re)

    @property
    def relaxation_time(self):
        return lib().ora_case_relaxation_time(self.h)

    @property
    def reference_time(self):
        return lib().ora_case_reference_time(self.h)

    @property
    def U0(self):
        return lib().ora_case_U0(self.h)

    @property
    def energy_scale(self):
        return lib().ora_case_energy_scale(self.h)

    def initial_populations(self, lam):
        f = np.zeros((Q, self.n))
        lib().ora_case_initial_populations(self.h, lam, _p(f))
        return f

    def exact(self, t):
        v = np.zeros((self.n, 3))
        lib().ora_case_exact(self.h, t, _p(v))
        return v

    # --- operators
    def collide_bgk(self, g, omega):
        g = np.ascontiguousarray(g, dtype=np.float64).copy()
        lib().ora_collide_bgk(self.h, _p(g), omega)
        return g

    def compute_equilibrium(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        feq = np.zeros_like(f)
        lib().ora_compute_equilibrium(self.h, _p(f), _p(feq))
        return feq

    def predictor_corrector(self, f, feq, incr, theta):
        f = np.ascontiguousarray(f, dtype=np.float64).copy()
        lib().ora_predictor_corrector(self.h, _p(f), _p(np.ascontiguousarray(feq)), _p(np.ascontiguousarray(incr)),
                                      theta)
        return f

    def advection(self, f, dt, lam, feq=None, e=D2Q9_E, leelin=False, remove_mass_flux=True):
        f = np.ascontiguousarray(f, dtype=np.float64)
        rhs = np.zeros((8, self.n))
        e = np.ascontiguousarray(e, dtype=np.float64)
        fe = _p(np.ascontiguousarray(feq)) if feq is not None else None
        lib().ora_advection_apply(self.h, _p(rhs), _p(f), fe, _p(e), dt, lam, int(leelin), int(remove_mass_flux))
        return rhs

    def vmult(self, x, c=0.0, e=(0.0, 0.0)):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        lib().ora_streaming_vmult(self.h, _p(y), _p(x), c, float(e[0]), float(e[1]))
        return y

    def mass_solve(self, r, x0=None, mass=MASS_CG, richardson=2, cg_tol=1e-8, cg_max=200, tg3=0.0, e=(0.0, 0.0)):
        r = np.ascontiguousarray(r, dtype=np.float64)
        x = np.zeros(self.n) if x0 is None else np.ascontiguousarray(x0, dtype=np.float64).copy()
        its = C.c_int(0)
        st = lib().ora_mass_solve(self.h, _p(x), _p(r), mass, richardson, cg_tol, cg_max, tg3, float(e[0]),
                                  float(e[1]), C.byref(its))
        return x, its.value, st

    def diagnostics(self, f, t, u_prev=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        up = np.zeros((self.n, 2)) if u_prev is None else np.ascontiguousarray(u_prev, dtype=np.float64)
        s = np.zeros(6)
        lib().ora_diagnostics(self.h, _p(f), t, _p(up), _p(s))
        return s

    def stream_function(self, u, v):
        """psi of the nodal velocity (u, v) on a walled box (include/lbfem/io.h:31-93);
        returns (psi, CG iterations, status)."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        psi = np.zeros(self.n)
        its = C.c_int(0)
        st = lib().ora_stream_function(self.h, _p(u), _p(v), _p(psi), C.byref(its))
        return psi, its.value, st


class Scheme:
    """Bardow / LeeLin stepper of the oracle (schemes.h, streaming.h)."""

    def __init__(self, disc, scheme=BARDOW, streaming=TG2, mass=MASS_CG, richardson=2, cg_tol=1e-8, cg_max=200):
        self.disc = disc
        self.h = lib().ora_scheme_create(disc.h, scheme, streaming, mass, richardson, cg_tol, cg_max)
        if not self.h:
            raise ValueError("invalid scheme settings")

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_scheme_destroy(self.h)
            self.h = None

    def set_time_step(self, dt, lam):
        lib().ora_scheme_set_time_step(self.h, dt, lam)

    def set_initial_populations(self):
        lib().ora_scheme_set_initial_populations(self.h)

    def set_populations(self, f):
        lib().ora_scheme_set_populations(self.h, _p(np.ascontiguousarray(f, dtype=np.float64)))

    def populations(self):
        f = np.zeros((Q, self.disc.n))
        lib().ora_scheme_get_populations(self.h, _p(f))
        return f

    def increments(self, which=0):
        x = np.zeros((8, self.disc.n))
        lib().ora_scheme_get_increments(self.h, which, _p(x))
        return x

    def step(self, n=1):
        st = lib().ora_scheme_step(self.h, n)
        if st != OK:
            raise RuntimeError(f"oracle step failed with status {st}")

    @property
    def cg_iterations(self):
        return lib().ora_scheme_cg_iterations(self.h)

    def last_cg_iterations(self):
        """The last step's CG steps per solve, shape (2 sweeps, 8 populations); 0 where not solved."""
        v = np.zeros(16, dtype=np.int32)
        lib().ora_scheme_last_cg_iterations(self.h, v.ctypes.data_as(_IP))
        return v.reshape(2, 8)


def moments(f):
    rho = f.sum(axis=0)
    return rho, (EX @ f) / rho, (EY @ f) / rho


def equilibrium(rho, ux, uy):
    uu = ux * ux + uy * uy
    eu = EX[:, None] * ux + EY[:, None] * uy
    return W[:, None] * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * uu)
