"""lbfem_amd: Python mirror of the reference's operator and scheme API
(include/lbfem/*.h of ivan-pi/lbm-fem) over the C ABI of liblbfem_hip.so
(include/lbfem_hip.h), for tests, benchmarks and the serving loop.

The library is built in-tree (lbm-fem_amd/build.py); importing this package
does not build it, and every compute call raises if the library or a GPU is
missing — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import test_cases
from .test_cases import (CouetteFlow, LidDrivenCavity, TaylorGreenVortex, TestCase, equilibrium, moments)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "liblbfem_hip.so")

OK, EINVAL, ENOCONV, EHIP = 0, 1, 2, 3
TILE_WIDE, TILE_TALL, TILE_SPLIT2, TILE_SPLIT4, TILE_SMALL4, TILE_WIDE40, TILE_WIDE48, TILE_SPLIT2_W40 = range(8)
MARCH3, MARCH4, MARCH8, MARCH16 = range(8, 12)  # the row-march kernel (Q1; Q2 mixed), rows per segment
TILE_VARIANTS = 12
TILE_LDS_VARIANTS = 8  # the shapes a batch runs
LEELIN, BARDOW = 0, 1
TG2, TG3, TG3_SPLIT = 0, 1, 2
MASS_CG, MASS_LUMPED, MASS_RICHARDSON = 0, 1, 2
FP64, MIXED = 0, 1
Q = 9
D2Q9_E = np.stack([test_cases.EX, test_cases.EY], axis=1).copy()


class LbfemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NoConvergence(LbfemError):
    """SolverControl::NoConvergence (include/lbfem/mass.h:38-43)."""


class MeshDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int), ("degree", C.c_int), ("n_periodic", C.c_int), ("L", C.c_double),
                ("stretch", C.c_double), ("distort", C.c_double)]


class DiscInfo(C.Structure):
    _fields_ = [("n_nodes", C.c_int64), ("n_independent", C.c_int64), ("stride", C.c_int64),
                ("n_per_dir", C.c_int32), ("n_walls", C.c_int32), ("h_min", C.c_double)]


class MassSettings(C.Structure):
    """include/lbfem/mass.h:33-48."""
    _fields_ = [("type", C.c_int), ("richardson", C.c_int), ("cg_tolerance", C.c_double),
                ("cg_max_iterations", C.c_int), ("fused", C.c_int)]


class SchemeSettings(C.Structure):
    _fields_ = [("scheme", C.c_int), ("streaming", C.c_int), ("mass", MassSettings), ("precision", C.c_int)]


_lib = None
_P = C.POINTER(C.c_double)
_vp = C.c_void_p


def lib():
    """The loaded liblbfem_hip.so; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"liblbfem_hip.so is not built ({LIB_PATH}); run lbm-fem_amd/build.py")
        L = C.CDLL(LIB_PATH)
        i, d, p, i32, i64 = C.c_int, C.c_double, _P, C.POINTER(C.c_int32), C.POINTER(C.c_int64)
        sig = {
            "lbfem_last_error": (C.c_size_t, [C.c_char_p, C.c_size_t]),
            "lbfem_device_count": (i, [C.POINTER(C.c_int)]),
            "lbfem_disc_create": (i, [C.POINTER(MeshDesc), C.POINTER(_vp)]),
            "lbfem_disc_destroy": (i, [_vp]),
            "lbfem_disc_get_info": (i, [_vp, C.POINTER(DiscInfo)]),
            "lbfem_disc_node_coordinates": (i, [_vp, p]),
            "lbfem_disc_constrained": (i, [_vp, i32]),
            "lbfem_disc_lumped_mass": (i, [_vp, p]),
            "lbfem_disc_mass_diagonal": (i, [_vp, p]),
            "lbfem_disc_set_walls": (i, [_vp, C.c_int32, i32, p]),
            "lbfem_disc_time_step": (i, [_vp, d, p]),
            "lbfem_advection_apply": (i, [_vp, i, _vp, _vp, _vp, p, d, d, i, _vp]),
            "lbfem_collide_bgk": (i, [_vp, _vp, d, _vp]),
            "lbfem_compute_equilibrium": (i, [_vp, _vp, _vp, _vp]),
            "lbfem_predictor_corrector": (i, [_vp, _vp, _vp, _vp, d, _vp]),
            "lbfem_streaming_vmult": (i, [_vp, _vp, _vp, d, p, _vp]),
            "lbfem_mass_solve": (i, [_vp, C.POINTER(MassSettings), _vp, _vp, p, d, C.POINTER(C.c_int), _vp]),
            "lbfem_mass_solve_noalloc": (i, [_vp, C.POINTER(MassSettings), _vp, _vp, p, d, C.POINTER(C.c_int),
                                             _vp]),
            "lbfem_scheme_create": (i, [_vp, C.POINTER(SchemeSettings), C.POINTER(_vp)]),
            "lbfem_scheme_destroy": (i, [_vp]),
            "lbfem_scheme_set_time_step": (i, [_vp, d, d]),
            "lbfem_scheme_set_initial_populations": (i, [_vp, p]),
            "lbfem_scheme_set_populations": (i, [_vp, p]),
            "lbfem_scheme_get_populations": (i, [_vp, p]),
            "lbfem_scheme_device_populations": (i, [_vp, C.POINTER(_vp)]),
            "lbfem_scheme_step": (i, [_vp, i]),
            "lbfem_scheme_synchronize": (i, [_vp]),
            "lbfem_scheme_cg_iterations": (i, [_vp, i64]),
            "lbfem_scheme_last_cg_iterations": (i, [_vp, i32]),
            "lbfem_scheme_enable_stage_timers": (i, [_vp, i]),
            "lbfem_scheme_stage_times": (i, [_vp, p]),
            "lbfem_diagnostics": (i, [_vp, _vp, _vp, _vp, p, _vp]),
            "lbfem_hbm_copy": (i, [_vp, _vp, C.c_size_t, _vp]),
            "lbfem_stream_function": (i, [C.POINTER(MeshDesc), p, p, p, C.POINTER(C.c_int)]),
            "lbfem_disc_stream_function": (i, [_vp, _vp, _vp, _vp, C.POINTER(C.c_int), _vp]),
            "lbfem_batch_create": (i, [C.POINTER(_vp), i, C.POINTER(_vp)]),
            "lbfem_batch_add": (i, [_vp, _vp]),
            "lbfem_batch_remove": (i, [_vp, _vp]),
            "lbfem_batch_member_status": (i, [_vp, i32]),
            "lbfem_batch_velocity_change": (i, [_vp, i32, p, p, i64]),
            "lbfem_scheme_velocity_change": (i, [_vp, p, p, i64]),
            "lbfem_scheme_diagnostics": (i, [_vp, p, d, p]),
            "lbfem_batch_destroy": (i, [_vp]),
            "lbfem_batch_set_graph_steps": (i, [_vp, i]),
            "lbfem_batch_step": (i, [_vp, i]),
            "lbfem_batch_synchronize": (i, [_vp]),
            "lbfem_batch_get_stream": (i, [_vp, C.POINTER(_vp)]),
            "lbfem_batch_info": (i, [_vp, C.POINTER(C.c_int), i64, i64]),
            "lbfem_batch_set_shape": (i, [_vp, i, i]),
            "lbfem_batch_autotune": (i, [_vp, p]),
            "lbfem_scheme_set_tile_variant": (i, [_vp, i]),
            "lbfem_scheme_get_stream": (i, [_vp, C.POINTER(_vp)]),
            "lbfem_scheme_get_tile_variant": (i, [_vp, C.POINTER(C.c_int)]),
            "lbfem_scheme_autotune": (i, [_vp, i, C.POINTER(C.c_int), p]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    """Names declared in include/lbfem_hip.h that the library must export."""
    hdr = os.path.join(os.path.dirname(os.path.dirname(HERE)), "include", "lbfem_hip.h")
    import re
    text = open(hdr).read() if os.path.exists(hdr) else ""
    return sorted(set(re.findall(r"\b(lbfem_[a-z_0-9]+)\s*\(", text)))


def last_error():
    buf = C.create_string_buffer(1024)
    lib().lbfem_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(code):
    if code != OK:
        msg = last_error()
        if code == ENOCONV:
            raise NoConvergence(code, msg)
        raise LbfemError(code, msg)


def device_count():
    c = C.c_int(0)
    check(lib().lbfem_device_count(C.byref(c)))
    return c.value


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


class Discretization:
    """include/lbfem/discretization.h:69-189 for the box meshes of a TestCase."""

    def __init__(self, tc: TestCase, n_cells: int, degree: int = 1):
        self.tc = tc
        desc = MeshDesc(n_cells, degree, tc.n_periodic, tc.L, tc.stretch, getattr(tc, "distort", 0.0))
        h = _vp()
        check(lib().lbfem_disc_create(C.byref(desc), C.byref(h)))
        self.h = h
        self._destroy = lib().lbfem_disc_destroy  # held, so that teardown at interpreter exit finds it
        info = DiscInfo()
        check(lib().lbfem_disc_get_info(self.h, C.byref(info)))
        self.N, self.p = n_cells, degree
        self.n_nodes = int(info.n_nodes)
        self.n_independent = int(info.n_independent)
        self.stride = int(info.stride)
        self.n_per_dir = int(info.n_per_dir)
        self.h_min = float(info.h_min)
        self.node = self.node_coordinates()
        mask, ux, uy = tc.wall_velocity(self.node[:, 0], self.node[:, 1])
        idx = np.nonzero(mask)[0].astype(np.int32)
        vel = np.ascontiguousarray(np.stack([ux[idx], uy[idx]], axis=1), dtype=np.float64)
        check(lib().lbfem_disc_set_walls(self.h, len(idx), idx.ctypes.data_as(C.POINTER(C.c_int32)), _p(vel)))
        self.walls = (idx, vel)

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def node_coordinates(self):
        xy = np.zeros((self.n_nodes, 2))
        check(lib().lbfem_disc_node_coordinates(self.h, _p(xy)))
        return xy

    def constrained(self):
        v = np.zeros(self.n_nodes, dtype=np.int32)
        check(lib().lbfem_disc_constrained(self.h, v.ctypes.data_as(C.POINTER(C.c_int32))))
        return v.astype(bool)

    def lumped_mass(self):
        v = np.zeros(self.n_nodes)
        check(lib().lbfem_disc_lumped_mass(self.h, _p(v)))
        return v

    def mass_diagonal(self):
        v = np.zeros(self.n_nodes)
        check(lib().lbfem_disc_mass_diagonal(self.h, _p(v)))
        return v

    def time_step(self, cfl):
        dt = C.c_double(0)
        check(lib().lbfem_disc_time_step(self.h, cfl, C.byref(dt)))
        return dt.value

    # --- device-vector helpers (torch tensors as device memory)
    def device_vector(self, blocks, torch_module=None):
        import torch
        return torch.zeros(blocks * self.stride, dtype=torch.float64, device="cuda")

    def to_device(self, host, blocks):
        """9 x n (or k x n) host array -> device vector with the disc's block stride."""
        import torch
        host = np.asarray(host, dtype=np.float64).reshape(blocks, self.n_nodes)
        pad = np.zeros((blocks, self.stride))
        pad[:, : self.n_nodes] = host
        return torch.from_numpy(pad.reshape(-1)).to("cuda")

    def to_host(self, dev, blocks):
        return dev.detach().cpu().numpy().reshape(blocks, self.stride)[:, : self.n_nodes].copy()


def _ptr(t):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    return C.c_void_p(stream)


# --------------------------------------------------------------- operators
def advection_apply(disc, kind, rhs, f, feq, dt, lam, e=D2Q9_E, remove_mass_flux=True, stream=None):
    """AdvectionOperator::apply (include/lbfem/advection.h:49-58) on device tensors."""
    e = np.ascontiguousarray(e, dtype=np.float64).reshape(-1)
    check(lib().lbfem_advection_apply(disc.h, kind, _ptr(rhs), _ptr(f), _ptr(feq), _p(e), dt, lam,
                                      int(remove_mass_flux), _stream(stream)))


def collide_bgk(disc, g, omega, stream=None):
    check(lib().lbfem_collide_bgk(disc.h, _ptr(g), omega, _stream(stream)))


def compute_equilibrium(disc, f, feq, stream=None):
    check(lib().lbfem_compute_equilibrium(disc.h, _ptr(f), _ptr(feq), _stream(stream)))


def predictor_corrector(disc, f, feq, incr, theta, stream=None):
    check(lib().lbfem_predictor_corrector(disc.h, _ptr(f), _ptr(feq), _ptr(incr), theta, _stream(stream)))


def streaming_vmult(disc, y, x, c=0.0, e=(0.0, 0.0), stream=None):
    ee = np.asarray(e, dtype=np.float64)
    check(lib().lbfem_streaming_vmult(disc.h, _ptr(y), _ptr(x), c, _p(ee), _stream(stream)))


def mass_solve(disc, x, r, mass=MASS_CG, richardson=2, cg_tolerance=1e-8, cg_max_iterations=200, e=(0.0, 0.0),
               tg3_coefficient=0.0, stream=None):
    """MassSolver::solve (include/lbfem/mass.h:184-213); returns the CG iterations."""
    st = MassSettings(mass, richardson, cg_tolerance, cg_max_iterations, 0)
    ee = np.asarray(e, dtype=np.float64)
    its = C.c_int(0)
    check(lib().lbfem_mass_solve(disc.h, C.byref(st), _ptr(x), _ptr(r), _p(ee), tg3_coefficient, C.byref(its),
                                 _stream(stream)))
    return its.value


def mass_solve_noalloc(disc, x, r, mass=MASS_CG, richardson=2, cg_tolerance=1e-8, cg_max_iterations=200,
                       e=(0.0, 0.0), tg3_coefficient=0.0, stream=None):
    """lbfem_mass_solve_noalloc: mass_solve on the library's fixed per-GPU work space (no allocation; at most
    1179648 nodes); the same x and iteration count bit for bit. Calls that overlap on one GPU race."""
    st = MassSettings(mass, richardson, cg_tolerance, cg_max_iterations, 0)
    ee = np.asarray(e, dtype=np.float64)
    its = C.c_int(0)
    check(lib().lbfem_mass_solve_noalloc(disc.h, C.byref(st), _ptr(x), _ptr(r), _p(ee), tg3_coefficient,
                                         C.byref(its), _stream(stream)))
    return its.value


# --------------------------------------------------------------- schemes
class Scheme:
    """include/lbfem/schemes.h:31-75 with the streaming of streaming.h."""

    scheme_id = BARDOW

    def __init__(self, disc: Discretization, streaming=TG2, mass=MASS_CG, richardson=2, cg_tolerance=1e-8,
                 cg_max_iterations=200, precision=FP64):
        self.disc = disc
        st = SchemeSettings(self.scheme_id, streaming,
                            MassSettings(mass, richardson, cg_tolerance, cg_max_iterations, 0), precision)
        h = _vp()
        check(lib().lbfem_scheme_create(disc.h, C.byref(st), C.byref(h)))
        self.h = h
        self._destroy = lib().lbfem_scheme_destroy
        self.dt = self.lam = None

    def __del__(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def set_time_step(self, dt, lam):
        check(lib().lbfem_scheme_set_time_step(self.h, dt, lam))
        self.dt, self.lam = dt, lam

    def set_initial_populations(self, tc=None):
        tc = tc or self.disc.tc
        xy = self.disc.node
        f0 = np.ascontiguousarray(tc.initial_populations(xy[:, 0], xy[:, 1], self.lam), dtype=np.float64)
        check(lib().lbfem_scheme_set_initial_populations(self.h, _p(f0)))

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        check(lib().lbfem_scheme_set_populations(self.h, _p(f)))

    @property
    def f(self):
        out = np.zeros((Q, self.disc.n_nodes))
        check(lib().lbfem_scheme_get_populations(self.h, _p(out)))
        return out

    def device_populations(self):
        ptr = _vp()
        check(lib().lbfem_scheme_device_populations(self.h, C.byref(ptr)))
        return ptr.value

    def step(self, n=1):
        check(lib().lbfem_scheme_step(self.h, n))

    def synchronize(self):
        check(lib().lbfem_scheme_synchronize(self.h))

    @property
    def cg_iterations(self):
        v = C.c_int64(0)
        check(lib().lbfem_scheme_cg_iterations(self.h, C.byref(v)))
        return v.value

    def last_cg_iterations(self):
        """The last step's CG steps per solve (SolverControl::last_step(), include/lbfem/mass.h:204-211),
        shape (2 sweeps, 8 populations); 0 where no CG ran."""
        v = np.zeros(16, dtype=np.int32)
        check(lib().lbfem_scheme_last_cg_iterations(self.h, v.ctypes.data_as(C.POINTER(C.c_int32))))
        return v.reshape(2, 8)

    def diagnostics(self, u_exact=None, exact_scale=1.0):
        """The six sums of cgdbe's diagnostics (apps/cgdbe/cgdbe.cc:353-368) from this scheme's populations on
        the device, against exact_scale times the exact velocity last given (2 per node, host)."""
        s = np.zeros(6)
        ue = None if u_exact is None else np.ascontiguousarray(u_exact, dtype=np.float64).reshape(-1)
        check(lib().lbfem_scheme_diagnostics(self.h, None if ue is None else _p(ue), float(exact_scale), _p(s)))
        return s

    def velocity_change(self):
        """(max |u - u_prev|, lumped-mass mean |u - u_prev|, non-finite nodes) against the velocity at the
        previous call, computed on the device (the first call records the velocity and returns inf)."""
        mx, mean, bad = C.c_double(0), C.c_double(0), C.c_int64(0)
        check(lib().lbfem_scheme_velocity_change(self.h, C.byref(mx), C.byref(mean), C.byref(bad)))
        return mx.value, mean.value, bad.value

    def enable_stage_timers(self, on=True):
        check(lib().lbfem_scheme_enable_stage_timers(self.h, int(on)))

    def stage_times(self):
        t = np.zeros(4)
        check(lib().lbfem_scheme_stage_times(self.h, _p(t)))
        return {"collision": t[0], "advection": t[1], "mass": t[2], "total": t[3]}

    @property
    def stream_handle(self):
        """The scheme's hipStream_t as an int (torch.cuda.ExternalStream takes it)."""
        h = _vp()
        check(lib().lbfem_scheme_get_stream(self.h, C.byref(h)))
        return h.value or 0

    @property
    def tile_variant(self):
        v = C.c_int(0)
        check(lib().lbfem_scheme_get_tile_variant(self.h, C.byref(v)))
        return v.value

    @tile_variant.setter
    def tile_variant(self, v):
        check(lib().lbfem_scheme_set_tile_variant(self.h, int(v)))

    def autotune(self, force=False):
        """Times the tile kernel's variants on this scheme, keeps the fastest; returns (variant, seconds)."""
        v = C.c_int(0)
        t = np.zeros(TILE_VARIANTS)
        check(lib().lbfem_scheme_autotune(self.h, int(force), C.byref(v), _p(t)))
        return v.value, t


class Bardow(Scheme):
    """include/lbfem/schemes.h:104-142."""
    scheme_id = BARDOW


class LeeLin(Scheme):
    """include/lbfem/schemes.h:77-100."""
    scheme_id = LEELIN


def diagnostics(disc, f_dev_ptr, u_exact=None, u_prev=None, stream=None):
    """The six sums of cgdbe's diagnostics (apps/cgdbe/cgdbe.cc:353-368) from device populations."""
    s = np.zeros(6)
    check(lib().lbfem_diagnostics(disc.h, C.c_void_p(f_dev_ptr), _ptr(u_exact), _ptr(u_prev), _p(s),
                                  _stream(stream)))
    return s


def disc_stream_function(disc, u, v, psi, stream=None):
    """lbfem_disc_stream_function: psi (device) of the device nodal velocity u, v on the walled box of disc
    (include/lbfem/io.h:31-93); returns the CG iterations."""
    its = C.c_int(0)
    check(lib().lbfem_disc_stream_function(disc.h, C.c_void_p(u.data_ptr()), C.c_void_p(v.data_ptr()),
                                            C.c_void_p(psi.data_ptr()), C.byref(its), _stream(stream)))
    return its.value


def hbm_copy(dst, src, stream=None):
    """lbfem_hbm_copy: device copy of src into dst (torch tensors of equal byte size, on the GPU)."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() != nbytes:
        raise LbfemError(EINVAL, "hbm_copy: tensors differ in size")
    check(lib().lbfem_hbm_copy(C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()), nbytes, _stream(stream)))


class Batch:
    """lbfem_batch: schemes stepped together. Lumped TG2 members: one launch per (degree, scheme, precision)
    group per step, steps replayed from captured HIP graphs; consistent-mass FP64 members: each stage of the
    step one launch over the (degree, scheme, streaming) group. Members are read after synchronize(); the
    batch keeps references to them, and is destroyed (close()) before they are."""

    def __init__(self, schemes=(), graph_steps=8):
        self.schemes = list(schemes)
        arr = (_vp * max(1, len(self.schemes)))(*[s.h.value for s in self.schemes])
        h = _vp()
        check(lib().lbfem_batch_create(arr, len(self.schemes), C.byref(h)))
        self.h = h
        self._destroy = lib().lbfem_batch_destroy
        if graph_steps != 8:
            check(lib().lbfem_batch_set_graph_steps(self.h, int(graph_steps)))

    def close(self):
        if getattr(self, "h", None):
            self._destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, n=1):
        check(lib().lbfem_batch_step(self.h, int(n)))

    def synchronize(self):
        check(lib().lbfem_batch_synchronize(self.h))

    def member_status(self):
        """per member: OK, or ENOCONV once its CG reached its iteration cap"""
        st = np.zeros(max(1, len(self.schemes)), dtype=np.int32)
        check(lib().lbfem_batch_member_status(self.h, st.ctypes.data_as(C.POINTER(C.c_int32))))
        return st[: len(self.schemes)]

    def add(self, scheme):
        """lbfem_batch_add: the scheme joins the running batch from its current populations."""
        check(lib().lbfem_batch_add(self.h, scheme.h))
        self.schemes.append(scheme)

    def remove(self, scheme):
        """lbfem_batch_remove: the scheme leaves the batch (its populations are then readable)."""
        check(lib().lbfem_batch_remove(self.h, scheme.h))
        self.schemes = [s for s in self.schemes if s is not scheme]

    def velocity_change(self, members=None):
        """Per member (in member order): max |u - u_prev|, mean |u - u_prev|, non-finite nodes, computed on the
        device in one launch; members (optional): the schemes to measure (the others are left untouched, their
        entries NaN / -1)."""
        n = len(self.schemes)
        mx, mean = np.full(n, np.nan), np.full(n, np.nan)
        bad = np.full(n, -1, dtype=np.int64)
        mask = None
        if members is not None:
            ids = {id(s) for s in members}
            mask = np.array([1 if id(s) in ids else 0 for s in self.schemes], dtype=np.int32)
        check(lib().lbfem_batch_velocity_change(
            self.h, None if mask is None else mask.ctypes.data_as(C.POINTER(C.c_int32)), _p(mx), _p(mean),
            bad.ctypes.data_as(C.POINTER(C.c_int64))))
        return mx, mean, bad

    def set_shape(self, variant=-1, grid=0):
        check(lib().lbfem_batch_set_shape(self.h, int(variant), int(grid)))

    def autotune(self):
        t = C.c_double(0.0)
        check(lib().lbfem_batch_autotune(self.h, C.byref(t)))
        return t.value

    def info(self):
        g, n, t = C.c_int(0), C.c_int64(0), C.c_int64(0)
        check(lib().lbfem_batch_info(self.h, C.byref(g), C.byref(n), C.byref(t)))
        return {"groups": g.value, "nodes": n.value, "tiles": t.value}

    @property
    def stream_handle(self):
        h = _vp()
        check(lib().lbfem_batch_get_stream(self.h, C.byref(h)))
        return h.value or 0


def stream_function(n_cells, degree, u, v, L=1.0, stretch=0.0):
    """include/lbfem/io.h:31-93: psi with u = -dpsi/dy, v = dpsi/dx, psi = 0 on the walls of the box
    (lexicographic nodal host arrays); returns (psi, CG iterations). Computed on the device."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    psi = np.zeros_like(u)
    its = C.c_int(0)
    desc = MeshDesc(int(n_cells), int(degree), 0, float(L), float(stretch), 0.0)
    check(lib().lbfem_stream_function(C.byref(desc), _p(u), _p(v), _p(psi), C.byref(its)))
    return psi, its.value
