"""The flows of the reference as host-side setup: periodicity, mesh clustering,
wall nodes, initial populations and exact solutions
(include/lbfem/test_case.h:23-83, include/lbfem/test_cases.h:21-245).

Vectorised over nodes with numpy; the populations are 9 x n arrays (block a
for population a), as the device expects them.
"""
from __future__ import annotations

import math

import numpy as np

Q = 9
CS2 = 1.0 / 3.0
EX = np.array([0, 1, 1, 0, -1, -1, -1, 0, 1], dtype=np.float64)
EY = np.array([0, 0, 1, 1, 1, 0, -1, -1, -1], dtype=np.float64)
W = np.array([4 / 9, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36, 1 / 9, 1 / 36])


def equilibrium(rho, ux, uy):
    """include/lbfem/d2q9.h:44-55, vectorised: 9 x n."""
    uu = ux * ux + uy * uy
    eu = EX[:, None] * ux + EY[:, None] * uy
    return W[:, None] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu)


def moments(f):
    """include/lbfem/d2q9.h:30-41, vectorised."""
    rho = f.sum(axis=0)
    return rho, (EX @ f) / rho, (EY @ f) / rho


class TestCase:
    """include/lbfem/test_case.h:23-83: lattice units, U0 = Ma c_s, nu = U0 L / Re."""

    n_periodic = 0
    steady = False
    stretch = 0.0
    distort = 0.0

    def __init__(self, mach):
        self.L = 1.0
        self.rho0 = 1.0
        self.U0 = mach * math.sqrt(CS2)
        self.nu = 0.0

    def set_reynolds(self, Re):
        self.nu = self.U0 * self.L / Re

    def relaxation_time(self):
        return self.nu / CS2

    def energy_scale(self):
        return self.U0 * self.U0

    def wall_velocity(self, x, y):
        """(mask, ux, uy) of the wall nodes among the points (x, y)."""
        z = np.zeros_like(x)
        return np.zeros(x.shape, dtype=bool), z, z

    def exact(self, x, y, t):
        return np.full_like(x, self.rho0), np.zeros_like(x), np.zeros_like(x)

    def initial_populations(self, x, y, lam):
        rho, ux, uy = self.exact(x, y, 0.0)
        return equilibrium(rho, ux, uy)


class TaylorGreenVortex(TestCase):
    """include/lbfem/test_cases.h:21-119 (Bardow et al., Eqs. 14-16), modes (n1, n2); distort > 0 maps the
    mesh by x -> x + eps L sin(2 pi x / L) sin(2 pi y / L) (1, 1) into general quadrilaterals
    (test_cases.h:53-65)."""

    n_periodic = 2

    def __init__(self, mach, n1=1, n2=4, distort=0.0):
        super().__init__(mach)
        self.n1, self.n2 = n1, n2
        self.distort = float(distort)

    @property
    def k1(self):
        return 2 * math.pi * self.n1 / self.L

    @property
    def k2(self):
        return 2 * math.pi * self.n2 / self.L

    def reference_time(self):
        return 1.0 / (self.nu * (self.k1 ** 2 + self.k2 ** 2))

    def energy_scale(self):
        r = self.k1 / self.k2
        return 0.25 * self.U0 ** 2 * (1 + r * r)

    def exact(self, x, y, t):
        F = math.exp(-self.nu * (self.k1 ** 2 + self.k2 ** 2) * t)
        x1, y2, r = self.k1 * x, self.k2 * y, self.k1 / self.k2
        dp = -0.25 * self.rho0 * self.U0 ** 2 * F * F * (np.cos(2 * x1) + r * r * np.cos(2 * y2))
        return (self.rho0 + dp / CS2, -self.U0 * F * np.cos(x1) * np.sin(y2),
                self.U0 * r * F * np.sin(x1) * np.cos(y2))

    def initial_populations(self, x, y, lam):
        rho, ux, uy = self.exact(x, y, 0.0)
        f = equilibrium(rho, ux, uy)
        x1, y2 = self.k1 * x, self.k2 * y
        dudx = self.U0 * self.k1 * np.sin(x1) * np.sin(y2)
        sh = self.U0 * (self.k1 ** 2 - self.k2 ** 2) / self.k2 * np.cos(x1) * np.cos(y2)
        for a in range(Q):
            f[a] -= lam / CS2 * W[a] * rho * ((EX[a] ** 2 - EY[a] ** 2) * dudx + EX[a] * EY[a] * sh)
        return f


class CouetteFlow(TestCase):
    """include/lbfem/test_cases.h:125-180: walls at y = 0 (rest) and y = L (U0), periodic in x."""

    n_periodic = 1

    def reference_time(self):
        return self.L * self.L / self.nu

    def wall_velocity(self, x, y):
        lo = np.abs(y) < 1e-12 * self.L
        hi = np.abs(y - self.L) < 1e-12 * self.L
        return lo | hi, np.where(hi, self.U0, 0.0), np.zeros_like(x)

    def exact(self, x, y, t):
        u = self.U0 * y / self.L
        for m in range(1, 401):
            lm = m * math.pi / self.L
            u = u + 2 * self.U0 * (-1.0 if m % 2 else 1.0) / (lm * self.L) * math.exp(-self.nu * lm * lm * t) * np.sin(lm * y)
        return np.full_like(x, self.rho0), u, np.zeros_like(x)

    def initial_populations(self, x, y, lam):
        one = np.ones_like(x)
        return equilibrium(self.rho0 * one, 0 * one, 0 * one)


class LidDrivenCavity(TestCase):
    """include/lbfem/test_cases.h:186-245: tanh-clustered mesh, the lid at y = L moves with U0,
    the two top corners belong to the side walls."""

    n_periodic = 0
    steady = True

    def __init__(self, mach, stretch=1.2):
        super().__init__(mach)
        self.stretch = stretch

    def reference_time(self):
        return self.L / self.U0

    def wall_velocity(self, x, y):
        L = self.L
        on = lambda a, a0: np.abs(a - a0) < 1e-12 * L
        side = on(x, 0) | on(x, L) | on(y, 0)
        lid = on(y, L) & ~side
        return side | lid, np.where(lid, self.U0, 0.0), np.zeros_like(x)


def make(case, mach=0.1, modes=(1, 4), stretch=1.2, distort=0.0):
    """apps/cgdbe/cgdbe.cc:92-103."""
    if case == "tgv":
        return TaylorGreenVortex(mach, *modes, distort=distort)
    if case == "couette":
        return CouetteFlow(mach)
    if case == "cavity":
        return LidDrivenCavity(mach, stretch)
    raise ValueError(f"--case: unknown value {case}")
