"""Seeded synthetic workloads: the batch of uneven simulations and the served
request stream. Every instance is drawn from splitmix64 of (base seed, index),
so a workload is reproduced from its seed alone."""
import math

import numpy as np

MASK = (1 << 64) - 1
CS = 1.0 / math.sqrt(3.0)


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & MASK
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class Stream:
    """Uniform doubles in [0, 1) from splitmix64, seeded per instance."""

    def __init__(self, seed):
        self.state = seed & MASK

    def uniform(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        z ^= z >> 31
        return (z >> 11) * (1.0 / (1 << 53))


def batch_instance(base_seed, i):
    """Instance i of the uneven batch: flow uniform over {tgv, couette, cavity}; N = floor(32 16^u) cells
    per direction (32..512); Q1 with probability 0.75, else Q2; Re = 50 64^u; U0 uniform in [0.02, 0.1]."""
    r = Stream(splitmix64(base_seed ^ (i * 0x632BE59BD9B4E019 & MASK)))
    case = ("tgv", "couette", "cavity")[min(2, int(3 * r.uniform()))]
    N = int(math.floor(32 * 16 ** r.uniform()))
    p = 1 if r.uniform() < 0.75 else 2
    Re = 50.0 * 64 ** r.uniform()
    U = 0.02 + 0.08 * r.uniform()
    return {"case": case, "N": N, "p": p, "Re": Re, "U": U, "mach": U / CS}


def batch(base_seed=20261203, n=96):
    return [batch_instance(base_seed, i) for i in range(n)]


def nodes(inst):
    return (inst["p"] * inst["N"] + 1) ** 2


def request_stream(base_seed, n, rate, sizes=(32, 64, 128, 256), steps=(2000, 20000), re=(100.0, 5000.0), q2=0.25,
                   steady=0.5, tol=1e-6):
    """n served requests: Poisson arrivals at `rate` per second; cavity or Couette; N from `sizes`, Q2 with
    probability q2, else Q1; Re log-uniform in re; a number of steps uniform in `steps` (multiples of 100),
    which is the request's length, or, with probability `steady`, the cap of a steady-state solve to
    tolerance `tol` (mode "steady")."""
    r = Stream(splitmix64(base_seed))
    t = 0.0
    out = []
    for i in range(n):
        t += -math.log(1.0 - r.uniform()) / rate
        case = "cavity" if r.uniform() < 0.5 else "couette"
        N = sizes[min(len(sizes) - 1, int(len(sizes) * r.uniform()))]
        p = 2 if r.uniform() < q2 else 1
        Re = re[0] * (re[1] / re[0]) ** r.uniform()
        U = 0.02 + 0.08 * r.uniform()
        k = int(steps[0] + (steps[1] - steps[0]) * r.uniform()) // 100 * 100
        req = {"id": i, "arrival": t, "case": case, "N": N, "p": p, "Re": Re, "U": U, "mach": U / CS,
               "steps": max(100, k), "mode": "steps"}
        if r.uniform() < steady:
            req["mode"] = "steady"
            req["tol"] = tol
        out.append(req)
    return out


def make_scheme(A, inst, precision, scheme=None, cfl=0.5, mass=None, cg_tolerance=1e-8):
    """A TG2 scheme for a workload instance (the lumped mass unless `mass` says otherwise), initial
    populations set. The default stepper is Bardow's: Lee & Lin's is unstable for dt / lambda below about 1
    (the low-Re, finely stretched cavities of these workloads) with either mass, in the oracle as here."""
    tc = A.test_cases.make(inst["case"], mach=inst["mach"], modes=(1, 1), stretch=1.2)
    tc.set_reynolds(inst["Re"])
    d = A.Discretization(tc, inst["N"], inst["p"])
    cls = A.LeeLin if scheme == A.LEELIN else A.Bardow
    s = cls(d, mass=A.MASS_LUMPED if mass is None else mass, precision=precision, cg_tolerance=cg_tolerance)
    s.set_time_step(d.time_step(cfl), tc.relaxation_time())
    s.set_initial_populations(tc)
    return s
