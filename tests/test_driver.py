"""The C++ driver cgdbe_hip (apps/cgdbe/cgdbe.cc over the C ABI): its command
line on the CPU; on the GPU the reference's CI check (cavity, 16^2 Q1, Re 100
then 400 by continuation, lumped mass, each to steady state; psi_max within 10 %
of Ghia, Ghia & Shin 1982, .github/workflows/ci.yml:50-58) and a checkpoint /
restart that continues bitwise."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "lbm-fem_amd", "lbfem_amd", "_lib", "cgdbe_hip")

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="driver not built (lbm-fem_amd/build.py)")


def run(args, cwd, timeout=300):
    return subprocess.run([DRIVER, *args], cwd=cwd, capture_output=True, text=True, timeout=timeout)


def test_help_and_bad_options(tmp_path):
    r = run(["--help"], tmp_path)
    assert r.returncode == 0 and "--reynolds" in r.stdout
    r = run(["--case", "pipe"], tmp_path)
    assert r.returncode == 1 and "--case" in r.stderr
    r = run(["--mass", "richardsonX"], tmp_path)
    assert r.returncode == 1


def psi_max(path):
    for line in open(path):
        if line.startswith("# primary vortex:"):
            return float(line.split()[5])
    raise AssertionError(f"no primary vortex line in {path}")


@pytest.mark.gpu
def test_ci_cavity_ghia(tmp_path):
    r = run(["--case", "cavity", "--refine", "4", "--reynolds", "100,400", "--mass", "lumped"], tmp_path)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    assert r.stdout.count("-> steady state") == 2
    for Re, ghia in ((100, 0.1034), (400, 0.1139)):
        p = psi_max(tmp_path / f"cavity_Re{Re}.dat")
        assert abs(p - ghia) / ghia < 0.1, (Re, p, ghia)


@pytest.mark.gpu
@pytest.mark.parametrize("args", [
    ["--case", "tgv", "--cells", "24", "--mass", "lumped"],
    ["--case", "couette", "--cells", "16", "--scheme", "leelin", "--mass", "cg", "--cg-tol", "1e-12"],
])
def test_checkpoint_restart_continues(tmp_path, args):
    common = [*args, "--no-output", "--reynolds", "200"]
    r = run([*common, "--steps", "60", "--checkpoint", "whole"], tmp_path)
    assert r.returncode == 0, r.stderr
    r = run([*common, "--steps", "30", "--checkpoint", "half"], tmp_path)
    assert r.returncode == 0, r.stderr
    r = run([*common, "--steps", "30", "--restart", "half", "--checkpoint", "rest"], tmp_path)
    assert r.returncode == 0, r.stderr
    a = np.fromfile(tmp_path / "whole.0")
    b = np.fromfile(tmp_path / "rest.0")
    assert a.size == b.size and np.isfinite(a).all()
    if "cg" in args:
        # the restarted CG starts from a zero increment history instead of the extrapolated one
        assert np.abs(a - b).max() < 1e-11 * np.abs(a).max()
    else:
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_distorted_vortex_published_error(tmp_path):
    """--distort through the driver: docs/results.md:226-229, 48^2 Q1, modes (1, 4), Re 100, Ma 0.05,
    t = 0.25 t_ref, Bardow, CG 1e-8, TG2 at CFL 0.25 on the +-25 % mesh: rel. L2 error 2.07e-2."""
    r = run(["--case", "tgv", "--cells", "48", "--mach", "0.05", "--modes", "1,4", "--reynolds", "100",
             "--tend", "0.25", "--distort", "0.04", "--no-output"], tmp_path)
    assert r.returncode == 0, r.stderr + r.stdout[-2000:]
    rows = [ln.split() for ln in r.stdout.splitlines()]
    rows = [x for x in rows if len(x) == 6 and x[0].isdigit()]
    err = float(rows[-1][4])
    assert abs(err - 2.07e-2) <= 0.5e-4, err
    # the summary's stage breakdown (cgdbe.cc:519-538): the CG run spends its time in the mass solves
    import re
    sec = {}
    for ln in r.stdout.splitlines():
        m = re.match(r"^\s*(collision|advection|mass solves)\s+([0-9.]+) s ", ln)
        if m:
            sec[m.group(1)] = float(m.group(2))
    assert set(sec) == {"collision", "advection", "mass solves"}, r.stdout[-1500:]
    assert sec["mass solves"] > sec["collision"] > 0 and sec["advection"] > 0, sec
