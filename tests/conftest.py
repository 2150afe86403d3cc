import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "lbm-fem_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

# The oracle's OpenMP build parallelises only loops with independent iterations (its
# populations and CG counts are bitwise the serial build's: tests/test_oracle_omp.py), so the
# tests check against it where it is built; LBFEM_ORACLE_OMP=0 selects the serial build.
if os.path.exists(os.path.join(ROOT, "oracle", "liblbfem_oracle_omp.so")):
    os.environ.setdefault("LBFEM_ORACLE_OMP", "1")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) to run")
    config.addinivalue_line("markers", "slow: long-running CPU test")
