"""The OpenMP build of the oracle (the multi-core CPU baseline) parallelises
only loops whose iterations are independent (cells, nodes), so its populations
are bitwise those of the serial oracle."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest


ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHILD = r"""
import json, os, sys
import numpy as np
sys.path.insert(0, os.path.join(sys.argv[1], "oracle"))
import oracle as O
out = {}
for case, scheme, stretch in ((O.CAVITY, O.LEELIN, 1.2), (O.TGV, O.BARDOW, 0.0), (O.COUETTE, O.BARDOW, 0.0)):
    d = O.Disc(12, 2, case, mach=0.1, modes=(1, 1), stretch=stretch, reynolds=200.0)
    s = O.Scheme(d, scheme, O.TG2, O.MASS_LUMPED)
    s.set_time_step(d.time_step(0.4), d.relaxation_time)
    s.set_initial_populations()
    s.step(5)
    out[f"{case}-{scheme}"] = s.populations().tolist()
# consistent mass: CG (the parallel operator, serial dot products), TG3 split and Richardson
for case, scheme, streaming, mass, p in ((O.CAVITY, O.LEELIN, O.TG2, O.MASS_CG, 1),
                                          (O.TGV, O.BARDOW, O.TG3_SPLIT, O.MASS_CG, 2),
                                          (O.CAVITY, O.BARDOW, O.TG2, O.MASS_RICHARDSON, 1)):
    d = O.Disc(10, p, case, mach=0.1, modes=(1, 1), stretch=1.2 if case == O.CAVITY else 0.0, reynolds=200.0)
    s = O.Scheme(d, scheme, streaming, mass, cg_tol=1e-10)
    s.set_time_step(d.time_step(0.4), d.relaxation_time)
    s.set_initial_populations()
    s.step(4)
    out[f"cg-{case}-{scheme}-{streaming}-{mass}"] = s.populations().tolist() + [[float(s.cg_iterations)]]
print(json.dumps(out))
"""


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "liblbfem_oracle_omp.so")),
                    reason="OpenMP oracle not built (make -C oracle)")
def test_openmp_oracle_bitwise_serial():
    runs = {}
    for omp in ("0", "1"):
        env = dict(os.environ, LBFEM_ORACLE_OMP=omp, OMP_NUM_THREADS="4")
        r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True,
                           timeout=300, check=True)
        runs[omp] = json.loads(r.stdout.strip().splitlines()[-1])
    assert runs["0"].keys() == runs["1"].keys()
    for k in runs["0"]:
        a, b = np.concatenate([np.ravel(r) for r in runs["0"][k]]), np.concatenate([np.ravel(r) for r in runs["1"][k]])
        assert np.isfinite(a).all()
        assert np.array_equal(a, b), k
