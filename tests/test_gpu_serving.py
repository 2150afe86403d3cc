"""The continuous-batching server: requests that arrive while others run are
admitted into the running batch, and every result is bitwise the one the same
request gets stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():
    pytest.skip("no GPU", allow_module_level=True)

import lbfem_amd as A  # noqa: E402
from lbfem_amd import serving, workloads  # noqa: E402


@pytest.mark.parametrize("precision", [A.MIXED, A.FP64])
def test_served_results_equal_alone(precision):
    reqs = workloads.request_stream(7, 10, rate=400.0, sizes=(16, 24, 40), steps=(200, 900), re=(100.0, 1000.0))
    srv = serving.Server(A, precision=precision, max_slots=4, tick_steps=150, keep_results=True)
    done, summary = srv.run(reqs)
    assert summary["requests"] == len(reqs)
    assert summary["ticks"] > 1
    by_id = {d["id"]: d for d in done}
    # the slot limit made later requests wait for earlier ones to finish, and the
    # batch took in requests while others were part way through their steps
    assert any(b["admit"] >= a["done"] for a in done for b in done)
    assert any(a["admit"] < b["admit"] < a["done"] for a in done for b in done)
    for r in reqs:
        s = workloads.make_scheme(A, r, precision)
        s.step(r["steps"])
        assert np.array_equal(by_id[r["id"]]["f"], s.f), r


def test_admission_budget_limits_concurrency():
    reqs = workloads.request_stream(11, 6, rate=1000.0, sizes=(32,), steps=(100, 300), q2=0.0)
    one = serving.request_bytes(reqs[0])
    srv = serving.Server(A, budget_bytes=1.5 * one, max_slots=96, tick_steps=100)
    done, summary = srv.run(reqs)
    assert summary["requests"] == 6
    # with room for one request at a time, each admission follows the previous completion
    d = sorted(done, key=lambda x: x["admit"])
    for a, b in zip(d, d[1:]):
        assert b["admit"] >= a["done"] - 1e-9
