"""The stream function (include/lbfem/io.h:31-93) of the library, on the host:
the velocity of psi = sin^2(pi x) sin^2(pi y) (zero on the walls) gives back
psi with the error of the Q_p interpolant's order, on uniform and stretched
meshes; a dropped term or a sign of either velocity component breaks it."""
import numpy as np
import pytest

import lbfem_amd as A


def field(N, p, stretch):
    v = np.array([0.5 * (1 + np.tanh(stretch * (2 * c / N - 1)) / np.tanh(stretch)) if stretch > 0 else c / N
                  for c in range(N + 1)])
    # GLL nodes of the degree on each cell (Q1: vertices; Q2: midpoints)
    ref = {1: [0.0], 2: [0.0, 0.5], 3: [0.0, 0.5 - np.sqrt(5) / 10, 0.5 + np.sqrt(5) / 10]}[p]
    xs = np.concatenate([[v[c] + (v[c + 1] - v[c]) * t for t in sorted(ref)] for c in range(N)] + [[v[N]]])
    X, Y = np.meshgrid(xs, xs)
    X, Y = X.ravel(), Y.ravel()
    psi = np.sin(np.pi * X) ** 2 * np.sin(np.pi * Y) ** 2
    u = -2 * np.pi * np.sin(np.pi * X) ** 2 * np.sin(np.pi * Y) * np.cos(np.pi * Y)
    w = 2 * np.pi * np.sin(np.pi * X) * np.cos(np.pi * X) * np.sin(np.pi * Y) ** 2
    return psi, u, w


@pytest.mark.parametrize("p,stretch", [(1, 0.0), (1, 1.2), (2, 0.0), (2, 1.2)])
def test_stream_function_converges(p, stretch):
    errs = []
    for N in (16, 32):
        psi_ex, u, v = field(N, p, stretch)
        psi, its = A.stream_function(N, p, u, v, stretch=stretch)
        assert its > 0
        errs.append(np.abs(psi - psi_ex).max())
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < {1: 6e-3, 2: 2e-4}[p], errs  # the stretched mesh is coarser at the centre
    assert rate > {1: 1.9, 2: 3.5}[p], (errs, rate)  # Q1 second order; Q2 fourth at the nodes


def test_stream_function_rejects_periodic():
    desc = A.MeshDesc(8, 1, 1, 1.0, 0.0)
    z = np.zeros(81)
    with pytest.raises(A.LbfemError):
        A.check(A.lib().lbfem_stream_function(A.C.byref(desc), A._p(z), A._p(z), A._p(z), None))


def test_stream_function_rejects_a_distorted_mesh():
    desc = A.MeshDesc(8, 1, 0, 1.0, 0.0, 0.02)
    z = np.zeros(81)
    with pytest.raises(A.LbfemError):
        A.check(A.lib().lbfem_stream_function(A.C.byref(desc), A._p(z), A._p(z), A._p(z), None))
