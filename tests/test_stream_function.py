"""The stream function (include/lbfem/io.h:31-93): (grad phi, grad psi) =
(dphi/dx, v) - (dphi/dy, u), psi = 0 on the walls, SolverCG to 1e-12 |b|.

The oracle's cell-wise restatement is pinned by mathematics: the velocity of
psi = x(1-x)y(1-y) (in the Q2 space, zero on the walls; its velocity is
interpolated exactly) gives psi back at the nodes to the CG tolerance on every
degree, uniform and stretched, and the velocity of sin^2(pi x) sin^2(pi y)
gives psi back with the error order of the Q_p interpolant; a dropped term or a
sign of either velocity component breaks both. The library's device solve is
then held to the oracle: psi within 1e-10 relative, CG iterations within +-1.
"""
import numpy as np
import pytest

import lbfem_amd as A
import oracle as O


def sin2_field(X, Y):
    psi = np.sin(np.pi * X) ** 2 * np.sin(np.pi * Y) ** 2
    u = -2 * np.pi * np.sin(np.pi * X) ** 2 * np.sin(np.pi * Y) * np.cos(np.pi * Y)
    v = 2 * np.pi * np.sin(np.pi * X) * np.cos(np.pi * X) * np.sin(np.pi * Y) ** 2
    return psi, u, v


def biquadratic_field(X, Y):
    return X * (1 - X) * Y * (1 - Y), -X * (1 - X) * (1 - 2 * Y), (1 - 2 * X) * Y * (1 - Y)


def oracle_disc(N, p, stretch):
    d = O.Disc(N, p, case=O.CAVITY, stretch=stretch)
    xy = d.nodes()
    return d, xy[:, 0], xy[:, 1]


# ------------------------------------------------------------------ oracle pins (CPU)
@pytest.mark.parametrize("p,stretch", [(1, 0.0), (1, 1.2), (2, 0.0), (2, 1.2), (3, 0.0), (3, 1.2)])
def test_oracle_stream_function_exact_for_a_biquadratic_psi(p, stretch):
    d, X, Y = oracle_disc(8, p, stretch)
    psi_ex, u, v = biquadratic_field(X, Y)
    psi, its, st = d.stream_function(u, v)
    assert st == O.OK and its > 0
    assert np.abs(psi - psi_ex).max() < 1e-12 * np.abs(psi_ex).max(), np.abs(psi - psi_ex).max()


@pytest.mark.parametrize("p,stretch", [(1, 0.0), (1, 1.2), (2, 0.0), (2, 1.2)])
def test_oracle_stream_function_converges(p, stretch):
    errs = []
    for N in (16, 32):
        d, X, Y = oracle_disc(N, p, stretch)
        psi_ex, u, v = sin2_field(X, Y)
        psi, its, st = d.stream_function(u, v)
        assert st == O.OK and its > 0
        assert np.abs(psi[(X == 0) | (X == 1) | (Y == 0) | (Y == 1)]).max() == 0.0  # zero on the walls
        errs.append(np.abs(psi - psi_ex).max())
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < {1: 6e-3, 2: 2e-4}[p], errs  # the stretched mesh is coarser at the centre
    assert rate > {1: 1.9, 2: 3.5}[p], (errs, rate)  # Q1 second order; Q2 fourth at the nodes


def test_oracle_stream_function_rejects_a_periodic_box():
    d = O.Disc(8, 1, case=O.TGV)
    z = np.zeros(d.n)
    assert d.stream_function(z, z)[2] == O.EINVAL


# ------------------------------------------------------------------ argument checks (CPU: rejected before any device use)
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


# ------------------------------------------------------------------ the device solve against the oracle (GPU)
def _gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.mark.gpu
@pytest.mark.parametrize("p,stretch,N", [(1, 0.0, 32), (1, 1.2, 48), (2, 1.2, 24), (3, 1.2, 12)])
def test_stream_function_matches_oracle(p, stretch, N):
    _gpu()
    d, X, Y = oracle_disc(N, p, stretch)
    _, u, v = sin2_field(X, Y)
    # a part that is not a curl (the fit is then not exact): a gradient and a wall-crossing flow
    u = u + 0.3 * np.cos(np.pi * X) * Y + 0.1
    v = v + 0.2 * X * X * np.sin(2 * np.pi * Y)
    psi_o, its_o, st = d.stream_function(u, v)
    assert st == O.OK
    psi, its = A.stream_function(N, p, u, v, stretch=stretch)
    print(f"N {N} p {p} stretch {stretch}: its {its} (oracle {its_o})", flush=True)
    assert abs(its - its_o) <= 1, (its, its_o)
    assert np.abs(psi - psi_o).max() <= 1e-10 * np.abs(psi_o).max(), np.abs(psi - psi_o).max()


@pytest.mark.gpu
@pytest.mark.parametrize("p,stretch", [(1, 0.0), (1, 1.2), (2, 0.0), (2, 1.2)])
def test_stream_function_converges(p, stretch):
    _gpu()
    errs = []
    for N in (16, 32):
        _, X, Y = oracle_disc(N, p, stretch)
        psi_ex, u, v = sin2_field(X, Y)
        psi, its = A.stream_function(N, p, u, v, stretch=stretch)
        assert its > 0
        errs.append(np.abs(psi - psi_ex).max())
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < {1: 6e-3, 2: 2e-4}[p], errs
    assert rate > {1: 1.9, 2: 3.5}[p], (errs, rate)


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 2, 3])
def test_stream_function_exact_for_a_biquadratic_psi(p):
    _gpu()
    _, X, Y = oracle_disc(8, p, 1.2)
    psi_ex, u, v = biquadratic_field(X, Y)
    psi, _ = A.stream_function(8, p, u, v, stretch=1.2)
    assert np.abs(psi - psi_ex).max() < 1e-12 * np.abs(psi_ex).max()


@pytest.mark.gpu
def test_device_entry_on_a_cavity_discretization():
    """lbfem_disc_stream_function on device tensors of a cavity discretization: the host-array entry's
    result bit for bit, psi of the zero field is zero with no iterations, and a periodic box is refused."""
    torch = _gpu()
    tc = A.test_cases.LidDrivenCavity(0.1, stretch=1.2)
    disc = A.Discretization(tc, 40, 2)
    X, Y = disc.node[:, 0], disc.node[:, 1]
    _, u, v = sin2_field(X, Y)
    u = u + 0.1 * Y
    ud = torch.tensor(u, device="cuda", dtype=torch.float64)
    vd = torch.tensor(v, device="cuda", dtype=torch.float64)
    psi = torch.full_like(ud, 7.0)  # overwritten, not a start value
    its = A.disc_stream_function(disc, ud, vd, psi)
    psi_h, its_h = A.stream_function(40, 2, u, v, stretch=1.2)
    assert its == its_h and (psi.cpu().numpy() == psi_h).all()
    z = torch.zeros_like(ud)
    assert A.disc_stream_function(disc, z, z, psi) == 0 and psi.abs().max().item() == 0.0
    periodic = A.Discretization(A.test_cases.TaylorGreenVortex(0.1), 8, 1)
    zp = torch.zeros(periodic.n_nodes, device="cuda", dtype=torch.float64)
    with pytest.raises(A.LbfemError):
        A.disc_stream_function(periodic, zp, zp, zp)
