"""Two networks of the reference's documentation examples, built from the library's models:

- docs/examples/heterogeneous_system.jl: an 8-vertex Kuramoto ring with one static vertex
  (output = its parameter, NoFeedForward) and one vertex with inertia;
- docs/examples/directed_and_weighted_graphs.jl: FitzHugh-Nagumo neurons on the directed,
  weighted 90-area brain atlas, coupled by electrical gap junctions.  The weight matrix is
  the example's own data file docs/examples/Norm_G_DTI.txt, kept as the fixture
  tests/golden/Norm_G_DTI.txt; edge i -> j carries G[i, j] (SimpleWeightedDiGraph(G)), in
  lexicographic edge order.

The oracle is pinned by each example's equations written out with numpy; the device
then follows the oracle.
"""
import os

import numpy as np
import pytest

from conftest import oracle_for, scaled_error

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _heterogeneous(ND):
    L = ND.Lib
    N = 8
    g = ND.watts_strogatz(N, 2, 0.0, seed=1)               # the ring of heterogeneous_system.jl:17
    omega = np.arange(1, N + 1) / N
    omega -= omega.sum() / N
    vm = [L.kuramoto_first() for _ in range(N)]
    vm[0] = L.static_vertex(theta_fix=omega[0])
    vm[4] = L.kuramoto_inertia()
    nw = ND.Network(g, vm, L.kuramoto_edge())
    p = nw.default_p()
    v, e = nw.ranges()
    for i in range(N):
        p[v[i, 2, 0] - 1] = omega[i]
    for k in range(g.ne()):
        p[e[k, 3, 0] - 1] = 3.0                            # K = 3 (heterogeneous_system.jl:24)
    x0 = np.arange(1, N + 1) / N
    x0 -= x0.sum() / N
    u = np.zeros(nw.dim)
    for i in range(1, N):
        u[v[i, 0, 0] - 1] = x0[i]
    u[v[4, 0, 0]] = 5.0                                    # omega of the inertia vertex
    return nw, g, u, p, omega


def _brain(ND):
    G = np.loadtxt(os.path.join(GOLDEN, "Norm_G_DTI.txt"), delimiter=",")
    src, dst = np.nonzero(G)                               # row-major: lexicographic (src, dst)
    g = ND.SimpleDiGraph(G.shape[0], src + 1, dst + 1)
    nw = ND.Network(g, ND.Lib.fhn_vertex(), ND.Lib.electrical_edge())
    p = nw.default_p()
    _, e = nw.ranges()
    p[e[:, 3, 0] - 1] = G[src, dst]                        # p.e[:, :weight] = edge_weights
    u = np.random.default_rng(42).normal(size=nw.dim) * 5
    return nw, g, u, p, G


def test_heterogeneous_example_equations(ND):
    nw, g, u, p, omega = _heterogeneous(ND)
    assert nw.dim == 8 and nw.pdim == 16
    v, _ = nw.ranges()
    theta = np.array([omega[0]] + [u[v[i, 0, 0] - 1] for i in range(1, 8)])
    flow = np.zeros(8)
    for s, d in zip(g.src - 1, g.dst - 1):
        y = 3.0 * np.sin(theta[s] - theta[d])
        flow[d] += y
        flow[s] -= y
    du = oracle_for(nw).rhs(u, p)
    for i in range(1, 8):
        if i == 4:                                         # inertia: dtheta = omega, domega = w0 - omega + P
            np.testing.assert_allclose(du[v[i, 0, 0] - 1: v[i, 0, 0] + 1],
                                       [5.0, omega[i] - 5.0 + flow[i]], rtol=1e-14, atol=1e-14)
        else:
            np.testing.assert_allclose(du[v[i, 0, 0] - 1], omega[i] + flow[i], rtol=1e-14, atol=1e-14)
    o, _ = oracle_for(nw).buffers(u, p)
    assert o[v[0, 1, 0] - 1] == omega[0]                   # the static vertex's output


def test_brain_example_equations(ND):
    nw, g, u, p, G = _brain(ND)
    n = G.shape[0]
    assert g.ne() == 7793 and nw.dim == 2 * n
    uu, vv = u[0::2], u[1::2]
    coup = np.array([sum(G[j, i] * (uu[j] - uu[i]) * 0.5 for j in range(n) if G[j, i] != 0) for i in range(n)])
    du = oracle_for(nw).rhs(u, p)
    np.testing.assert_allclose(du[0::2], uu - uu ** 3 / 3 - vv + coup, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(du[1::2], (uu - 0.5) * 0.05, rtol=1e-14, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["heterogeneous", "brain"])
def test_docs_examples_on_the_device(ND, dev, which):
    import torch
    nw, g, u, p, _ = (_heterogeneous if which == "heterogeneous" else _brain)(ND)
    on = oracle_for(nw)
    ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
    du = torch.full((nw.dim,), np.nan, dtype=torch.float64, device=dev)
    nw(du, ud, pd, 0.0)
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p), on.scale(u, p))
    assert err.max() <= 1e-12, err.max()
    # the neurons start at |u| up to ~15 (randn * 5), where -u^2 sets the stiffness: a
    # smaller step keeps RK4 inside its stability region
    dt = 0.01 if which == "heterogeneous" else 0.001
    ur = ud.clone()
    nw.rk4(ur, pd, 0.0, dt, 200)
    np.testing.assert_allclose(ur.cpu().numpy(), on.rk4(u, p, 0.0, dt, 200), rtol=1e-10, atol=1e-10)
