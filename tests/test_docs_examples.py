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


def _truss(ND):
    """stress_on_truss.jl:88-150: N = 5 bays, vertices 1 and 4 fixed, vertex 11 heavier.
    The example keeps positions as Float32 points; here the same coordinates in Float64."""
    L = ND.Lib
    N = 5
    es = set()
    for i in range(1, N + 1):
        es.add((i, i + N))
        if i < N:
            es |= {(i + 1, i + N), (i, i + 1), (i + N, i + N + 1)}
    es.add((2 * N, 2 * N + 1))
    es = sorted(tuple(sorted(x)) for x in es)
    g = ND.SimpleGraph(2 * N + 1, [a for a, _ in es], [b for _, b in es])
    pos0 = np.array([[i, 0.0] for i in range(N)] + [[(i + 1) + 0.2, 1.0] for i in range(N)] + [[N + 1.0, -1.0]])
    fixed = {0, 3}
    vm = [L.truss_fixed_vertex() if i in fixed else L.truss_free_vertex() for i in range(2 * N + 1)]
    nw = ND.Network(g, vm, L.beam())
    v, e = nw.ranges()
    p = nw.default_p()
    u = np.zeros(nw.dim)
    for i in range(2 * N + 1):
        if i in fixed:
            p[v[i, 2, 0] - 1: v[i, 2, 0] + 1] = pos0[i]
        else:
            u[v[i, 0, 0] + 1: v[i, 0, 0] + 3] = pos0[i]                   # x, y
    p[v[10, 2, 0] - 1: v[10, 2, 0] + 1] = [200.0, 100.0]                 # M, gamma of vertex 11
    for k, (a, b) in enumerate(zip(g.src - 1, g.dst - 1)):
        p[e[k, 3, 0]] = np.linalg.norm(pos0[a] - pos0[b])                # L
    rng = np.random.default_rng(5)
    for i in range(2 * N + 1):                                           # move the free nodes a little
        if i not in fixed:
            u[v[i, 0, 0] - 1: v[i, 0, 0] + 3] += rng.normal(scale=[0.1, 0.1, 0.01, 0.01])
    return nw, g, u, p, fixed


def test_truss_example_equations(ND):
    nw, g, u, p, fixed = _truss(ND)
    v, e = nw.ranges()
    n = g.nv()
    pos = np.zeros((n, 2))
    for i in range(n):
        pos[i] = p[v[i, 2, 0] - 1: v[i, 2, 0] + 1] if i in fixed else u[v[i, 0, 0] + 1: v[i, 0, 0] + 3]
    F = np.zeros((n, 2))
    for k, (a, b) in enumerate(zip(g.src - 1, g.dst - 1)):
        K, Lk = p[e[k, 3, 0] - 1], p[e[k, 3, 0]]
        dxy = pos[b] - pos[a]
        d = np.sqrt(dxy[0] * dxy[0] + dxy[1] * dxy[1])
        f = K * (Lk - d) * dxy / d
        F[b] += f
        F[a] -= f
    du = oracle_for(nw).rhs(u, p)
    for i in range(n):
        if i in fixed:
            continue
        M, gam, gr = p[v[i, 2, 0] - 1: v[i, 2, 0] + 2]
        x0 = v[i, 0, 0] - 1
        vel = u[x0: x0 + 2]
        want = np.r_[(F[i] - gam * vel) / M - [0.0, gr], vel]
        np.testing.assert_allclose(du[x0: x0 + 4], want, rtol=1e-9, atol=1e-9 * np.abs(F).max())
    o, agg = oracle_for(nw).buffers(u, p)
    for i in fixed:
        np.testing.assert_array_equal(o[v[i, 1, 0] - 1: v[i, 1, 0] + 1], pos[i])
    for i in range(n):                                   # each vertex's aggregation range
        np.testing.assert_allclose(agg[v[i, 3, 0] - 1: v[i, 3, 0] + 1], F[i], rtol=1e-9, atol=1e-9 * np.abs(F).max())


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
@pytest.mark.parametrize("which", ["heterogeneous", "brain", "truss"])
def test_docs_examples_on_the_device(ND, dev, which):
    import torch
    nw, g, u, p, _ = dict(heterogeneous=_heterogeneous, brain=_brain, truss=_truss)[which](ND)
    on = oracle_for(nw)
    ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
    du = torch.full((nw.dim,), np.nan, dtype=torch.float64, device=dev)
    nw(du, ud, pd, 0.0)
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p), on.scale(u, p))
    assert err.max() <= 1e-12, err.max()
    # the neurons start at |u| up to ~15 (randn * 5), where -u^2 sets the stiffness: a
    # smaller step keeps RK4 inside its stability region
    dt = 0.01 if which == "heterogeneous" else (0.001 if which == "brain" else 1e-4)
    ur = ud.clone()
    nw.rk4(ur, pd, 0.0, dt, 200)
    np.testing.assert_allclose(ur.cpu().numpy(), on.rk4(u, p, 0.0, dt, 200), rtol=1e-10, atol=1e-10)
    # buffers and JVPs (the truss: vector edge outputs, computed vertex outputs at depth 2)
    from conftest import fd_jvp
    o, agg = nw.get_buffers(ud, pd)
    o_ref, agg_ref = on.buffers(u, p)
    np.testing.assert_allclose(o.cpu().numpy(), o_ref, rtol=1e-13, atol=1e-13 * np.abs(o_ref).max())
    np.testing.assert_allclose(agg.cpu().numpy(), agg_ref, rtol=1e-12, atol=1e-12 * np.abs(agg_ref).max())
    vv = np.random.default_rng(3).normal(size=nw.dim)
    jv = torch.empty_like(du)
    nw.jvp(du, jv, ud, torch.from_numpy(vv).to(dev), pd, 0.0)
    ref = fd_jvp(on, u, vv, p, h=1e-4)
    np.testing.assert_allclose(jv.cpu().numpy(), ref, rtol=1e-6, atol=1e-6 * max(1.0, np.abs(ref).max()))
