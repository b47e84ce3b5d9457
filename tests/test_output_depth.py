"""Networks whose vertices and edges carry d > 1 output components: the second half of
docs/examples/getting_started_with_network_dynamics.jl:218-291 -- two independent
diffusions (x and phi) on one graph, vertices dv .= esum with outputs 1:2, edges
e .= v_s .- v_d, AntiSymmetric.

The oracle is pinned by the example's own point: each component evolves as the
one-component diffusion network on the same graph, bit for bit, and as du = -L x.
"""
import numpy as np
import pytest

from conftest import fd_jvp, oracle_for, scaled_error


def _nets(ND, d, n=40, seed=1):
    L = ND.Lib
    g = ND.watts_strogatz(n, 4, 0.4, seed=seed)
    nw = ND.Network(g, L.diffusion_vertex_multi(d), L.diffusion_edge_multi(d))
    nw1 = ND.Network(g, L.diffusion_vertex(), L.bench.diffusion_edge())
    return g, nw, nw1


@pytest.mark.parametrize("d", [2, 3])
def test_components_evolve_as_independent_diffusions(ND, d):
    g, nw, nw1 = _nets(ND, d)
    assert nw.info.vdepth == d and nw.info.edepth == d and nw.dim == d * g.nv()
    rng = np.random.default_rng(d)
    u = rng.normal(size=nw.dim)
    on, on1 = oracle_for(nw), oracle_for(nw1)
    du = on.rhs(u, np.zeros(0))
    n = g.nv()
    Lap = np.zeros((n, n))
    for s, t in zip(g.src - 1, g.dst - 1):
        Lap[s, s] += 1
        Lap[t, t] += 1
        Lap[s, t] -= 1
        Lap[t, s] -= 1
    for j in range(d):
        np.testing.assert_array_equal(du[j::d], on1.rhs(u[j::d].copy(), np.zeros(0)))
        np.testing.assert_allclose(du[j::d], -Lap @ u[j::d], rtol=1e-13, atol=1e-13)
    r = on.rk4(u, np.zeros(0), 0.0, 0.01, 100)
    for j in range(d):
        np.testing.assert_array_equal(r[j::d], on1.rk4(u[j::d].copy(), np.zeros(0), 0.0, 0.01, 100))
    # the structural Jacobian pattern holds every entry that moves
    P = ND.sparsity.jac_prototype(nw).toarray()
    J = np.empty((nw.dim, nw.dim))
    for k in range(nw.dim):
        e = np.zeros(nw.dim)
        e[k] = 1e-6
        J[:, k] = (on.rhs(u + e, np.zeros(0)) - on.rhs(u - e, np.zeros(0))) / 2e-6
    assert not np.argwhere((np.abs(J) > 1e-8) & ~P).size


@pytest.mark.gpu
@pytest.mark.parametrize("d", [2, 3])
def test_output_depth_on_the_device(ND, dev, d):
    import torch
    g, nw, _ = _nets(ND, d, n=3000, seed=d)
    rng = np.random.default_rng(10 + d)
    u = rng.normal(size=nw.dim)
    p0 = np.zeros(0)
    on = oracle_for(nw)
    ud, pd = torch.from_numpy(u).to(dev), torch.zeros(1, dtype=torch.float64, device=dev)
    du = torch.full((nw.dim,), np.nan, dtype=torch.float64, device=dev)
    nw(du, ud, pd, 0.0)
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p0), on.scale(u, p0))
    assert err.max() <= 1e-12, err.max()
    # buffers, JVP, RK4, the perturbed loop without channels
    o, agg = nw.get_buffers(ud, pd)
    o_ref, agg_ref = on.buffers(u, p0)
    np.testing.assert_array_equal(o.cpu().numpy(), o_ref)
    np.testing.assert_allclose(agg.cpu().numpy(), agg_ref, rtol=1e-13, atol=1e-13)
    v = rng.normal(size=nw.dim)
    jv = torch.empty_like(du)
    nw.jvp(du, jv, ud, torch.from_numpy(v).to(dev), pd, 0.0)
    np.testing.assert_allclose(jv.cpu().numpy(), fd_jvp(on, u, v, p0), rtol=1e-8, atol=1e-8)
    ur = ud.clone()
    nw.rk4(ur, pd, 0.0, 0.01, 100)
    np.testing.assert_allclose(ur.cpu().numpy(), on.rk4(u, p0, 0.0, 0.01, 100), rtol=0, atol=1e-11)
    nw(du, ud, pd, 0.0, perturb=np.zeros(0), perturb_maps={})
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p0), on.scale(u, p0))
    assert err.max() <= 1e-12, err.max()


@pytest.mark.gpu
def test_output_depth_refuses_other_models(ND, dev):
    import torch
    L = ND.Lib
    g = ND.SimpleGraph(3, [1, 2], [2, 3])
    # a Kuramoto edge on two-component outputs is not componentwise
    nw = ND.Network(g, L.diffusion_vertex_multi(2), ND.EdgeModel(g=ND.AntiSymmetric(ND.components.kuramoto_edge_g),
                                                                 outdim=2, psym=["K"]))
    with pytest.raises(ND.UnsupportedError, match="componentwise"):
        nw(torch.empty(6, dtype=torch.float64, device=dev), torch.zeros(6, dtype=torch.float64, device=dev),
           torch.ones(2, dtype=torch.float64, device=dev), 0.0)


def _with_noop_edges(ND, n=400, seed=4):
    """Depth-2 diffusion with every fifth edge a NoOp edge (test/aggregators_test.jl:20): a
    NoOp edge writes no output, so the rows it reaches are NaN on both sides."""
    L = ND.Lib
    g = ND.watts_strogatz(n, 4, 0.4, seed=seed)
    noop = ND.EdgeModel(g=ND.AntiSymmetric(ND.components.noop_edge), outdim=2)
    em = [noop if k % 5 == 0 else L.diffusion_edge_multi(2) for k in range(g.ne())]
    return g, ND.Network(g, L.diffusion_vertex_multi(2), em)


def test_noop_edges_at_depth_two_oracle(ND):
    g, nw = _with_noop_edges(ND)
    u = np.random.default_rng(1).normal(size=nw.dim)
    du = oracle_for(nw).rhs(u, np.zeros(0))
    hit = np.zeros(g.nv(), bool)
    for k, (s, t) in enumerate(zip(g.src - 1, g.dst - 1)):
        if k % 5 == 0:
            hit[s] = hit[t] = True
    assert hit.any() and not hit.all()
    for j in range(2):
        assert np.array_equal(np.isnan(du[j::2]), hit)


@pytest.mark.gpu
def test_noop_edges_at_depth_two_on_the_device(ND, dev):
    import torch
    g, nw = _with_noop_edges(ND)
    rng = np.random.default_rng(2)
    u = rng.normal(size=nw.dim)
    p0 = np.zeros(0)
    on = oracle_for(nw)
    ud, pd = torch.from_numpy(u).to(dev), torch.zeros(1, dtype=torch.float64, device=dev)
    du = torch.full((nw.dim,), 7.0, dtype=torch.float64, device=dev)
    nw(du, ud, pd, 0.0)
    ref = on.rhs(u, p0)
    got = du.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    ok = ~np.isnan(ref)
    err = scaled_error(got[ok], ref[ok], on.scale(u, p0)[ok])
    assert err.max() <= 1e-12, err.max()
    v = rng.normal(size=nw.dim)
    jv = torch.full_like(du, 7.0)
    nw.jvp(du, jv, ud, torch.from_numpy(v).to(dev), pd, 0.0)
    jr = fd_jvp(on, u, v, p0)
    assert np.array_equal(np.isnan(jv.cpu().numpy()), np.isnan(jr))
    np.testing.assert_allclose(jv.cpu().numpy()[ok], jr[ok], rtol=1e-8, atol=1e-8)
