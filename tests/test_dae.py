"""Time integration with algebraic states: RK4 on the differential states with the
constraints solved at every stage.  Pinned by the exact solution of a linear diffusion
DAE (the Kron-reduced Laplacian flow exp(-L_red t)), by fourth-order convergence and by
a momentum balance on a swing network with algebraic loads."""
import numpy as np
import pytest


def _diffusion_dae(ND, n, seed):
    rng = np.random.default_rng(seed)
    g = ND.watts_strogatz(n, 4, 0.3, seed=seed)
    alg = rng.random(n) < 0.3
    dv = ND.VertexModel(f=ND.components.diffusionvertex, dim=1, g=1)
    av = ND.VertexModel(f=ND.components.diffusionvertex, dim=1, g=1, mass_matrix=0)
    nw = ND.Network(g, [av if a else dv for a in alg], ND.Lib.bench.diffusion_edge())
    return nw, g, alg, rng


@pytest.mark.gpu
def test_linear_diffusion_dae_matches_kron_reduced_flow(ND, dev):
    from scipy.linalg import expm
    nw, g, alg, rng = _diffusion_dae(ND, 60, 7)
    n = g.nv()
    L = np.zeros((n, n))
    s, d = g.src - 1, g.dst - 1
    np.add.at(L, (s, d), -1.0)
    np.add.at(L, (d, s), -1.0)
    L[np.diag_indices(n)] = -L.sum(axis=1)
    di, ai = np.flatnonzero(~alg), np.flatnonzero(alg)
    Lred = L[np.ix_(di, di)] - L[np.ix_(di, ai)] @ np.linalg.solve(L[np.ix_(ai, ai)], L[np.ix_(ai, di)])
    v, _ = nw.ranges()
    x0 = rng.normal(size=n)
    u0 = np.empty(nw.dim)
    u0[v[:, 0, 0] - 1] = x0
    u, info = ND.dae.rk4(nw, u0, np.zeros(0), 0.0, 0.005, 200)
    got = u.cpu().numpy()[v[:, 0, 0] - 1]
    xd = expm(-Lred) @ x0[di]
    xa = -np.linalg.solve(L[np.ix_(ai, ai)], L[np.ix_(ai, di)] @ xd)
    np.testing.assert_allclose(got[di], xd, atol=1e-9)
    np.testing.assert_allclose(got[ai], xa, atol=1e-9)
    assert info["newton_iters"] > 0


@pytest.mark.gpu
def test_swing_with_algebraic_loads_order_and_balance(ND, dev):
    import torch
    grid = ND.swing_load_grid(200, seed=61)
    nw = grid.network(algebraic_loads=True)
    p = grid.flat_p()
    rng = np.random.default_rng(3)
    n = grid.g.nv()
    th0 = rng.normal(scale=0.05, size=n)
    om0 = rng.normal(scale=0.1, size=n)
    u0 = grid.flat_u(th0, om0)
    runs = {}
    for dt in (0.02, 0.01, 0.005):
        u, _ = ND.dae.rk4(nw, u0, p, 0.0, dt, int(round(1.0 / dt)))
        runs[dt] = u.cpu().numpy()
    e1 = np.abs(runs[0.02] - runs[0.005]).max()
    e2 = np.abs(runs[0.01] - runs[0.005]).max()
    assert e1 / e2 > 10.0, (e1, e2)             # 4th order: (2^4 - 1)/(1 - 2^-4) ~ 17
    # the constraints hold and sum_g M_g dw_g = sum_all P - sum_g D_g w_g (the line flows cancel)
    uf = torch.from_numpy(runs[0.005]).to(dev)
    du = torch.empty_like(uf)
    nw(du, uf, torch.from_numpy(p).to(dev), 1.0)
    duh = du.cpu().numpy()
    v, _ = nw.ranges()
    gi, li = np.flatnonzero(grid.is_gen), np.flatnonzero(~grid.is_gen)
    assert np.abs(duh[v[li, 0, 0] - 1]).max() <= 1e-11
    w = runs[0.005][v[gi, 0, 0]]
    lhs = (grid.M[gi] * duh[v[gi, 0, 0]]).sum()
    rhs = grid.P.sum() - (grid.D[gi] * w).sum()
    assert abs(lhs - rhs) <= 1e-10 * (np.abs(grid.P).sum() + 1)


def test_dae_integrator_refuses_dense_mass_matrices(ND):
    f = ND.DeviceFunction("fv", "vertex_f", ND._abi.ND_VM_NOOP, None, 0)
    nw = ND.Network(ND.path_graph(3), ND.VertexModel(f=f, dim=2, g=1, mass_matrix=np.array([[1.0, 1.0], [0.0, 0.0]])),
                    ND.Lib.bench.diffusion_edge())
    with pytest.raises(ND.UnsupportedError):
        ND.dae.rk4(nw, np.zeros(nw.dim), np.zeros(0), 0.0, 0.1, 1)
