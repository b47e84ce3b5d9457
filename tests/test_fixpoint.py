"""Batched Newton fixpoints, checked by the network RHS (isfixpoint, linear_analysis.jl:8-15)."""
import numpy as np
import pytest


def test_swing_load_grid_construction(ND):
    gr = ND.swing_load_grid(800, seed=3)
    n = gr.g.nv()
    assert n > 600 and abs(gr.P.sum()) < 1e-12
    p = gr.flat_p()
    nw = gr.network()
    assert len(p) == nw.pdim and np.isfinite(p).all()
    assert nw.dim == n + gr.is_gen.sum()


@pytest.mark.gpu
def test_ring_fixpoint_known_answer(ND, dev):
    # test/initialization_test.jl:15-38: ring of 4, P = [1,-1,1,-1], K = 6 -> each line carries 0.5
    import torch
    g = ND.cycle_graph(4)
    fs = ND.FixpointSolver([g])
    th = torch.zeros(4, dtype=torch.float64, device=dev)
    P = torch.tensor([1.0, -1.0, 1.0, -1.0], dtype=torch.float64, device=dev)
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, P, torch.ones_like(P), torch.full((4,), 6.0, dtype=torch.float64, device=dev)))
    t = th.cpu().numpy()
    d = np.arcsin(0.5 / 6)
    assert st["retcode"][0] == 0 and st["residual"][0] <= 1e-10
    np.testing.assert_allclose(np.diff(t), [-d, d, -d], atol=1e-12)


@pytest.mark.gpu
def test_batched_grids_are_fixpoints(ND, dev):
    import torch
    grids = [ND.swing_load_grid(int(n), seed=100 + i) for i, n in enumerate(np.linspace(300, 5000, 10).astype(int))]
    fs = ND.FixpointSolver([gr.g for gr in grids])
    th = torch.zeros(int(fs.voff[-1]), dtype=torch.float64, device=dev)
    P = torch.from_numpy(np.concatenate([gr.P for gr in grids])).to(dev)
    S = torch.from_numpy(np.concatenate([gr.scale for gr in grids])).to(dev)
    K = torch.from_numpy(np.concatenate([gr.K for gr in grids])).to(dev)
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, P, S, K))
    ok = st["retcode"] == 0
    # a grid without a synchronous state (a line would need |flow| > K) must be reported,
    # never returned as converged
    assert ok.sum() >= 8, st
    assert (st["residual"][~ok] > 1e-10).all()
    t = th.cpu().numpy()
    for i, gr in enumerate(grids):
        if not ok[i]:
            continue
        ti = t[fs.voff[i]:fs.voff[i + 1]]
        assert ND.isfixpoint_residual(gr, ti, dev) <= 1e-10
        assert 1 <= st["newton_iters"][i] <= 20


@pytest.mark.gpu
def test_forcing_cap(ND, dev):
    """eta_max caps the inexact-Newton forcing term min(eta_max, residual): a looser cap
    takes fewer CG iterations per Newton step and more steps, and every grid still ends
    at the requested residual, in the nonlinear and in the linearized (DC) problem."""
    import torch
    grids = [ND.swing_load_grid(int(n), seed=400 + i) for i, n in enumerate([600, 1500, 3000])]
    fs = ND.FixpointSolver([gr.g for gr in grids])
    d = lambda xs: torch.from_numpy(np.concatenate(xs)).to(dev)
    P, S, K = d([gr.P for gr in grids]), d([gr.scale for gr in grids]), d([gr.K for gr in grids])
    for lin in (False, True):
        out = {}
        for eta in (1e-2, 0.5):
            th = torch.zeros(int(fs.voff[-1]), dtype=torch.float64, device=dev)
            st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, P, S, K, linearized=lin, eta_max=eta))
            assert (st["retcode"] == 0).all() and (st["residual"] <= 1e-10).all(), (lin, eta, st)
            out[eta] = st
        per_step = lambda st: st["cg_iters"].sum() / st["newton_iters"].sum()
        assert per_step(out[0.5]) < per_step(out[1e-2]), (lin, out)
        assert out[0.5]["newton_iters"].sum() > out[1e-2]["newton_iters"].sum(), (lin, out)


@pytest.mark.gpu
def test_linearization_at_a_fixpoint(ND, dev):
    # linearize_network / jacobian_eigenvals at a batched-Newton fixpoint of a swing/load
    # grid: the state is a fixpoint, one zero eigenvalue (global phase), the rest stable
    import torch
    LA = ND.linear_analysis
    grid = ND.swing_load_grid(150, seed=41)
    th = torch.zeros(grid.g.nv(), dtype=torch.float64, device=dev)
    fs = ND.FixpointSolver([grid.g])
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, torch.from_numpy(grid.P).to(dev),
                                                   torch.from_numpy(grid.scale).to(dev),
                                                   torch.from_numpy(grid.K).to(dev)))
    assert st["retcode"][0] == 0
    nw = grid.network()
    u = torch.from_numpy(grid.flat_u(th.cpu().numpy())).to(dev)
    p = torch.from_numpy(grid.flat_p()).to(dev)
    assert LA.isfixpoint(nw, u, p, tol=1e-9)
    ev = LA.jacobian_eigenvals(nw, u, p)
    scale = np.abs(ev).max()
    assert abs(ev[-1]) <= 1e-9 * scale                # the phase-shift mode
    assert ev[:-1].real.max() < 0                     # every other mode decays
    # a perturbed angle breaks the fixpoint
    u2 = u.clone()
    u2[0] += 0.1
    assert not LA.isfixpoint(nw, u2, p, tol=1e-9)


@pytest.mark.gpu
def test_dae_linearization_is_the_kron_reduction(ND, dev):
    # Loads with mass 0 are algebraic: linearize_network returns (M, A) with the device
    # Jacobian, reduce_dae eliminates the load angles.  Independently of both: with the
    # line Laplacian L (off-diagonal -K cos(theta_j - theta_i)), eliminating the loads
    # is the Kron reduction L_red = L_gg - L_gl L_ll^-1 L_lg, and the generators obey
    # d(delta) = omega, d(omega) = -(L_red delta + D omega) / M.
    import torch
    LA = ND.linear_analysis
    grid = ND.swing_load_grid(120, seed=43)
    n = grid.g.nv()
    th = torch.zeros(n, dtype=torch.float64, device=dev)
    fs = ND.FixpointSolver([grid.g])
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, torch.from_numpy(grid.P).to(dev),
                                                   torch.from_numpy(grid.scale).to(dev),
                                                   torch.from_numpy(grid.K).to(dev)))
    assert st["retcode"][0] == 0
    theta = th.cpu().numpy()
    nw = grid.network(algebraic_loads=True)
    u = torch.from_numpy(grid.flat_u(theta)).to(dev)
    p = torch.from_numpy(grid.flat_p()).to(dev)
    assert LA.isfixpoint(nw, u, p, tol=1e-9)
    sys = LA.linearize_network(nw, u, p)
    assert isinstance(sys.M, ND.Diagonal)
    assert sys.constraint_dim() == int((~grid.is_gen).sum())
    red = LA.reduce_dae(sys)

    s, d = grid.g.src - 1, grid.g.dst - 1
    w = grid.K * np.cos(theta[d] - theta[s])
    L = np.zeros((n, n))
    np.add.at(L, (s, d), -w)
    np.add.at(L, (d, s), -w)
    L[np.diag_indices(n)] = -L.sum(axis=1)
    gi, li = np.flatnonzero(grid.is_gen), np.flatnonzero(~grid.is_gen)
    Lred = L[np.ix_(gi, gi)] - L[np.ix_(gi, li)] @ np.linalg.solve(L[np.ix_(li, li)], L[np.ix_(li, gi)])
    col = {}
    for r, (_, vi, name) in enumerate(red.sym):
        col[(vi, name)] = r
    J = np.zeros_like(red.A)
    for a, i in enumerate(gi):
        v = int(i) + 1
        J[col[(v, "delta")], col[(v, "omega")]] = 1.0
        J[col[(v, "omega")], col[(v, "omega")]] = -grid.D[i] / grid.M[i]
        for b, j in enumerate(gi):
            J[col[(v, "omega")], col[(int(j) + 1, "delta")]] = -Lred[a, b] / grid.M[i]
    np.testing.assert_allclose(red.A, J, rtol=0, atol=1e-9 * np.abs(J).max())

    ev = LA.jacobian_eigenvals(sys)
    ref = np.linalg.eigvals(J)
    assert len(ev) == 2 * len(gi)
    left = list(ref)
    for x in ev:                                      # nearest-neighbour matching
        k = int(np.argmin(np.abs(np.asarray(left) - x)))
        assert abs(left[k] - x) <= 1e-7 * np.abs(ref).max()
        left.pop(k)
    assert abs(ev[-1]) <= 1e-9 * np.abs(ev).max()    # the phase-shift mode
    assert ev[:-1].real.max() < 0
    assert LA.is_linear_stable(nw, u, p, marginally_stable=True, atol=1e-9, tol=1e-9)
    pf = LA.participation_factors(sys)
    np.testing.assert_allclose(pf["pfactors"].sum(axis=0), 1.0, rtol=1e-9)
    assert pf["sym"] == red.sym


def _ring_network(ND, P, K=6.0):
    # test/initialization_test.jl:15-38: swing vertices (M = 1, D = 0.1) on
    # watts_strogatz(4, 2, 0.0) -- the 4-cycle -- with K = 6 lines
    g = ND.cycle_graph(4)
    nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    v, e = nw.ranges()
    p = np.zeros(nw.pdim)
    for i in range(4):
        p[v[i, 2, 0] - 1: v[i, 2, 1]] = [1.0, 0.1, P[i]]     # [M, D, Pm]
    p[e[:, 3, 0] - 1] = K
    return g, nw, p


@pytest.mark.gpu
def test_find_fixpoint_ring_known_answer(ND, dev):
    # P = [1, -1, 1, -1]: every line carries 1/2, so K sin(theta_src - theta_dst) = +-1/2
    g, nw, p = _ring_network(ND, [1.0, -1.0, 1.0, -1.0])
    x, info = ND.find_fixpoint(nw, np.zeros(nw.dim), p)
    assert info["residual"] <= 1e-10 and info["iters"] <= 10
    u = x.cpu().numpy()
    v, _ = nw.ranges()
    th, om = u[v[:, 0, 0] - 1], u[v[:, 0, 0]]
    np.testing.assert_allclose(om, 0.0, atol=1e-12)
    flows = 6.0 * np.sin(th[g.src - 1] - th[g.dst - 1])
    np.testing.assert_allclose(np.abs(flows), 0.5, atol=1e-11)
    np.testing.assert_allclose(np.abs(np.diff(th)), np.arcsin(0.5 / 6.0), atol=1e-11)
    import torch
    assert ND.linear_analysis.isfixpoint(nw, x, torch.from_numpy(p).to(dev))


@pytest.mark.gpu
@pytest.mark.parametrize("linsolve", ["dense", "sparse", "gmres"])
def test_find_fixpoint_dae_grid_matches_batched_solver(ND, dev, linsolve):
    # algebraic loads (mass 0), flat start; dense least squares, sparse LU of the colored
    # Jacobian and JVP-GMRES steps all reach the batched solver's fixpoint up to the
    # global phase
    import torch
    grid = ND.swing_load_grid(300, seed=47)
    n = grid.g.nv()
    th = torch.zeros(n, dtype=torch.float64, device=dev)
    fs = ND.FixpointSolver([grid.g])
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, torch.from_numpy(grid.P).to(dev),
                                                   torch.from_numpy(grid.scale).to(dev),
                                                   torch.from_numpy(grid.K).to(dev)))
    assert st["retcode"][0] == 0
    nw = grid.network(algebraic_loads=True)
    x, info = ND.find_fixpoint(nw, np.zeros(nw.dim), grid.flat_p(), linsolve=linsolve)
    assert info["residual"] <= 1e-10
    v, _ = nw.ranges()
    got = x.cpu().numpy()[v[:, 0, 0] - 1]
    ref = th.cpu().numpy()
    np.testing.assert_allclose(got - got.mean(), ref - ref.mean(), atol=1e-8)


@pytest.mark.gpu
def test_find_fixpoint_errors(ND, dev):
    g, nw, p = _ring_network(ND, [1.0, -1.0, 1.0, -1.0])
    x0 = np.zeros(nw.dim)
    x0[2] = np.nan
    with pytest.raises(ND.NetworkInitError, match=r"Variables:  VIndex\(2, :delta\)"):
        ND.find_fixpoint(nw, x0, p)
    p2 = p.copy()
    p2[-1] = np.nan
    with pytest.raises(ND.NetworkInitError, match=r"Parameters: EIndex\(4, :K\)"):
        ND.find_fixpoint(nw, np.zeros(nw.dim), p2)
    # injections beyond what K = 0.1 lines can carry: no synchronous state exists
    _, nw3, p3 = _ring_network(ND, [1.0, -1.0, 1.0, -1.0], K=0.1)
    with pytest.raises(ND.NetworkInitError, match="Could not find fixpoint"):
        ND.find_fixpoint(nw3, np.zeros(nw3.dim), p3, maxiters=30)
    with pytest.raises(ND.ArgumentError):
        ND.find_fixpoint(nw, np.zeros(nw.dim + 1), p)
