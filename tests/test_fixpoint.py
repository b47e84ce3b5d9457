"""Batched Newton fixpoints, judged by the network RHS (isfixpoint, linear_analysis.jl:8-15)."""
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
