"""Warm-start network: CPU reference pinned by finite differences; GPU forward/backward
kernels against it; a trained network against the flat start (Newton iterations)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mpnn_ref  # noqa: E402


def _tiny_problem(n=40, nfeat=8, seed=0):
    import scipy.sparse as sp
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(n):
        for j in rng.choice(n, 3, replace=False):
            if i != j:
                rows += [i, j]
                cols += [j, i]
    M = sp.coo_matrix((rng.uniform(1, 2, len(rows)), (rows, cols)), shape=(n, n)).tocsr()
    M = (M + M.T) * 0.5
    M.sort_indices()
    deg = np.asarray(M.sum(axis=1)).ravel()
    A = sp.diags(1 / deg) @ M
    A = A.tocsr()
    A.sort_indices()
    npar = nfeat * 64 + 64 + 4 * (64 * 128 + 64) + 65
    params = rng.normal(0, 0.15, npar)
    X = rng.normal(size=(n, nfeat))
    y = rng.normal(size=n)
    return A, params, X, y


def test_reference_backward_matches_finite_differences():
    A, params, X, y = _tiny_problem()
    loss, g = mpnn_ref.loss_and_grad(params, X, A, y)
    rng = np.random.default_rng(1)
    idx = rng.choice(len(params), 60, replace=False)
    for i in idx:
        e = np.zeros_like(params)
        e[i] = 1e-6
        fd = (mpnn_ref.loss_and_grad(params + e, X, A, y)[0] - mpnn_ref.loss_and_grad(params - e, X, A, y)[0]) / 2e-6
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


@pytest.mark.gpu
@pytest.mark.parametrize("n", [300, 100_000])
def test_gpu_forward_backward_match_reference(ND, dev, n):
    # n = 100000 rows: past every 16-bit grid and copy-height limit; the workspace starts
    # as NaN so any entry the kernels fail to write shows up
    import torch
    A, params, X, y = _tiny_problem(n=n, seed=2)
    At = A.T.tocsr()
    At.sort_indices()
    # the kernels take A's pattern with transposed weights on the same pattern
    rp = A.indptr.astype(np.int32)
    nbr = A.indices.astype(np.int32)
    a_f = A.data.astype(np.float32)
    a_b = np.asarray(At[np.repeat(np.arange(A.shape[0]), np.diff(A.indptr)), A.indices]).ravel().astype(np.float32)
    p32 = params.astype(np.float32)
    L = ND._abi.lib()
    n = A.shape[0]
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    rp_d, nb_d, af_d, ab_d, X_d, p_d, y_d = d(rp), d(nbr), d(a_f), d(a_b), d(X.astype(np.float32)), d(p32), d(y.astype(np.float32))
    work = torch.full((int(L.nd_mp_workspace_floats(n)),), float("nan"), dtype=torch.float32, device=dev)
    out = torch.full((n,), float("nan"), dtype=torch.float32, device=dev)
    grads = torch.full_like(p_d, float("nan"))
    lp = torch.zeros(n // 4 + 2, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    ND._abi.check(L.nd_mp_forward(n, rp_d.data_ptr(), nb_d.data_ptr(), af_d.data_ptr(), X_d.data_ptr(), 8,
                                  p_d.data_ptr(), work.data_ptr(), out.data_ptr(), s))
    ND._abi.check(L.nd_mp_backward(n, rp_d.data_ptr(), nb_d.data_ptr(), ab_d.data_ptr(), X_d.data_ptr(), 8,
                                   p_d.data_ptr(), y_d.data_ptr(), work.data_ptr(), grads.data_ptr(), lp.data_ptr(), s))
    torch.cuda.synchronize()
    ref_out, _ = mpnn_ref.forward(p32.astype(np.float64), X.astype(np.float32).astype(np.float64), A)
    loss, gref = mpnn_ref.loss_and_grad(p32.astype(np.float64), X.astype(np.float32).astype(np.float64), A,
                                        y.astype(np.float32).astype(np.float64))
    np.testing.assert_allclose(out.cpu().numpy(), ref_out, rtol=2e-4, atol=2e-5)
    g = grads.cpu().numpy()
    assert np.linalg.norm(g - gref) <= 2e-4 * np.linalg.norm(gref)
    assert abs(lp.sum().item() / n - loss) <= 1e-4 * loss


def test_source_term_identity(ND):
    # L_K theta* = P + s(theta*) at any fixpoint: checked on the 4-node ring of
    # test/initialization_test.jl:15-38 style (balanced P, known fixpoint by Newton on the CPU)
    W = ND.warmstart
    g = ND.swing_load_grid(300, seed=11)
    n = g.g.nv()
    s, d = g.g.src - 1, g.g.dst - 1
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl

    def lap(w):
        A = sp.coo_matrix((np.r_[w, w], (np.r_[s, d], np.r_[d, s])), shape=(n, n)).tocsr()
        return (sp.diags(np.asarray(A.sum(1)).ravel()) - A).tocsc()

    def F(th):
        D = th[d] - th[s]
        f = g.K * np.sin(D)
        return g.P + np.bincount(s, f, n) - np.bincount(d, f, n)

    th = np.zeros(n)
    for _ in range(30):
        r = F(th)
        if np.abs(r).max() < 1e-13:
            break
        dth = np.zeros(n)
        dth[1:] = spl.spsolve(lap(g.K * np.cos(th[d] - th[s]))[1:, 1:], r[1:])
        th = th + dth
    assert np.abs(F(th)).max() < 1e-12
    lhs = lap(g.K) @ th
    np.testing.assert_allclose(lhs, g.P + W.source(g, th), rtol=0, atol=1e-10)


@pytest.mark.gpu
def test_trained_network_reduces_newton_iterations(ND, dev):
    import torch
    W = ND.warmstart
    train = [ND.swing_load_grid(int(n), seed=5000 + i) for i, n in enumerate(np.random.default_rng(0).integers(500, 1500, 48))]
    test = [ND.swing_load_grid(int(n), seed=9000 + i) for i, n in enumerate(np.random.default_rng(1).integers(500, 1500, 12))]
    th_star, st, _ = W.solve_grids(train, dev)
    ok = st["retcode"] == 0
    train = [g for g, o in zip(train, ok) if o]
    th_star = [W.center(t) for t, o in zip(th_star, ok) if o]
    dc, s0, X, _ = W.physics_inputs(train, dev)
    tgt = np.concatenate([W.source(g, t) - a for g, t, a in zip(train, th_star, s0)])
    net = W.WarmStartNet(dev, seed=0)
    Xa = np.concatenate(X)
    net.xmean, net.xstd = Xa.mean(axis=0), Xa.std(axis=0) + 1e-12
    net.tscale = 1.0 / tgt.std()
    batch = W.GraphBatch(train, dev)
    Xd = net.normalise(Xa)
    yd = torch.from_numpy((tgt * net.tscale).astype(np.float32)).to(dev)
    first = net.train_step(batch, Xd, yd, lr=2e-3).item()
    for _ in range(300):
        last = net.train_step(batch, Xd, yd, lr=2e-3)
    assert last.item() < 0.8 * first
    # held-out grids: flat start vs the learned warm start; every result must pass the acceptance test
    flat_th, flat_st, _ = W.solve_grids(test, dev)
    keep = flat_st["retcode"] == 0
    test = [g for g, o in zip(test, keep) if o]
    dce, s0e, Xe, _ = W.physics_inputs(test, dev)
    warm0, _ = W.warm_start(test, s0e, W.predict(net, test, Xe, dev), dev)
    warm_th, warm_st, _ = W.solve_grids(test, dev, theta0=np.concatenate(warm0))
    assert (warm_st["retcode"] == 0).all()
    for g, t in zip(test, warm_th):
        assert ND.isfixpoint_residual(g, t, dev) <= 1e-10
    assert warm_st["newton_iters"].mean() <= 0.75 * flat_st["newton_iters"][keep].mean()


@pytest.mark.gpu
def test_device_features_and_warm_start_match_host(ND, dev):
    # the GPU pipeline (GridBatch) computes the same source terms, features and warm
    # start as the host functions the network is trained with
    W = ND.warmstart
    grids = [ND.swing_load_grid(int(n), seed=7000 + i) for i, n in enumerate((400, 900, 1600))]
    gb = W.GridBatch(grids, dev)
    dc_host, s0_host, X_host, _ = W.physics_inputs(grids, dev)
    th_dc, _, _ = gb.solve(linearized=True)
    th_dc = gb.center(th_dc)
    np.testing.assert_allclose(th_dc.cpu().numpy(), np.concatenate(dc_host), rtol=0, atol=1e-12)
    s0, X = gb.features(th_dc)
    np.testing.assert_allclose(s0.cpu().numpy(), np.concatenate(s0_host), rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(X.cpu().numpy(), np.concatenate(X_host), rtol=1e-9, atol=1e-12)
    th_w, _ = W.warm_start_device(gb, None, tol=1e-10)
    th_w_host, _ = W.warm_start(grids, s0_host, None, dev)
    np.testing.assert_allclose(th_w.cpu().numpy(), np.concatenate(th_w_host), rtol=0, atol=1e-9)


@pytest.mark.gpu
def test_learned_start_falls_back_to_the_flat_start(ND, dev):
    """solve_with_fallback: grids whose start fails (here a NaN prediction, as a diverged
    network would give) are solved again from the flat start and then equal the flat
    solve bit for bit; the others keep their warm-started solution; every result passes
    the network-RHS residual check; a grid without a fixpoint (injections far beyond its lines'
    capacity) raises NetworkInitError."""
    import copy
    import torch
    from networkdynamics_jl_amd import warmstart as W
    from networkdynamics_jl_amd.fixpoint import isfixpoint_residual, swing_load_grid
    grids = [swing_load_grid(int(n), seed=900 + i) for i, n in enumerate([300, 450, 600, 350, 500, 400])]
    gb = W.GridBatch(grids, dev)
    start, _ = W.warm_start_device(gb, None)
    bad = start.clone()
    broken = np.array([False, True, False, False, True, False])
    for i in np.flatnonzero(broken):
        bad[int(gb.voff[i]):int(gb.voff[i + 1])] = float("nan")
    th, st, fell = W.solve_with_fallback(gb, bad)
    assert np.array_equal(fell, broken)
    assert (st["retcode"] == 0).all() and (st["residual"] <= 1e-10).all()
    flat, fst, _ = gb.solve()
    warm, wst, _ = gb.solve(theta0=start)
    thh, flh, wmh = th.cpu().numpy(), flat.cpu().numpy(), warm.cpu().numpy()
    for i, gr in enumerate(grids):
        sl = slice(int(gb.voff[i]), int(gb.voff[i + 1]))
        assert np.array_equal(thh[sl], flh[sl] if broken[i] else wmh[sl]), i
        assert st["newton_iters"][i] == (fst if broken[i] else wst)["newton_iters"][i]
        assert isfixpoint_residual(gr, thh[sl], dev) <= 1e-10
    # an overloaded grid: the flat start fails too
    over = copy.copy(grids[2])
    over.P = grids[2].P * 200.0
    gb2 = W.GridBatch([grids[0], over], dev)
    with pytest.raises(ND.NetworkInitError, match="grid indices \\[1\\]"):
        W.solve_with_fallback(gb2, torch.zeros(gb2.n, dtype=torch.float64, device=dev) + float("nan"))


@pytest.mark.gpu
def test_pipeline_at_c5_grid_sizes(ND, dev):
    """The whole C5 pipeline (experiments.c5_warmstart, which bench.py runs at 20,000 /
    1,000 grids) at its grid-size distribution N ~ U{500..5000}, on 120 training and 24
    held-out grids (generated at 680-4943 vertices, 20 of them above 1500 and 5 above
    4000; those without a flat-start fixpoint are dropped) and a 3-epoch network:
    every start (flat, DC, physics, learned, and learned_solve's pipeline with its
    flat fallback) must end at a fixpoint that the network RHS accepts (max |du| <=
    1e-10), the physics-informed starts must need fewer Newton iterations than the flat
    start, and the reported ratios must be the quotients of the reported means."""
    from networkdynamics_jl_amd.experiments import c5_warmstart
    out = c5_warmstart(dev, ntrain=120, neval=24, epochs=3, batch_grids=40, seed=7, workers=1)
    n = out["eval_grids"]
    assert n + out["eval_dropped"] == 24 and n >= 16
    assert out["accepted_flat"] == n
    for name in ("dc", "physics", "learned"):
        assert out[f"{name}_accepted"] == n, name
        assert out[f"{name}_retcodes"][0] == n, name
    assert out["pipeline_accepted"] == n
    assert out["loss_last"] < out["loss_first"]
    assert out["dc_newton_mean"] < out["flat_newton_mean"]
    assert out["physics_newton_mean"] < out["flat_newton_mean"]
    assert out["learned_newton_mean"] < out["flat_newton_mean"]
    assert out["newton_ratio_learned"] == pytest.approx(out["flat_newton_mean"] / out["learned_newton_mean"])


@pytest.mark.gpu
def test_start_residual(ND, dev):
    """GridBatch.start_residual (nd_ws_residual, then a per-grid max) is the fixpoint
    residual max_i |F_i / scale_i|, F_i = P_i + sum_j K_ij sin(theta_j - theta_i): exactly
    max |P / scale| at the flat start, below the tolerance at the solver's fixpoints, and
    numpy's at a random state."""
    import torch
    from networkdynamics_jl_amd import warmstart as W
    from networkdynamics_jl_amd.fixpoint import swing_load_grid
    grids = [swing_load_grid(int(n), seed=1300 + i) for i, n in enumerate([400, 700, 550])]
    gb = W.GridBatch(grids, dev)
    z = torch.zeros(gb.n, dtype=torch.float64, device=dev)
    assert np.array_equal(gb.start_residual(z), gb.flat_residual)
    flat, fst, _ = gb.solve()
    assert (fst["retcode"] == 0).all()
    assert (gb.start_residual(flat) <= 1e-9).all()
    rng = np.random.default_rng(13)
    th = rng.uniform(-np.pi, np.pi, gb.n)
    r = gb.start_residual(torch.from_numpy(th).to(dev))
    for i, gr in enumerate(grids):
        t = th[gb.voff[i]:gb.voff[i + 1]]
        s, d = gr.g.src - 1, gr.g.dst - 1
        F = gr.P.copy()
        np.add.at(F, s, gr.K * np.sin(t[d] - t[s]))
        np.add.at(F, d, gr.K * np.sin(t[s] - t[d]))
        np.testing.assert_allclose(r[i], np.max(np.abs(F / gr.scale)), rtol=1e-12)


@pytest.mark.gpu
def test_network_entry_points_refuse_bad_arguments(ND, dev):
    """The warm-start entry points check their arguments on the host before any launch:
    a params or workspace pointer that is not 16-byte aligned (the products and the
    aggregation load 16 bytes at a time), a null pointer, an empty batch."""
    import torch
    L = ND._abi.lib()
    chk = ND._abi.check
    n = 64
    rp = torch.arange(0, n + 1, dtype=torch.int32, device=dev)
    nbr = torch.arange(n, dtype=torch.int32, device=dev)
    a = torch.ones(n, dtype=torch.float32, device=dev)
    X = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    params = torch.zeros(int(L.nd_mp_param_count(8)) + 4, dtype=torch.float32, device=dev)
    work = torch.zeros(int(L.nd_mp_workspace_floats(n)) + 4, dtype=torch.float32, device=dev)
    out = torch.empty(n, dtype=torch.float32, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    fwd = lambda pp, ww, nn=n: chk(L.nd_mp_forward(nn, rp.data_ptr(), nbr.data_ptr(), a.data_ptr(), X.data_ptr(), 8,
                                                   pp, ww, out.data_ptr(), s))
    fwd(params.data_ptr(), work.data_ptr())
    torch.cuda.synchronize()
    for pp, ww in ((params.data_ptr() + 4, work.data_ptr()), (params.data_ptr(), work.data_ptr() + 8)):
        with pytest.raises(ND.ArgumentError, match="16-byte aligned"):
            fwd(pp, ww)
    with pytest.raises(ND.ArgumentError):
        fwd(params.data_ptr(), work.data_ptr(), 0)
    th = torch.zeros(n, dtype=torch.float64, device=dev)
    with pytest.raises(ND.ArgumentError):
        chk(L.nd_ws_features(n, rp.data_ptr(), nbr.data_ptr(), None, th.data_ptr(), th.data_ptr(), th.data_ptr(),
                             th.data_ptr(), th.data_ptr(), s))
    with pytest.raises(ND.ArgumentError):
        chk(L.nd_ws_residual(0, rp.data_ptr(), nbr.data_ptr(), th.data_ptr(), th.data_ptr(), th.data_ptr(),
                             th.data_ptr(), th.data_ptr(), s))


def _ws_features(ND, dev, rows, theta, P, gen):
    """nd_ws_features on a CSR given as per-node lists of (neighbour, weight)"""
    import torch
    L = ND._abi.lib()
    n = len(rows)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    nbr = np.array([j for r in rows for j, _ in r] or [0], np.int32)
    ww = np.array([x for r in rows for _, x in r] or [0.0])
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    d_rp, d_nbr, d_w, d_th, d_P, d_gen = T(rp), T(nbr), T(ww), T(theta), T(P), T(gen)
    s0 = torch.empty(n, dtype=torch.float64, device=dev)
    X = torch.empty((n, 8), dtype=torch.float64, device=dev)
    scale = torch.ones(n, dtype=torch.float64, device=dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)
    s = torch.cuda.current_stream(dev).cuda_stream
    ND._abi.check(L.nd_ws_features(n, d_rp.data_ptr(), d_nbr.data_ptr(), d_w.data_ptr(), d_th.data_ptr(),
                                   d_P.data_ptr(), d_gen.data_ptr(), s0.data_ptr(), X.data_ptr(), s))
    ND._abi.check(L.nd_ws_residual(n, d_rp.data_ptr(), d_nbr.data_ptr(), d_w.data_ptr(), d_th.data_ptr(),
                                   d_P.data_ptr(), scale.data_ptr(), r.data_ptr(), s))
    torch.cuda.synchronize()
    return s0.cpu().numpy(), X.cpu().numpy(), r.cpu().numpy()


@pytest.mark.gpu
def test_features_of_isolated_nodes(ND, dev):
    """Nodes without entries next to a connected part: the other nodes' features, sources
    and residuals are those of the graph without them, bit for bit, and an isolated node's
    row holds P and the generator flag with zeros elsewhere (log degree and neighbour mean
    included), its residual |P| / scale."""
    K = {(0, 1): 1.5, (1, 2): 0.7, (0, 2): 2.0, (2, 3): 1.1, (3, 4): 0.9}
    rng = np.random.default_rng(5)

    def rows_of(n, relabel):
        rows = [[] for _ in range(n)]
        for (a, b), k in K.items():
            rows[relabel[a]].append((relabel[b], k))
            rows[relabel[b]].append((relabel[a], k))
        return [sorted(r) for r in rows]

    th5, P5, g5 = rng.uniform(-0.5, 0.5, 5), rng.normal(size=5), (rng.uniform(size=5) < 0.5).astype(float)
    sA, XA, rA = _ws_features(ND, dev, rows_of(5, list(range(5))), th5, P5, g5)
    new = [0, 2, 3, 4, 6]                       # isolated nodes at 1 and 5
    iso = [1, 5]
    th7, P7, g7 = np.full(7, 0.3), np.zeros(7), np.zeros(7)
    th7[new], P7[new], g7[new] = th5, P5, g5
    P7[iso], g7[iso] = [0.25, -1.5], [1.0, 0.0]
    sB, XB, rB = _ws_features(ND, dev, rows_of(7, new), th7, P7, g7)
    assert np.all(np.isfinite(XB)) and np.all(np.isfinite(sB))
    np.testing.assert_array_equal(XB[new], XA)
    np.testing.assert_array_equal(sB[new], sA)
    np.testing.assert_array_equal(rB[new], rA)
    for i in iso:
        np.testing.assert_array_equal(XB[i], [P7[i], g7[i], 0, 0, 0, 0, 0, 0])
        assert sB[i] == 0.0 and rB[i] == abs(P7[i])


def test_host_features_of_isolated_vertex(ND):
    """The host features (the training data's path) on a grid with a vertex without lines:
    the other vertices' rows equal those of the grid without it, bit for bit, and its row
    holds P and the generator flag with zeros elsewhere."""
    from types import SimpleNamespace
    from networkdynamics_jl_amd import warmstart as W

    def grid(nv, src, dst, P, gen):
        g = SimpleNamespace(src=np.array(src), dst=np.array(dst), nv=lambda: nv)
        return SimpleNamespace(g=g, K=np.array([1.5, 0.7, 2.0, 1.1]), P=np.array(P), is_gen=np.array(gen))

    th = np.array([0.1, -0.2, 0.05, 0.3])
    a = grid(4, [1, 2, 1, 3], [2, 3, 3, 4], [1.0, -0.5, -0.25, -0.25], [1, 0, 0, 0])
    b = grid(5, [1, 2, 1, 3], [2, 3, 3, 5], [1.0, -0.5, -0.25, 0.4, -0.25], [1, 0, 0, 1, 0])
    XA = W.features(a, th)
    XB = W.features(b, np.array([0.1, -0.2, 0.05, 0.7, 0.3]))
    assert np.all(np.isfinite(XB))
    np.testing.assert_array_equal(XB[[0, 1, 2, 4]], XA)
    np.testing.assert_array_equal(XB[3], [0.4, 1.0, 0, 0, 0, 0, 0, 0])
