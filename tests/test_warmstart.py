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
    th_dc, _, _ = W.solve_grids(train, dev, linearized=True)
    th_dc = [W.center(t) for t in th_dc]
    net = W.WarmStartNet(dev, seed=0)
    X = np.concatenate([W.features(g, t) for g, t in zip(train, th_dc)])
    net.xmean, net.xstd = X.mean(axis=0), X.std(axis=0) + 1e-12
    tgt = np.concatenate([s - d for s, d in zip(th_star, th_dc)])
    net.tscale = 1.0 / tgt.std()
    batch = W.GraphBatch(train, dev)
    Xd = net.normalise(X)
    yd = torch.from_numpy((tgt * net.tscale).astype(np.float32)).to(dev)
    first = net.train_step(batch, Xd, yd, lr=2e-3).item()
    for _ in range(300):
        last = net.train_step(batch, Xd, yd, lr=2e-3)
    assert last.item() < 0.5 * first
    # held-out grids: flat start vs DC + network
    flat_th, flat_st, _ = W.solve_grids(test, dev)
    keep = flat_st["retcode"] == 0
    test = [g for g, o in zip(test, keep) if o]
    dc, _, _ = W.solve_grids(test, dev, linearized=True)
    warm0 = W.predict(net, test, [W.center(t) for t in dc], dev)
    warm_th, warm_st, _ = W.solve_grids(test, dev, theta0=np.concatenate(warm0))
    assert (warm_st["retcode"] == 0).all()
    for g, t in zip(test, warm_th):
        assert ND.isfixpoint_residual(g, t, dev) <= 1e-10
    assert warm_st["newton_iters"].mean() < flat_st["newton_iters"][keep].mean()
