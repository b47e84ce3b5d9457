"""Structural Jacobian sparsity, column coloring and colored sparse Jacobians.

The pattern is checked against the oracle's finite-difference Jacobian (every entry
that moves must be in the pattern), the coloring against its definition (no row sees a
color twice), and on the GPU the colored Jacobian against the dense one built column by
column from the same JVP kernel."""
import numpy as np
import pytest

from conftest import oracle_for
from test_external_inputs import _random_pairs
from test_oracle import _complete4, _random_mixed


def _fd_jacobian(on, u, p, h=1e-6):
    n = len(u)
    J = np.empty((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = h
        J[:, j] = (on.rhs(u + e, p) - on.rhs(u - e, p)) / (2 * h)
    return J


def _cases(ND):
    nw4 = _complete4(ND)
    rng = np.random.default_rng(1)
    yield "complete4", nw4, rng.normal(size=nw4.dim), rng.uniform(0.5, 2.0, nw4.pdim)
    nw, u, p = _random_mixed(ND, n=60, seed=5)
    yield "mixed", nw, u, p
    nw, p, u = ND.workloads.swing_ba(80, 2, seed=2)
    yield "swing_ba", nw, u, p
    nw, u, p = _random_pairs(ND, 25, 3)
    yield "external_inputs", nw, u, p


def test_pattern_covers_the_numerical_jacobian(ND):
    for name, nw, u, p in _cases(ND):
        P = ND.sparsity.jac_prototype(nw).toarray()
        J = _fd_jacobian(oracle_for(nw), u, p)
        moved = np.abs(J) > 1e-7 * max(1.0, np.abs(J).max())
        missing = np.argwhere(moved & ~P)
        assert len(missing) == 0, (name, missing[:5])
        # and it is far from dense
        assert P.sum() < 0.5 * P.size or nw.dim < 20, name


def test_column_coloring_is_valid(ND):
    for name, nw, u, p in _cases(ND):
        P = ND.sparsity.jac_prototype(nw)
        colors, ncolor = ND.sparsity.color_columns(P)
        assert colors.min() >= 0 and colors.max() == ncolor - 1
        Pr = P.tocsr()
        for r in range(P.shape[0]):
            c = colors[Pr.indices[Pr.indptr[r]:Pr.indptr[r + 1]]]
            assert len(np.unique(c)) == len(c), (name, r)
        assert ncolor <= np.diff(Pr.indptr).max() ** 2 + 1
    # a grid-like network needs few colors
    grid = ND.swing_load_grid(2000, seed=5)
    P = ND.sparsity.jac_prototype(grid.network())
    _, ncolor = ND.sparsity.color_columns(P)
    assert ncolor <= 60, ncolor


@pytest.mark.gpu
def test_colored_jacobian_equals_dense(ND, dev):
    import torch
    for name, nw, u, p in _cases(ND):
        ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
        Js = ND.sparsity.sparse_jacobian(nw, ud, pd, 0.0).toarray()
        Jd = nw.jacobian(ud, pd, 0.0, colored=False).cpu().numpy()
        np.testing.assert_allclose(Js, Jd, rtol=1e-15, atol=1e-15 * np.abs(Jd).max(), err_msg=name)


@pytest.mark.gpu
def test_find_fixpoint_sparse_lu_on_a_large_dae_grid(ND, dev):
    import torch
    grid = ND.swing_load_grid(4000, seed=53)
    n = grid.g.nv()
    th = torch.zeros(n, dtype=torch.float64, device=dev)
    fs = ND.FixpointSolver([grid.g])
    st = ND.FixpointSolver.stats_to_numpy(fs.solve(th, torch.from_numpy(grid.P).to(dev),
                                                   torch.from_numpy(grid.scale).to(dev),
                                                   torch.from_numpy(grid.K).to(dev)))
    assert st["retcode"][0] == 0
    nw = grid.network(algebraic_loads=True)
    assert nw.dim > 2048                                # the automatic choice is the sparse path
    x, info = ND.find_fixpoint(nw, np.zeros(nw.dim), grid.flat_p())
    assert info["residual"] <= 1e-10 and info["iters"] <= 12
    v, _ = nw.ranges()
    got = x.cpu().numpy()[v[:, 0, 0] - 1]
    ref = th.cpu().numpy()
    np.testing.assert_allclose(got - got.mean(), ref - ref.mean(), atol=1e-8)
