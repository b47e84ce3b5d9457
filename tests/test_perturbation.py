"""nw(du, u, p, t; perturb, perturb_maps) (src/coreloop.jl:1,49-87) and the input/output
linearization built on it (linearize_network(s0; in, out), src/linear_analysis.jl:86-122,
_classify_perturbation_channels :629-680).

The oracle's four perturbation points are pinned by exact equivalences and closed forms:
a vertex-output perturbation of first-order Kuramoto vertices is a state shift (their f
does not read the state), a vertex-input perturbation adds to a diffusion vertex's
derivative, an edge-input perturbation shifts one edge's view of its source, and an
edge-output perturbation adds to one side's input.  The device's perturbed loop and its
dual numbers are then checked against the oracle and its finite differences.
"""
import numpy as np
import pytest

from conftest import oracle_for, scaled_error


def _kuramoto_pair(ND, K=2.0):
    c = ND.components
    vm = ND.VertexModel(f=c.kuramoto_vertex, sym=["theta"], psym=["omega"], g=1, insym=["P_in"])
    em = ND.EdgeModel(g=ND.AntiSymmetric(c.kuramoto_edge_g), outsym=["P"], psym=["K"], insym=["theta"])
    nw = ND.Network(ND.SimpleGraph(2, [1], [2]), vm, em)
    return nw, np.array([0.3, -0.5]), np.array([1.0, -0.7, K])


def test_vertex_output_perturbation_is_a_state_shift(ND):
    g = ND.erdos_renyi(60, 200, seed=3)
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    rng = np.random.default_rng(3)
    u, p = rng.uniform(-np.pi, np.pi, nw.dim), rng.uniform(0.5, 2.0, nw.pdim)
    on = oracle_for(nw)
    v, _ = nw.ranges()
    k, d = 7, 0.25
    du, o, _ = on.rhs_perturbed(u, p, [d], dict(vo_map=[(int(v[k, 1, 0]), 1)]))
    us = u.copy()
    us[v[k, 0, 0] - 1] += d
    np.testing.assert_array_equal(du, on.rhs(us, p))
    assert o[v[k, 1, 0] - 1] == us[v[k, 0, 0] - 1]


def test_vertex_input_and_edge_perturbations_closed_forms(ND):
    nw, u, p = _kuramoto_pair(ND)
    on = oracle_for(nw)
    om1, om2, K = p
    y = K * np.sin(u[0] - u[1])                    # dst output of the one edge (1, 2)
    base = on.rhs(u, p)
    np.testing.assert_allclose(base, [om1 - y, om2 + y], rtol=1e-15, atol=1e-15)
    # vertex input of vertex 2 (aggregation index 2): its acc gains d
    du, _, agg = on.rhs_perturbed(u, p, [0.1], dict(vi_map=[(2, 1)]))
    np.testing.assert_allclose(du, [om1 - y, om2 + (0.1 + y)], rtol=1e-15, atol=1e-15)
    assert agg[1] == 0.1 + y
    # edge input, src side (gather index 1): the edge sees theta_1 + d
    ys = K * np.sin(u[0] + 0.2 - u[1])
    du, _, _ = on.rhs_perturbed(u, p, [0.2], dict(ei_map=[(1, 1)]))
    np.testing.assert_allclose(du, [om1 - ys, om2 + ys], rtol=1e-15, atol=1e-15)
    # edge output, dst side (output index 4: vertex outputs 1, 2, then the edge's src 3, dst 4)
    du, o, _ = on.rhs_perturbed(u, p, [0.3], dict(eo_map=[(4, 1)]))
    np.testing.assert_allclose(du, [om1 - y, om2 + (y + 0.3)], rtol=1e-15, atol=1e-15)
    # and the channels of one map add in map order
    du, _, _ = on.rhs_perturbed(u, p, [0.3, 0.4], dict(eo_map=[(3, 1), (3, 2)]))
    np.testing.assert_allclose(du, [om1 + (-y + 0.3 + 0.4), om2 + y], rtol=1e-15, atol=1e-15)


def test_perturbation_channel_classification(ND):
    # linear_analysis_test.jl:178-189: one representative of every class
    nw, _, _ = _kuramoto_pair(ND)
    V, E = ND.VIndex, ND.EIndex
    chans = [V(1, "theta"), V(2, "P_in"), E(1, "₋P"), E(1, "P"), E(1, "src₊theta"), E(1, "dst₊theta"),
             V(2, "omega"), E(1, "K")]
    d0, maps = ND.linear_analysis.perturbation_channels(nw, chans)
    assert np.array_equal(d0, np.zeros(8))
    assert maps["vo_map"] == [(1, 1)] and maps["vi_map"] == [(2, 2)]
    assert maps["eo_map"] == [(3, 3), (4, 4)] and maps["ei_map"] == [(1, 5), (2, 6)]
    assert maps["p_map"] == [(2, 7), (3, 8)]
    with pytest.raises(ND.ArgumentError, match="not an input or output"):
        ND.linear_analysis.perturbation_channels(nw, [V(1, "nope")])


def _random_maps(nw, nch, rng):
    nv = nw.nv()
    lo, la, lg = int(nw.info.lastidx_out), int(nw.info.lastidx_aggr), int(nw.info.lastidx_gbuf)
    maps = dict(vo_map=[], vi_map=[], eo_map=[], ei_map=[])
    for k in range(1, nch + 1):
        kind = ("vo_map", "vi_map", "eo_map", "ei_map")[k % 4]
        b = {"vo_map": lambda: rng.integers(1, nv + 1), "vi_map": lambda: rng.integers(1, la + 1),
             "eo_map": lambda: rng.integers(nv + 1, lo + 1), "ei_map": lambda: rng.integers(1, lg + 1)}[kind]()
        maps[kind].append((int(b), k))
    return maps


def _fd(fun, h=1e-3):
    return (8.0 * (fun(h) - fun(-h)) - (fun(2 * h) - fun(-2 * h))) / (12.0 * h)


@pytest.mark.gpu
def test_perturbed_loop_matches_oracle(ND, dev):
    import torch
    from test_oracle import _random_mixed
    rng = np.random.default_rng(21)
    nw, u, p = _random_mixed(ND, n=400, seed=21)
    on = oracle_for(nw)
    maps = _random_maps(nw, 24, rng)
    delta = rng.normal(size=24)
    ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
    du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    # without channels the buffered loop is the plain RHS
    nw(du, ud, pd, 0.0, perturb=np.zeros(0), perturb_maps={})
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p), on.scale(u, p))
    assert err.max() <= 1e-12, err.max()
    du_ref, o_ref, agg_ref = on.rhs_perturbed(u, p, delta, maps)
    nw(du, ud, pd, 0.0, perturb=delta, perturb_maps=maps)
    S = on.scale(u, p) + np.abs(du_ref)
    err = scaled_error(du.cpu().numpy(), du_ref, S)
    assert err.max() <= 1e-12, err.max()
    jv = torch.empty_like(du)
    o, do, agg, dagg = nw.perturbed_jvp(du, jv, ud, None, pd, 0.0, delta, None, maps, buffers=True)
    np.testing.assert_allclose(o.cpu().numpy(), o_ref, rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(agg.cpu().numpy(), agg_ref, rtol=1e-13, atol=1e-13)
    assert np.all(jv.cpu().numpy() == 0.0)


@pytest.mark.gpu
def test_perturbed_duals_match_finite_differences(ND, dev):
    import torch
    from test_oracle import _random_mixed
    rng = np.random.default_rng(22)
    nw, u, p = _random_mixed(ND, n=200, seed=22)
    on = oracle_for(nw)
    maps = _random_maps(nw, 12, rng)
    delta = rng.normal(size=12)
    v, dd, dp = rng.normal(size=nw.dim), rng.normal(size=12), rng.normal(size=nw.pdim)
    ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
    du, jv = (torch.empty(nw.dim, dtype=torch.float64, device=dev) for _ in range(2))
    o, do, agg, dagg = nw.perturbed_jvp(du, jv, ud, torch.from_numpy(v).to(dev), pd, 0.0, delta, dd, maps,
                                        buffers=True, dp=torch.from_numpy(dp).to(dev))

    def along(h, what):
        r = on.rhs_perturbed(u + h * v, p + h * dp, delta + h * dd, maps)
        return r[what]
    np.testing.assert_allclose(jv.cpu().numpy(), _fd(lambda h: along(h, 0)), rtol=1e-7, atol=1e-7)
    do_fd = _fd(lambda h: np.nan_to_num(along(h, 1)))
    np.testing.assert_allclose(np.nan_to_num(do.cpu().numpy()), do_fd, rtol=1e-7, atol=1e-7)
    np.testing.assert_allclose(dagg.cpu().numpy(), _fd(lambda h: along(h, 2)), rtol=1e-7, atol=1e-7)


@pytest.mark.gpu
def test_linearize_network_in_out(ND, dev):
    c = ND.components
    V, E = ND.VIndex, ND.EIndex
    vm = ND.VertexModel(f=c.kuramoto_inertia, sym=["delta", "omega"], psym=["M", "D", "Pm"], g=c.StateMask(1),
                        insym=["P_in"])
    em = ND.EdgeModel(g=ND.AntiSymmetric(c.kuramoto_edge_g), outsym=["P"], psym=["K"], insym=["theta"])
    g = ND.watts_strogatz(12, 4, 0.3, seed=2)
    nw = ND.Network(g, vm, em)
    rng = np.random.default_rng(2)
    u = rng.normal(size=nw.dim)
    p = np.r_[np.stack([rng.uniform(1, 3, 12), rng.uniform(0.1, 1, 12), rng.normal(size=12)], 1).ravel(),
              rng.uniform(1, 3, g.ne())]
    on = oracle_for(nw)
    ins = [V(1, "delta"), V(2, "P_in"), E(1, "₋P"), E(2, "P"), E(1, "src₊theta"), E(3, "dst₊theta"),
           V(4, "M"), E(2, "K")]
    outs = [V(3, "omega"), V(1, "delta"), V(2, "P_in"), E(1, "P"), E(2, "src₊theta")]
    sys = ND.linear_analysis.linearize_network(nw, u, p, 0.0, in_=ins, out=outs)
    assert sys.B.shape == (nw.dim, 8) and sys.C.shape == (5, nw.dim) and sys.D.shape == (5, 8)
    _, maps = ND.linear_analysis.perturbation_channels(nw, ins)
    pmap = {k: b for b, k in maps["p_map"]}
    obs = [ND.linear_analysis._observable(nw, r) for r in outs]

    def run(x, pp, d):
        du, o, agg = on.rhs_perturbed(x, pp, d, maps)
        y = []
        for ob in obs:
            if ob[0] == "u":
                y.append(x[ob[1]])
            elif ob[0] == "o":
                y.append(o[ob[1]])
            elif ob[0] == "agg":
                y.append(agg[ob[1]])
            else:
                y.append(o[ob[2]] + sum(d[k - 1] for gi, k in maps["ei_map"] if gi == ob[1]))
        return du, np.array(y)

    for c_ in range(8):
        e = np.zeros(8)
        dp = np.zeros(nw.pdim)
        if c_ + 1 in pmap:
            dp[pmap[c_ + 1] - 1] = 1.0
        else:
            e[c_] = 1.0
        B_fd = _fd(lambda h: run(u, p + h * dp, h * e)[0])
        D_fd = _fd(lambda h: run(u, p + h * dp, h * e)[1])
        np.testing.assert_allclose(sys.B[:, c_], B_fd, rtol=1e-7, atol=1e-7, err_msg=str(ins[c_]))
        np.testing.assert_allclose(sys.D[:, c_], D_fd, rtol=1e-7, atol=1e-7, err_msg=str(ins[c_]))
    for j in range(nw.dim):
        ej = np.zeros(nw.dim)
        ej[j] = 1.0
        C_fd = _fd(lambda h: run(u + h * ej, p, np.zeros(8))[1])
        np.testing.assert_allclose(sys.C[:, j], C_fd, rtol=1e-7, atol=1e-7)
    A_ref = ND.linear_analysis.linearize_network(nw, u, p).A
    np.testing.assert_array_equal(sys.A, A_ref)
    # SISO: a scalar transfer function equal to the MIMO entry (linear_analysis_test.jl:156-170)
    siso = ND.linear_analysis.linearize_network(nw, u, p, 0.0, in_=V(2, "P_in"), out=V(3, "omega"))
    mimo = ND.linear_analysis.linearize_network(nw, u, p, 0.0, in_=[V(2, "P_in")], out=[V(3, "omega")])
    s = 10j * 2 * np.pi
    assert np.isscalar(siso(s)) and mimo(s).shape == (1, 1)
    assert np.isclose(siso(s), mimo(s)[0, 0], rtol=1e-14)
    with pytest.raises(ND.ArgumentError, match="Must specify `out`"):
        ND.linear_analysis.linearize_network(nw, u, p, 0.0, in_=[V(1, "delta")])
    with pytest.raises(ND.ArgumentError, match="Must specify `in`"):
        ND.linear_analysis.linearize_network(nw, u, p, 0.0, out=[V(1, "delta")])


@pytest.mark.gpu
def test_linearize_dae_transfer_function_survives_reduce_dae(ND, dev):
    # linear_analysis_test.jl:195-205: G_full(s) == G_reduced(s) for a DAE network
    V = ND.VIndex
    grid = ND.swing_load_grid(40, seed=11)
    nw = grid.network(algebraic_loads=True)
    x, _ = ND.find_fixpoint(nw, np.zeros(nw.dim), grid.flat_p())
    u = x.cpu().numpy()
    syms = nw.variable_symbols()
    gen = next(i for i, m in enumerate(nw.vertexm) if m.dim == 2) + 1
    load = next(i for i, m in enumerate(nw.vertexm) if m.dim == 1) + 1
    ins = [V(gen, nw.vertexm[gen - 1].psym[-1]), V(load, nw.vertexm[load - 1].psym[-1])]
    outs = [V(gen, nw.vertexm[gen - 1].sym[1]), V(load, nw.vertexm[load - 1].sym[0])]
    full = ND.linear_analysis.linearize_network(nw, u, grid.flat_p(), 0.0, in_=ins, out=outs)
    red = ND.linear_analysis.reduce_dae(full)
    assert full.constraint_dim() > 0 and red.dim < full.dim and len(syms) == full.dim
    for s in np.array([0.1j, 1j, 10j, 100j, -1 + 1j, -0.5 + 5j, -1 - 3j, 3j, 7j, 50j]) * 2 * np.pi:
        np.testing.assert_allclose(full(s), red(s), rtol=1e-8, atol=1e-12)
