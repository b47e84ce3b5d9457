"""GPU parity of nw(du, u, p, t) through the C ABI against the CPU oracle.

Tolerance (FP64): |du_gpu - du_ref| <= 1e-12 * max(|du_ref|, S) per element,
S = sum of the magnitudes of the element's terms (oracle scale).  Segments of
at most hub_degree entries are summed in the reference order, so their only
difference from the oracle is the last ulp of sin; longer segments use a fixed
wave-wide tree (still far inside 1e-12 * S).
"""
import numpy as np
import pytest

from conftest import oracle_for, scaled_error

RTOL = 1e-12
pytestmark = pytest.mark.gpu


def _dev_rhs(nw, u, p, dev, t=0.0):
    import torch
    du = torch.full((nw.dim,), float("nan"), dtype=torch.float64, device=dev)
    nw(du, torch.from_numpy(np.ascontiguousarray(u)).to(dev),
       torch.from_numpy(np.ascontiguousarray(p if len(p) else np.zeros(1))).to(dev), t)
    torch.cuda.synchronize()
    return du.cpu().numpy()


def _check(nw, u, p, dev, rtol=RTOL):
    """Both ways of gathering neighbour outputs (straight from u, and through the compact
    per-call copy with slot-ordered parameters) against the oracle."""
    on = oracle_for(nw)
    ref = on.rhs(u, p)
    S = on.scale(u, p)
    mode = nw.tuning()["gather_mode"]
    for gm in (2, 3):
        nw.tune(gather_mode=gm)
        got = _dev_rhs(nw, u, p, dev)
        assert np.all(np.isfinite(got))
        err = scaled_error(got, ref, S)
        assert err.max() <= rtol, f"gather mode {gm}: max scaled error {err.max():.3e} at {err.argmax()}"
    nw.tune(gather_mode=mode)
    return got, ref


def test_kuramoto_er_c1(ND, dev):
    # first-order Kuramoto, ER N=1000 E=4000, K=3 (config 1)
    nw, p, u0 = ND.workloads.kuramoto_er(1000, 4000, 3.0, seed=1)
    _check(nw, u0, p, dev)


def test_swing_ba_20k(ND, dev):
    nw, p, u0 = ND.workloads.swing_ba(20_000, 2, seed=2)
    _check(nw, u0, p, dev)


def test_swing_ba_hub_thresholds(ND, dev):
    nw, p, u0 = ND.workloads.swing_ba(20_000, 2, seed=3)
    on = oracle_for(nw)
    ref, S = on.rhs(u0, p), on.scale(u0, p)
    for hub in (1, 8, 64, 1 << 30):
        nw.tune(hub_degree=hub)
        got = _dev_rhs(nw, u0, p, dev)
        assert scaled_error(got, ref, S).max() <= RTOL, hub
    # hub_degree beyond every segment: reference summation order everywhere; only sin ulps differ
    nw.tune(hub_degree=1 << 30)
    got = _dev_rhs(nw, u0, p, dev)
    assert np.array_equal(got[0::2], ref[0::2])


def test_heterogeneous_rgg_c3_small(ND, dev):
    nw, p, u0 = ND.workloads.heterogeneous_rgg(20_000, seed=4)
    assert nw.info.nvbatches == 2
    _check(nw, u0, p, dev)


def test_complete4_multibatch_with_ode_edges(ND, dev):
    from test_oracle import _complete4
    nw = _complete4(ND)
    rng = np.random.default_rng(5)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2, nw.pdim)
    _check(nw, u, p, dev)


@pytest.mark.parametrize("seed,directed", [(1, True), (2, False)])
def test_random_mixed_batches(ND, dev, seed, directed):
    from test_oracle import _random_mixed
    nw, u, p = _random_mixed(ND, n=3000, seed=seed, directed=directed)
    _check(nw, u, p, dev)


def _six_vertex_batches(ND, n, seed, edges):
    """Six vertex batches of the fused kernel set (swing twice, through two distinct
    device functions, unit-damping swing, first-order Kuramoto, diffusion, load): more
    batches than one grouped launch takes."""
    import copy
    L, c = ND.Lib, ND.components
    swing2 = c.VertexModel(f=copy.copy(c.kuramoto_inertia), sym=["delta", "omega"], psym=["M", "D", "Pm"],
                           g=c.StateMask(1), name="swing2")
    vt = [L.kuramoto_second(), swing2, L.bench.kuramoto_vertex_2d(), L.kuramoto_first(), L.diffusion_vertex(),
          L.load_vertex()]
    rng = np.random.default_rng(seed)
    g = ND.watts_strogatz(n, 4, 0.6, seed=seed)
    vm = [vt[i] for i in rng.integers(0, len(vt), n)]
    em = [edges[i] for i in rng.integers(0, len(edges), g.ne())] if len(edges) > 1 else edges[0]
    nw = ND.Network(g, vm, em)
    assert len(list(nw.batches(False))) == 6
    return nw, rng.normal(size=nw.dim), rng.uniform(0.5, 2.0, nw.pdim)


def test_many_vertex_batches_grouped_launches(ND, dev):
    L = ND.Lib
    for edges in ([L.static_line_edge()], [L.kuramoto_edge(), L.diffusion_edge(), L.diffusion_odeedge()]):
        nw, u, p = _six_vertex_batches(ND, 4000, 7, edges)
        _check(nw, u, p, dev)
        # fused RK stages go through the grouped launches as well (several batches: the
        # four-launches-per-step path)
        ref = oracle_for(nw).rk4(u, p, 0.0, 0.001, 50)
        np.testing.assert_allclose(_dev_rk4(nw, u, p, dev, 0.001, 50), ref, rtol=0, atol=1e-11)


def test_diffusion_closed_form_gpu(ND, dev):
    import scipy.sparse as sp
    g = ND.barabasi_albert(2000, 5, seed=2)
    nw = ND.Network(g, ND.Lib.diffusion_test.odevertex(), ND.Lib.diffusion_test.staticedge())
    n = g.nv()
    A = sp.coo_matrix((np.ones(g.ne()), (g.src - 1, g.dst - 1)), shape=(n, n))
    A = (A + A.T).tocsr()
    Lap = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
    x = np.random.default_rng(0).normal(size=n)
    got = _dev_rhs(nw, x, np.zeros(0), dev)
    np.testing.assert_allclose(got, -(Lap @ x), rtol=0, atol=1e-11)


def test_ode_edge_networks(ND, dev):
    g = ND.watts_strogatz(500, 6, 0.3, seed=1)
    for em in (ND.Lib.bench.diffusion_dedge(), ND.Lib.diffusion_test.odeedge(), ND.Lib.diffusion_odeedge()):
        nw = ND.Network(g, ND.Lib.bench.diffusion_vertex(), em)
        rng = np.random.default_rng(1)
        _check(nw, rng.normal(size=nw.dim), rng.uniform(1, 2, nw.pdim), dev)


def test_benchmark_cases_models(ND, dev):
    # benchmark/benchmark_cases.jl: kuramoto homogeneous / heterogeneous, diffusion static edge
    g = ND.watts_strogatz(3000, 4, 0.8, seed=1)
    rng = np.random.default_rng(2)
    B = ND.Lib.bench
    v1, v2 = B.kuramoto_vertex_1d(), B.kuramoto_vertex_2d()
    vm = [v1 if x else v2 for x in rng.random(3000) < 0.5]
    for nw in (ND.Network(g, B.kuramoto_vertex_2d(), B.static_kuramoto_edge()),
               ND.Network(g, vm, B.static_kuramoto_edge()),
               ND.Network(g, B.diffusion_vertex(), B.diffusion_edge())):
        _check(nw, rng.normal(size=nw.dim), rng.normal(size=nw.pdim), dev)


def test_deterministic_and_inputs_untouched(ND, dev):
    import torch
    nw, p, u0 = ND.workloads.swing_ba(50_000, 2, seed=6)
    nw.tune(hub_degree=8)
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    a = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    nw(a, u, pd, 0.0)
    nw(b, u, pd, 0.0)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert np.array_equal(u.cpu().numpy(), u0) and np.array_equal(pd.cpu().numpy(), p)   # testutils.jl:43-44


def test_size_and_type_errors(ND, dev):
    import torch
    nw, p, u0 = ND.workloads.swing_ba(1000, 2, seed=7)
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    with pytest.raises(ND.ArgumentError):                      # coreloop.jl:2-4
        nw(du[:-1], u, pd, 0.0)
    with pytest.raises(ND.ArgumentError):                      # coreloop.jl:5-7
        nw(du, u, pd[:-1], 0.0)
    with pytest.raises(ND.ArgumentError):                      # network_structure.jl:146-151
        nw(du.float(), u.float(), pd.float(), 0.0)
    # host arrays: the HIP path takes device vectors only (no CPU fallback)
    with pytest.raises(ND.ArgumentError, match="HIP device"):
        nw(du.cpu(), u.cpu(), pd.cpu(), 0.0)
    with pytest.raises(ND.ArgumentError, match="torch tensor"):
        nw(np.empty(nw.dim), u0, p, 0.0)
    # the same checks in the C ABI itself, for a binding that passes raw pointers
    import ctypes as C
    L = ND._abi.lib()
    args = (C.c_void_p(du.data_ptr()), C.c_void_p(u.data_ptr()), C.c_void_p(pd.data_ptr()), 0.0)
    assert L.nd_rhs(nw.plan, *args, nw.dim - 1, nw.dim, nw.pdim, None) == ND._abi.ND_ERR_ARGUMENT
    assert "du" in L.nd_last_error().decode()
    assert L.nd_rhs(nw.plan, *args, nw.dim, nw.dim, nw.pdim + 1, None) == ND._abi.ND_ERR_ARGUMENT
    assert L.nd_rhs(nw.plan, None, args[1], args[2], 0.0, nw.dim, nw.dim, nw.pdim, None) == ND._abi.ND_ERR_ARGUMENT
    assert L.nd_rhs_jvp(nw.plan, args[0], None, args[1], None, args[2], 0.0, nw.dim, nw.dim, nw.pdim,
                        None) == ND._abi.ND_ERR_ARGUMENT


def test_get_buffers_and_primitives(ND, dev):
    import torch
    from test_oracle import _random_mixed
    nw, u, p = _random_mixed(ND, n=2000, seed=8)
    on = oracle_for(nw)
    ud, pd = torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev)
    o, agg = nw.get_buffers(ud, pd, 0.0)
    oref, aref = on.buffers(u, p)
    o, agg = o.cpu().numpy(), agg.cpu().numpy()
    assert np.array_equal(np.isnan(o), np.isnan(oref))
    m = ~np.isnan(oref)
    np.testing.assert_allclose(o[m], oref[m], rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(agg, aref, rtol=1e-12, atol=1e-12)
    # gather! (gbufs.jl:25) bit-exact
    ob = torch.from_numpy(np.nan_to_num(oref)).to(dev)
    gb = torch.empty(nw.info.lastidx_gbuf, dtype=torch.float64, device=dev)
    nw.gather(gb, ob)
    torch.cuda.synchronize()
    assert np.array_equal(gb.cpu().numpy(), on.gather(np.nan_to_num(oref)))


def test_aggregate_primitive_depth2(ND, dev):
    # aggregate! on the mixed depth-2 network of test/aggregators_test.jl: additive, and
    # bit-exact with SequentialAggregator once every segment is summed in order
    import torch
    rng = np.random.default_rng(11)
    n = 5000
    g0 = ND.watts_strogatz(n, 4, 0.8, seed=1)
    flip = rng.random(g0.ne()) < 0.5
    g = ND.SimpleDiGraph(n, np.where(flip, g0.dst, g0.src), np.where(flip, g0.src, g0.dst))
    C = ND.components
    vt = [ND.VertexModel(dim=d, g=[1, 2]) for d in (2, 3, 4)]
    et = [ND.EdgeModel(g=W(C.noop_edge), outdim=2) for W in (ND.Symmetric, ND.AntiSymmetric, ND.Directed)]
    et.append(ND.EdgeModel(g=C.noop_edge2, outdim=2))
    nw = ND.Network(g, [vt[i] for i in rng.integers(0, 3, n)], [et[i] for i in rng.integers(0, 4, g.ne())])
    on = oracle_for(nw)
    states = rng.random(on.lastidx_out)
    ref = on.aggregate(np.zeros(on.lastidx_aggr), states, "seq")
    sd = torch.from_numpy(states).to(dev)
    nw.tune(hub_degree=1 << 30)
    agg = torch.zeros(on.lastidx_aggr, dtype=torch.float64, device=dev)
    nw.aggregate(agg, sd)
    torch.cuda.synchronize()
    assert np.array_equal(agg.cpu().numpy(), ref)
    off = torch.arange(1, on.lastidx_aggr + 1, dtype=torch.float64, device=dev)
    nw.tune(hub_degree=2)
    nw.aggregate(off, sd)
    torch.cuda.synchronize()
    # additivity (aggregators_test.jl:69-79, there with isapprox): exact up to the ulp of the offset
    np.testing.assert_allclose(off.cpu().numpy() - ref, np.arange(1, on.lastidx_aggr + 1), rtol=4e-16, atol=1e-12)


def test_swing_c2_full_size_properties(ND, dev):
    # config 2 at full size (N = 1e6): size-independent checks on the GPU result
    import torch
    nw, p, u0 = ND.workloads.swing_ba(1_000_000, 2, seed=1)
    n = nw.nv()
    got = _dev_rhs(nw, u0, p, dev)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got[0::2], u0[1::2])               # dtheta = omega exactly
    M, D, P = p[0:3 * n:3], p[1:3 * n:3], p[2:3 * n:3]
    w = u0[1::2]
    # sum_i M_i dw_i = sum_i P_i - D_i w_i: the coupling cancels pairwise (AntiSymmetric)
    lhs = np.sum(M * got[1::2])
    rhs = np.sum(P - D * w)
    scale = np.sum(np.abs(M * got[1::2])) + np.sum(np.abs(P)) + np.sum(np.abs(D * w))
    assert abs(lhs - rhs) <= 1e-12 * scale
    # sampled vertices against the oracle restated on their neighbourhoods: full-size oracle run
    on = oracle_for(nw)
    ref, S = on.rhs(u0, p), on.scale(u0, p)
    assert scaled_error(got, ref, S).max() <= RTOL


def _dev_rk4(nw, u0, p, dev, dt, nsteps):
    """nd_rk4 with the neighbour outputs read from u and through the compact copy; both
    must agree (the results are the same arithmetic, gathered from different buffers)."""
    import torch
    mode = nw.tuning()["gather_mode"]
    out = []
    for gm in (2, 3):
        nw.tune(gather_mode=gm)
        u = torch.from_numpy(np.array(u0)).to(dev)
        nw.rk4(u, torch.from_numpy(np.ascontiguousarray(p if len(p) else np.zeros(1))).to(dev), 0.0, dt, nsteps)
        torch.cuda.synchronize()
        out.append(u.cpu().numpy())
    nw.tune(gather_mode=mode)
    assert np.array_equal(out[0], out[1])
    return out[1]


def test_rk4_kuramoto_c1(ND, dev):
    # config 1: RK4 dt = 0.01 to t = 10 against the oracle's textbook RK4.  The dynamics
    # contract (a 1e-15 relative perturbation of u0 stays below 1e-15 at t = 10), so the
    # only difference, sin's last ulp per stage, stays at round-off level.
    nw, p, u0 = ND.workloads.kuramoto_er(1000, 4000, 3.0, seed=1)
    ref = oracle_for(nw).rk4(u0, p, 0.0, 0.01, 1000)
    got = _dev_rk4(nw, u0, p, dev, 0.01, 1000)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_rk4_swing_and_ode_edges(ND, dev):
    nw, p, u0 = ND.workloads.swing_ba(20_000, 2, seed=2)
    ref = oracle_for(nw).rk4(u0, p, 0.0, 0.001, 200)
    np.testing.assert_allclose(_dev_rk4(nw, u0, p, dev, 0.001, 200), ref, rtol=0, atol=1e-10)
    from test_oracle import _complete4
    nw = _complete4(ND)
    rng = np.random.default_rng(5)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2, nw.pdim)
    ref = oracle_for(nw).rk4(u, p, 0.0, 0.002, 100)
    np.testing.assert_allclose(_dev_rk4(nw, u, p, dev, 0.002, 100), ref, rtol=0, atol=1e-12)


def test_edge_parameter_changes_between_calls(ND, dev):
    # the slot-ordered copy of the edge parameters must follow every change of p: a few
    # edges, all edges, alternating parameter vectors, -0.0 vs 0.0, and a hub re-tune
    import torch
    nw, p, u0 = ND.workloads.swing_ba(20_000, 2, seed=8)
    nw.tune(gather_mode=3)
    assert nw.tuning()["gather_mode"] == 3
    on = oracle_for(nw)
    n, E = nw.nv(), nw.ne()
    rng = np.random.default_rng(8)
    u = torch.from_numpy(u0).to(dev)
    du = torch.empty_like(u)

    def check(ph):
        nw(du, u, torch.from_numpy(ph).to(dev), 0.0)
        torch.cuda.synchronize()
        err = scaled_error(du.cpu().numpy(), on.rhs(u0, ph), on.scale(u0, ph))
        assert err.max() <= RTOL

    p1 = p.copy()
    check(p1)
    p2 = p1.copy()
    idx = rng.choice(E, 50, replace=False)
    p2[3 * n + idx] = rng.uniform(5, 10, 50)
    check(p2)
    p3 = p1.copy()
    p3[3 * n:] = rng.uniform(5, 10, E)
    for ph in (p3, p1, p3, p2):
        check(ph)
    p4 = p1.copy()
    p4[3 * n + idx[:5]] = 0.0
    check(p4)
    p4[3 * n + idx[:5]] = -0.0
    check(p4)
    nw.tune(hub_degree=8)
    check(p2)
    nw.tune(hub_degree=48)
    check(p1)
    # same tensor modified in place between calls
    pd = torch.from_numpy(p1.copy()).to(dev)
    nw(du, u, pd, 0.0)
    pd[3 * n + int(idx[0])] = 123.0
    nw(du, u, pd, 0.0)
    torch.cuda.synchronize()
    p5 = p1.copy()
    p5[3 * n + int(idx[0])] = 123.0
    assert scaled_error(du.cpu().numpy(), on.rhs(u0, p5), on.scale(u0, p5)).max() <= RTOL
    # RK4 after a parameter change
    ref = on.rk4(u0, p3, 0.0, 0.001, 20)
    np.testing.assert_allclose(_dev_rk4(nw, u0, p3, dev, 0.001, 20), ref, rtol=0, atol=1e-10)


def test_static_line_parameter_cache(ND, dev):
    # two parameters per edge (admittance, phase shift), several vertex batches
    nw, p, u0 = ND.workloads.heterogeneous_rgg(20_000, seed=9)
    assert nw.info.nebatches == 1
    nw.tune(gather_mode=3)
    _check(nw, u0, p, dev)
    lo, hi = int(nw.ranges()[1][0, 3, 0]) - 1, nw.pdim     # one static-line batch: params contiguous
    p2 = p.copy()
    p2[lo + 1:hi:2] = 0.05                      # phase shifts only
    _check(nw, u0, p2, dev)
    _check(nw, u0, p, dev)


def test_edge_cases_no_edges_isolated_star(ND, dev):
    # no edges at all; isolated vertices next to a component; a star whose centre is a
    # big hub (block-per-vertex path) with cached parameters
    import torch
    rng = np.random.default_rng(12)
    vm, em = ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge()
    for g in (ND.SimpleGraph(5), ND.SimpleGraph(1),
              ND.SimpleGraph(40, [1, 2, 3, 10], [2, 3, 4, 11]),
              ND.SimpleGraph(3001, np.ones(3000, np.int64), np.arange(2, 3002))):
        nw = ND.Network(g, vm, em)
        p = np.r_[np.stack([rng.uniform(1, 5, g.nv()), rng.uniform(0.1, 1, g.nv()), rng.normal(size=g.nv())],
                           1).ravel(), rng.uniform(5, 10, g.ne())]
        u = rng.normal(size=nw.dim)
        _check(nw, u, p, dev)


def test_heterogeneous_rgg_c3_full_size(ND, dev):
    # the heterogeneous network at its benchmark size (N = 2e5, two vertex batches)
    nw, p, u0 = ND.workloads.heterogeneous_rgg(200_000, seed=1)
    _check(nw, u0, p, dev)


def _dev_jvp(nw, u, v, p, dev):
    import torch
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    mode = nw.tuning()["gather_mode"]
    res = []
    for gm in (2, 3):
        nw.tune(gather_mode=gm)
        du = torch.full((nw.dim,), float("nan"), dtype=torch.float64, device=dev)
        jv = torch.full((nw.dim,), float("nan"), dtype=torch.float64, device=dev)
        nw.jvp(du, jv, d(u), d(v), d(p if len(p) else np.zeros(1)), 0.0)
        torch.cuda.synchronize()
        res.append((du.cpu().numpy(), jv.cpu().numpy()))
    nw.tune(gather_mode=mode)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    return res[1]


def test_jvp_closed_forms(ND, dev):
    # dual-number evaluation against the closed-form Jacobians (Kuramoto incidence form, swing)
    from test_oracle import kuramoto_jvp_closed_form, swing_jvp_closed_form
    rng = np.random.default_rng(21)
    nw, p, u0 = ND.workloads.kuramoto_er(1000, 4000, 3.0, seed=1)
    v = rng.normal(size=nw.dim)
    du, jv = _dev_jvp(nw, u0, v, p, dev)
    ref, S = kuramoto_jvp_closed_form(nw.g, u0, p[nw.nv():], v)
    assert scaled_error(jv, ref, S).max() <= RTOL
    on = oracle_for(nw)
    assert scaled_error(du, on.rhs(u0, p), on.scale(u0, p)).max() <= RTOL
    nw, p, u0 = ND.workloads.swing_ba(20_000, 2, seed=22)
    n = nw.nv()
    v = rng.normal(size=nw.dim)
    du, jv = _dev_jvp(nw, u0, v, p, dev)
    ref, S = swing_jvp_closed_form(nw.g, u0, p[0:3 * n:3], p[1:3 * n:3], p[3 * n:], v)
    assert scaled_error(jv, ref, S).max() <= RTOL
    # the value part is the RHS, every segment in the reference order
    on = oracle_for(nw)
    assert scaled_error(du, on.rhs(u0, p), on.scale(u0, p)).max() <= RTOL
    nw.tune(hub_degree=1 << 30)
    assert np.array_equal(du[0::2], _dev_rhs(nw, u0, p, dev)[0::2])


@pytest.mark.parametrize("seed,directed", [(31, False), (32, True)])
def test_jvp_mixed_networks_against_finite_differences(ND, dev, seed, directed):
    # several vertex and edge batches, every wrapper, ODE edges: against the Richardson
    # finite difference of the oracle (truncation and rounding ~1e-9 of the term scale)
    from conftest import fd_jvp
    from test_oracle import _random_mixed
    nw, u, p = _random_mixed(ND, n=400, directed=directed, seed=seed)
    rng = np.random.default_rng(seed)
    v = rng.normal(size=nw.dim)
    du, jv = _dev_jvp(nw, u, v, p, dev)
    on = oracle_for(nw)
    ref = fd_jvp(on, u, v, p)
    S = np.abs(fd_jvp(on, u, np.abs(v), p)) + np.abs(ref) + 1.0
    assert np.all(np.isfinite(jv))
    assert (np.abs(jv - ref) / S).max() <= 1e-7
    assert scaled_error(du, on.rhs(u, p), on.scale(u, p)).max() <= RTOL


def test_jvp_argument_errors(ND, dev):
    # the size checks of nw(du, u, p, t) (coreloop.jl:2-7) hold for the dual evaluation too
    import torch
    nw, p, u0 = ND.workloads.kuramoto_er(100, 300, 3.0, seed=2)
    d = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dev)
    u, pd = d(u0), d(p)
    good = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    with pytest.raises(ND.ArgumentError):
        nw.jvp(good, good.clone(), d(u0[:-1]), good.clone(), pd)
    with pytest.raises(ND.ArgumentError):
        nw.jvp(good, good.clone(), u, good.clone(), d(p[:-1]))
    with pytest.raises(ND.ArgumentError):
        nw.jvp(good, good.clone(), u, torch.empty(nw.dim + 1, dtype=torch.float64, device=dev), pd)


def test_dense_jacobian_kuramoto(ND, dev):
    # the dense Jacobian of first-order Kuramoto: symmetric, rows sum to zero (global phase
    # shift), negative semidefinite near synchrony, equal to the closed form
    import torch
    from test_oracle import _incidence
    nw, p, u0 = ND.workloads.kuramoto_er(200, 600, 3.0, seed=3)
    th = np.random.default_rng(3).normal(scale=0.2, size=nw.nv())
    J = nw.jacobian(torch.from_numpy(th).to(dev), torch.from_numpy(p).to(dev)).cpu().numpy()
    B = _incidence(nw.g).toarray()
    K = p[nw.nv():]
    ref = -(B * (K * np.cos(B.T @ th))) @ B.T
    np.testing.assert_allclose(J, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    np.testing.assert_allclose(J, J.T, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert np.abs(J.sum(axis=1)).max() <= 1e-12 * np.abs(ref).max()
    ev = np.linalg.eigvalsh(0.5 * (J + J.T))
    assert ev.max() <= 1e-10 * np.abs(ev).max()


@pytest.mark.parametrize("em,n", [("kuramoto", 1500), ("static_line", 1500), ("kuramoto", 1000)])
def test_rk4_single_workgroup_path(ND, dev, em, n):
    # small single-batch networks: nd_rk4 runs all steps in one 1024-thread workgroup (one
    # vertex per thread up to 1024 vertices, two above); same answer as the oracle and as
    # the four-launches-per-step path
    rng = np.random.default_rng(51)
    g = ND.barabasi_albert(n, 2, seed=51)
    n, m = g.nv(), g.ne()
    edge = ND.Lib.kuramoto_edge() if em == "kuramoto" else ND.Lib.static_line_edge()
    nw = ND.Network(g, ND.Lib.kuramoto_second(), edge)
    pe = rng.uniform(5, 10, m) if em == "kuramoto" else np.stack([rng.uniform(5, 10, m), rng.uniform(0, 0.1, m)], 1).ravel()
    P = rng.normal(size=n)
    p = np.r_[np.stack([rng.uniform(1, 5, n), rng.uniform(0.1, 1, n), P - P.mean()], 1).ravel(), pe]
    u0 = np.empty(2 * n)
    u0[0::2], u0[1::2] = rng.uniform(-np.pi, np.pi, n), rng.normal(size=n)
    ref = oracle_for(nw).rk4(u0, p, 0.0, 0.002, 250)
    assert nw.tuning()["rk4_path"] == 1
    one = _dev_rk4(nw, u0, p, dev, 0.002, 250)
    nw.tune(rk4_path=2)
    four = _dev_rk4(nw, u0, p, dev, 0.002, 250)
    np.testing.assert_allclose(one, ref, rtol=0, atol=1e-10)
    np.testing.assert_allclose(four, ref, rtol=0, atol=1e-10)


@pytest.mark.gpu
def test_device_sine_against_numpy(ND, dev):
    # The models' sine (a Cody-Waite reduction with Taylor kernels below |x| = 2^20 pi/2,
    # the library sine beyond): disjoint pairs of first-order Kuramoto vertices with
    # omega = 0 and K = 1 put exactly sin(theta_s - theta_d) into the dst side's du and
    # its negation into the src side's.  Against numpy's sine: within 2 ulp everywhere,
    # within 1 ulp for |x| < 40, and NaN for non-finite arguments.
    import torch
    rng = np.random.default_rng(81)
    near = np.arange(-3000, 3001) * (np.pi / 2)
    xs = np.concatenate([rng.uniform(-np.pi, np.pi, 20000), rng.uniform(-40, 40, 20000),
                         rng.uniform(-1e5, 1e5, 10000), rng.uniform(1.5e6, 1.8e6, 10000),
                         10.0 ** rng.uniform(7, 300, 4000), np.nextafter(near, np.inf), np.nextafter(near, -np.inf),
                         [0.0, 1e-300, 5e-324, 1647099.3291652855, np.nextafter(1647099.3291652855, 0), -1e308]])
    n = len(xs)
    g = ND.SimpleGraph(2 * n, np.arange(1, 2 * n, 2), np.arange(2, 2 * n + 1, 2))
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    u = np.zeros(2 * n)
    u[0::2] = xs
    p = np.r_[np.zeros(2 * n), np.ones(n)]
    du = torch.empty(2 * n, dtype=torch.float64, device=dev)
    nw(du, torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev), 0.0)
    got = du.cpu().numpy()
    ref = np.sin(xs)
    np.testing.assert_array_equal(got[0::2], -got[1::2])
    ulp = np.spacing(np.maximum(np.abs(ref), np.finfo(float).tiny))
    err = np.abs(got[1::2] - ref) / ulp
    assert err.max() <= 2.0, (err.max(), xs[np.argmax(err)])
    small = np.abs(xs) < 40
    assert err[small].max() <= 1.0
    inf = torch.tensor([np.inf, 0.0, -np.inf, 0.0, np.nan, 0.0], dtype=torch.float64, device=dev)
    nw3 = ND.Network(ND.SimpleGraph(6, [1, 3, 5], [2, 4, 6]), ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    d3 = torch.empty(6, dtype=torch.float64, device=dev)
    nw3(d3, inf, torch.from_numpy(np.r_[np.zeros(6), np.ones(3)]).to(dev), 0.0)
    assert torch.isnan(d3).all()


@pytest.mark.gpu
def test_rk4_dense_network_beyond_one_workgroup(ND, dev):
    # a complete graph K_120 (7140 edges) fits the one-workgroup kernel by its model kinds
    # and size, but its topology needs more LDS than a workgroup has: nd_rk4 steps it
    # through the stage launches instead of refusing, and Network.tsit5 steps it on the host
    rng = np.random.default_rng(57)
    g = ND.complete_graph(120)
    n, m = g.nv(), g.ne()
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    p = np.r_[rng.normal(size=n), rng.uniform(0.5, 1.0, m) / n]
    u0 = rng.uniform(-np.pi, np.pi, n)
    ref = oracle_for(nw).rk4(u0, p, 0.0, 0.01, 200)
    assert nw.tuning()["rk4_path"] == 1
    got = _dev_rk4(nw, u0, p, dev, 0.01, 200)
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-10)
    import torch
    u = torch.from_numpy(u0).to(dev)
    info = nw.tsit5(u, torch.from_numpy(p).to(dev), (0.0, 1.0))
    assert info["path"] == "host steps" and info["retcode"] == 0


def test_plain_pointer_round_trip_through_the_c_abi(ND, dev):
    # what a host shim without a GPU array package does: nd_malloc, nd_memcpy_h2d,
    # nd_rhs on the default stream, nd_device_synchronize, nd_memcpy_d2h, nd_free
    import ctypes as C
    L = ND._abi.lib()
    chk = ND._abi.check
    nw, p, u0 = ND.workloads.swing_ba(5000, 2, seed=61)
    ptrs = {}
    for name, arr in (("u", u0), ("p", p), ("du", np.zeros(nw.dim))):
        h = C.c_void_p()
        chk(L.nd_malloc(arr.nbytes, C.byref(h)))
        ptrs[name] = h
        chk(L.nd_memcpy_h2d(h, np.ascontiguousarray(arr).ctypes.data, arr.nbytes))
    try:
        chk(L.nd_rhs(nw.plan, ptrs["du"], ptrs["u"], ptrs["p"], 0.0, nw.dim, nw.dim, nw.pdim, None))
        chk(L.nd_device_synchronize())
        got = np.empty(nw.dim)
        chk(L.nd_memcpy_d2h(got.ctypes.data, ptrs["du"], got.nbytes))
    finally:
        for h in ptrs.values():
            chk(L.nd_free(h))
    on = oracle_for(nw)
    assert scaled_error(got, on.rhs(u0, p), on.scale(u0, p)).max() <= RTOL
    with pytest.raises(ND.ArgumentError):
        chk(L.nd_memcpy_h2d(None, u0.ctypes.data, 8))
