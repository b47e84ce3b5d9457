"""Pins the CPU oracle (oracle/nd_oracle.c) before it is trusted as the checker.

Each check is chosen so a plausible restatement mistake (wrong registration
order, swapped src/dst, dropped sign, wrong summation order) fails it:
known answers from the reference's own tests, the hand-derived layout of the
reference's complete_graph(4) network, cross-style / cross-aggregator
agreement, closed forms, dense brute force and conservation invariants.
"""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import oracle_for, scaled_error

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_batchstride_known_answers(oracle):
    # test/utils_test.jl:147-178
    assert oracle.batchstride_fullstride([3]) == 3
    assert oracle.batchstride_fullstride([3, 1]) == 4
    assert oracle.batchstride_range(2, [3], 1) == (2, 4)
    assert oracle.batchstride_range(2, [3], 2) == (5, 7)
    assert oracle.batchstride_range(2, [3, 1], 1, 1) == (2, 4)
    assert oracle.batchstride_range(2, [3, 1], 1, 2) == (5, 5)
    assert oracle.batchstride_range(2, [3, 1], 2, 1) == (6, 8)
    assert oracle.batchstride_range(2, [3, 1], 2, 2) == (9, 9)


def _complete4(ND):
    g = ND.complete_graph(4)
    L = ND.Lib
    vf = [L.kuramoto_second(), L.diffusion_vertex(), L.kuramoto_second(), L.diffusion_vertex()]
    ef = [L.diffusion_odeedge(), L.kuramoto_edge(), L.kuramoto_edge(), L.diffusion_edge_fid(),
          L.diffusion_odeedge(), L.diffusion_edge_fid()]
    return ND.Network(g, vf, ef)


def test_complete4_layout_matches_hand_derivation(ND):
    fx = json.load(open(os.path.join(GOLDEN, "complete4_layout.json")))
    on = oracle_for(_complete4(ND))
    assert (on.dim, on.pdim, on.lastidx_out, on.lastidx_aggr, on.lastidx_gbuf) == (
        fx["dim"], fx["pdim"], fx["lastidx_out"], fx["lastidx_aggr"], fx["lastidx_gbuf"])
    v, e = on.ranges()
    for j, name in enumerate(["v_data", "v_out", "v_para", "v_aggr"]):
        assert v[:, j, :].tolist() == fx[name], name
    for j, name in enumerate(["e_data", "e_out_src", "e_out_dst", "e_para", "e_gbuf_src", "e_gbuf_dst"]):
        assert e[:, j, :].tolist() == fx[name], name
    a, gm = on.maps()
    assert [on.agg_first, on.agg_last] == fx["aggregation_range"]
    assert a.tolist() == fx["aggregation_map"]
    assert gm.tolist() == fx["gather_map"]
    d, ptr, s = on.inverse_map()
    inv = [[int(d[w]), s[ptr[w]:ptr[w + 1]].tolist()] for w in range(len(d))]
    assert inv == fx["inverse_map"]
    assert [sorted(on.batch(False, b)["indices"].tolist()) for b in range(on.nvb)] == fx["vertex_batches"]
    assert [on.batch(True, b)["indices"].tolist() for b in range(on.neb)] == fx["edge_batches"]
    # variable_index(EIndex(1, :e_dst)) == 7  (sym = [:e_dst, :e_src])
    assert e[0, 0, 0] == fx["e_dst_state_of_edge1"]


def _random_mixed(ND, n=300, seed=3, directed=True):
    """Random mix of wrappers / batches like test/aggregators_test.jl:21-30."""
    rng = np.random.default_rng(seed)
    L = ND.Lib
    g = ND.watts_strogatz(n, 4, 0.8, seed=seed)
    if directed:
        flip = rng.random(g.ne()) < 0.5
        s = np.where(flip, g.dst, g.src)
        d = np.where(flip, g.src, g.dst)
        g = ND.SimpleDiGraph(n, s, d)
    vt = [L.kuramoto_second(), L.kuramoto_first(), L.diffusion_vertex(), L.load_vertex()]
    et = [L.kuramoto_edge(), L.diffusion_edge(), L.diffusion_edge_fid(), L.diffusion_odeedge(),
          L.static_line_edge(),
          ND.EdgeModel(g=ND.Symmetric(ND.components.kuramoto_edge_g), pdim=1),
          ND.EdgeModel(g=ND.Directed(ND.components.diffusionedge_weighted), pdim=1)]
    vm = [vt[i] for i in rng.integers(0, len(vt), n)]
    em = [et[i] for i in rng.integers(0, len(et), g.ne())]
    nw = ND.Network(g, vm, em)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2.0, nw.pdim)
    return nw, u, p


def test_eager_equals_lazy_bitwise(ND):
    # Eager vs Lazy GBufProvider agree (test/testutils.jl:10-64)
    nw, u, p = _random_mixed(ND)
    on = oracle_for(nw)
    a = on.rhs(u, p, 0.0, buffered=True)
    b = on.rhs(u, p, 0.0, buffered=False)
    assert np.array_equal(a, b)


def test_aggregators_agree_and_are_additive(ND):
    # test/aggregators_test.jl:58-79 on a mixed network with output depth 2
    rng = np.random.default_rng(11)
    n = 2000
    g0 = ND.watts_strogatz(n, 4, 0.8, seed=1)
    flip = rng.random(g0.ne()) < 0.5
    g = ND.SimpleDiGraph(n, np.where(flip, g0.dst, g0.src), np.where(flip, g0.src, g0.dst))
    C = ND.components
    vt = [ND.VertexModel(dim=d, g=[1, 2]) for d in (2, 3, 4)]
    et = [ND.EdgeModel(g=W(C.noop_edge), outdim=2) for W in (ND.Symmetric, ND.AntiSymmetric, ND.Directed)]
    et.append(ND.EdgeModel(g=C.noop_edge2, outdim=2))   # Fiducial(gss, gss): two independent outputs
    nw = ND.Network(g, [vt[i] for i in rng.integers(0, 3, n)], [et[i] for i in rng.integers(0, 4, g.ne())])
    on = oracle_for(nw)
    states = rng.random(on.lastidx_out)
    res = [on.aggregate(np.zeros(on.lastidx_aggr), states, m) for m in ("seq", "inv", "naive")]
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    off = np.arange(1, on.lastidx_aggr + 1, dtype=float)
    got = on.aggregate(off.copy(), states, "seq")
    np.testing.assert_allclose(got - res[0], off, rtol=0, atol=1e-12)
    # threaded restatement == sequential, bitwise
    assert np.array_equal(on.aggregate(np.zeros(on.lastidx_aggr), states, "inv", nthreads=4), res[0])


def test_threaded_equals_sequential_bitwise(ND):
    nw, u, p = _random_mixed(ND, n=500, directed=False)
    on = oracle_for(nw)
    assert np.array_equal(on.rhs(u, p), on.rhs_threaded(u, p, nthreads=4))


def test_diffusion_closed_form(ND):
    # du = -L x (test/diffusion_test.jl:33, 80-90)
    g = ND.barabasi_albert(200, 5, seed=2)
    nw = ND.Network(g, ND.Lib.diffusion_test.odevertex(), ND.Lib.diffusion_test.staticedge())
    on = oracle_for(nw)
    n = g.nv()
    A = sp.coo_matrix((np.ones(g.ne()), (g.src - 1, g.dst - 1)), shape=(n, n))
    A = (A + A.T).tocsr()
    Lap = sp.diags(np.asarray(A.sum(axis=1)).ravel()) - A
    rng = np.random.default_rng(0)
    for _ in range(30):
        x = rng.normal(size=n)
        du = on.rhs(x, np.zeros(0))
        np.testing.assert_allclose(du, -(Lap @ x), rtol=0, atol=1e-12)


def test_kuramoto_incidence_closed_form(ND):
    # dtheta = omega - B (K o sin(B^T theta)) with oriented incidence B (oldexamples/paper.jl:10-11)
    g = ND.erdos_renyi(300, 1200, seed=5)
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    on = oracle_for(nw)
    n, m = g.nv(), g.ne()
    rng = np.random.default_rng(1)
    th = rng.uniform(-np.pi, np.pi, n)
    om = rng.normal(size=n)
    K = rng.uniform(1, 3, m)
    B = sp.coo_matrix((np.r_[np.ones(m), -np.ones(m)], (np.r_[g.src - 1, g.dst - 1], np.r_[np.arange(m), np.arange(m)])),
                      shape=(n, m)).tocsr()
    ref = om - B @ (K * np.sin(B.T @ th))
    du = on.rhs(th, np.r_[om, K])
    S = on.scale(th, np.r_[om, K])
    assert scaled_error(du, ref, S).max() < 1e-14


def test_swing_dense_bruteforce_and_invariant(ND):
    g = ND.barabasi_albert(60, 2, seed=9)
    nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    on = oracle_for(nw)
    n = g.nv()
    rng = np.random.default_rng(2)
    M, D, P, K = rng.uniform(1, 5, n), rng.uniform(0.1, 1, n), rng.normal(size=n), rng.uniform(5, 10, g.ne())
    p = np.empty(nw.pdim)
    p[0:3 * n:3], p[1:3 * n:3], p[2:3 * n:3], p[3 * n:] = M, D, P, K
    u = np.empty(2 * n)
    u[0::2], u[1::2] = rng.uniform(-np.pi, np.pi, n), rng.normal(size=n)
    du = on.rhs(u, p)
    Kd = np.zeros((n, n))
    Kd[g.src - 1, g.dst - 1] = K
    Kd = Kd + Kd.T
    th, w = u[0::2], u[1::2]
    acc = np.array([sum(Kd[i, j] * np.sin(th[j] - th[i]) for j in range(n) if Kd[i, j] != 0) for i in range(n)])
    np.testing.assert_array_equal(du[0::2], w)
    np.testing.assert_allclose(du[1::2], (P - D * w + acc) / M, rtol=1e-13, atol=1e-13)
    # sum_i M_i dw_i = sum P - sum D w  (antisymmetric coupling cancels)
    lhs = np.sum(M * du[1::2])
    assert abs(lhs - (P.sum() - np.sum(D * w))) < 1e-11 * np.sum(np.abs(M * du[1::2]) + 1)


def _swing_load_bruteforce(th, w, is_gen, M, D, P, tau, s, d, B, phi):
    """The heterogeneous network's equations written out pair by pair: the static line
    from s to d carries y = B sin(th_s - th_d - phi) into d and -y into s (AntiSymmetric,
    test/ComponentLibrary.jl:51-57 with a phase shift); swing vertices
    dth = w, dw = (P - D w + acc) / M (test/ComponentLibrary.jl:59-67); load vertices
    tau dth = P + acc."""
    n = len(th)
    acc = np.zeros(n)
    for i in range(n):
        for e in range(len(s)):
            if d[e] == i:
                acc[i] += B[e] * np.sin(th[s[e]] - th[i] - phi[e])
            if s[e] == i:
                acc[i] -= B[e] * np.sin(th[i] - th[d[e]] - phi[e])
    dth = np.where(is_gen, w, (P + acc) / tau)
    dw = (P - D * w + acc) / M
    return dth, dw


def test_load_vertex_and_phase_shifted_line_bruteforce(ND):
    """The heterogeneous swing/load network's first-order load vertex and phase-shifted
    static line (build-defined kinds) against their equations written out pair by pair,
    with phase shifts and time constants far from 0 and 1 so that a sign error on phi or
    a product instead of a quotient by tau shows; plus the invariant
    sum_gen M dw + sum_load tau dth = sum P - sum_gen D w: the src end of each line takes
    exactly the negated dst output (AntiSymmetric), so the coupling cancels, phase shift
    or not."""
    rng = np.random.default_rng(31)
    n = 40
    g = ND.random_geometric(n, 0.3, seed=31)
    is_gen = rng.random(n) < 0.4
    swing, load = ND.Lib.kuramoto_second(), ND.Lib.load_vertex()
    nw = ND.Network(g, [swing if x else load for x in is_gen], ND.Lib.static_line_edge())
    on = oracle_for(nw)
    v, e = nw.ranges()
    M, D, P, tau = rng.uniform(1, 5, n), rng.uniform(0.1, 1, n), rng.normal(size=n), rng.uniform(0.3, 3, n)
    B, phi = rng.uniform(1, 20, g.ne()), rng.uniform(0.2, 0.6, g.ne())
    p = np.full(nw.pdim, np.nan)
    u = np.full(nw.dim, np.nan)
    th, w = rng.uniform(-np.pi, np.pi, n), np.where(is_gen, rng.normal(size=n), 0.0)
    for i in range(n):
        pf, sf = v[i, 2, 0] - 1, v[i, 0, 0] - 1
        u[sf] = th[i]
        if is_gen[i]:
            p[pf:pf + 3] = M[i], D[i], P[i]      # psym [M, D, Pm] (ComponentLibrary.jl:64-66)
            u[sf + 1] = w[i]
        else:
            p[pf:pf + 2] = tau[i], P[i]
    for k in range(g.ne()):
        p[e[k, 3, 0] - 1:e[k, 3, 0] + 1] = B[k], phi[k]
    assert not np.isnan(p).any() and not np.isnan(u).any()
    du = on.rhs(u, p)
    dth_ref, dw_ref = _swing_load_bruteforce(th, w, is_gen, M, D, P, tau, g.src - 1, g.dst - 1, B, phi)
    dth = du[v[:, 0, 0] - 1]
    np.testing.assert_allclose(dth, dth_ref, rtol=1e-13, atol=1e-13)
    gen = np.flatnonzero(is_gen)
    np.testing.assert_allclose(du[v[gen, 0, 0]], dw_ref[gen], rtol=1e-13, atol=1e-13)
    # invariant: the coupling cancels line by line
    lhs = np.sum(M[gen] * du[v[gen, 0, 0]]) + np.sum(np.where(is_gen, 0.0, tau * dth))
    rhs = P.sum() - np.sum(D[gen] * w[gen])
    assert abs(lhs - rhs) < 1e-12 * (np.abs(P).sum() + B.sum())
    # the mistakes the check is there for
    wrong_phi = _swing_load_bruteforce(th, w, is_gen, M, D, P, tau, g.src - 1, g.dst - 1, B, -phi)[0]
    wrong_tau = _swing_load_bruteforce(th, w, is_gen, M, D, P, 1.0 / tau, g.src - 1, g.dst - 1, B, phi)[0]
    assert np.abs(dth - wrong_phi).max() > 1e-2 and np.abs(dth - wrong_tau).max() > 1e-2


def test_swing_ring_fixpoint(ND):
    # test/initialization_test.jl:15-38: ring of 4, P = [1,-1,1,-1], K = 6: each line carries 0.5
    g = ND.cycle_graph(4)
    nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    on = oracle_for(nw)
    d = np.arcsin(0.5 / 6)
    u = np.array([d, 0, 0, 0, d, 0, 0, 0], float)
    p = np.r_[np.tile([1.0, 0.1, 0.0], 4), np.full(4, 6.0)]
    p[2::3][:4] = [1, -1, 1, -1]
    du = on.rhs(u, p)
    assert np.abs(du).max() < 1e-14
    u2 = u.copy()
    u2[0] += 1e-3
    assert np.abs(on.rhs(u2, p)).max() > 1e-4


def test_size_checks(ND):
    nw, u, p = _random_mixed(ND, n=50)
    on = oracle_for(nw)
    assert on.check_sizes(nw.dim, nw.dim, nw.pdim) == 0
    assert on.check_sizes(nw.dim - 1, nw.dim, nw.pdim) == 1
    assert on.check_sizes(nw.dim, nw.dim, nw.pdim + 1) == 2
    with pytest.raises(ValueError):
        on.rhs(u[:-1], p)


def test_rk4_order_of_convergence(ND):
    # textbook RK4: global error ~ dt^4
    g = ND.erdos_renyi(50, 120, seed=1)
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    on = oracle_for(nw)
    rng = np.random.default_rng(4)
    p = np.r_[rng.normal(size=50), np.full(120, 1.0)]
    u0 = rng.uniform(-np.pi, np.pi, 50)
    ref = on.rk4(u0, p, 0.0, 1e-3, 1000)
    e1 = np.abs(on.rk4(u0, p, 0.0, 0.04, 25) - ref).max()
    e2 = np.abs(on.rk4(u0, p, 0.0, 0.02, 50) - ref).max()
    assert 12 < e1 / e2 < 20


def _incidence(g):
    n, m = g.nv(), g.ne()
    return sp.coo_matrix((np.r_[np.ones(m), -np.ones(m)], (np.r_[g.src - 1, g.dst - 1], np.r_[np.arange(m), np.arange(m)])),
                         shape=(n, m)).tocsr()


def kuramoto_jvp_closed_form(g, th, K, v):
    """J v of dtheta = omega - B (K o sin(B^T theta)): -B (K cos(B^T theta) o B^T v), with the
    per-element magnitude of its terms."""
    B = _incidence(g)
    c = K * np.cos(B.T @ th)
    Bv = B.T @ v
    return -(B @ (c * Bv)), abs(B) @ np.abs(c * Bv)


def swing_jvp_closed_form(g, u, M, D, K, v):
    """J v of the swing network (state theta, omega interleaved)."""
    th, vth, vom = u[0::2], v[0::2], v[1::2]
    a, sa = kuramoto_jvp_closed_form(g, th, K, vth)
    out, sc = np.empty_like(v), np.empty_like(v)
    out[0::2], sc[0::2] = vom, np.abs(vom)
    out[1::2] = 1.0 / M * (-D * vom + a)
    sc[1::2] = (np.abs(D * vom) + sa) / M
    return out, sc


def test_fd_jvp_checker_on_closed_forms(ND):
    # the finite-difference checker used for the GPU Jacobian-vector products agrees with
    # the closed-form Jacobians of the Kuramoto and swing networks
    from conftest import fd_jvp
    rng = np.random.default_rng(4)
    g = ND.erdos_renyi(200, 800, seed=6)
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    th, om, K = rng.uniform(-np.pi, np.pi, g.nv()), rng.normal(size=g.nv()), rng.uniform(1, 3, g.ne())
    v = rng.normal(size=g.nv())
    ref, S = kuramoto_jvp_closed_form(g, th, K, v)
    assert scaled_error(fd_jvp(oracle_for(nw), th, v, np.r_[om, K]), ref, S).max() < 1e-8
    g = ND.barabasi_albert(300, 2, seed=7)
    nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    n = g.nv()
    M, D, P, K = rng.uniform(1, 5, n), rng.uniform(0.1, 1, n), rng.normal(size=n), rng.uniform(5, 10, g.ne())
    u, v = rng.normal(size=2 * n), rng.normal(size=2 * n)
    p = np.r_[np.stack([M, D, P], 1).ravel(), K]
    ref, S = swing_jvp_closed_form(g, u, M, D, K, v)
    assert scaled_error(fd_jvp(oracle_for(nw), u, v, p), ref, S).max() < 1e-8
