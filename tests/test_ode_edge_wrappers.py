"""ODE edges whose output is a StateMask inside AntiSymmetric / Symmetric / Directed
(component_functions.jl:81-175: y_dst = x[mask], y_src = -y_dst / y_dst / nothing), and
ODE edges that read external inputs (test/external_inputs_test.jl:41-53).

The oracle is pinned three ways:
- a one-state AntiSymmetric lag edge carries exactly the dynamics of the two-state
  Fiducial edge of test/ComponentLibrary.jl:30-40 started at e_src = -e_dst (sin is odd,
  so the pair stays antisymmetric bit for bit);
- closed forms of the three wrappers on a single directed edge;
- the reference test's e_noff edge (de = ext, AntiSymmetric(1), extin = EIndex(2, :a)):
  hand-derived layout and ExtMap, and de equal to the referenced state;
- lag lines next to vertices whose output is computed (a parameter, NoFeedForward, as the
  static vertex of docs/examples/heterogeneous_system.jl): the equations written out.
"""
import numpy as np
import pytest

from conftest import fd_jvp, oracle_for, scaled_error


def _pair_networks(ND, n=40, seed=2):
    L = ND.Lib
    g = ND.watts_strogatz(n, 4, 0.5, seed=seed)
    nw2 = ND.Network(g, L.kuramoto_second(), L.diffusion_odeedge())      # e_dst, e_src; Fiducial
    nw1 = ND.Network(g, L.kuramoto_second(), L.ode_lag_edge())          # e; AntiSymmetric(1)
    rng = np.random.default_rng(seed)
    nv, ne = g.nv(), g.ne()
    x = np.empty(2 * nv)
    x[0::2], x[1::2] = rng.uniform(-np.pi, np.pi, nv), rng.normal(size=nv)
    e = rng.normal(size=ne)
    u2 = np.r_[x, np.stack([e, -e], 1).ravel()]
    u1 = np.r_[x, e]
    p = np.r_[np.stack([rng.uniform(1, 5, nv), rng.uniform(0.1, 1, nv), rng.normal(size=nv)], 1).ravel(),
              rng.uniform(0.5, 2.0, ne)]
    return nw1, nw2, u1, u2, p


def test_antisymmetric_lag_edge_equals_fiducial_pair(ND):
    nw1, nw2, u1, u2, p = _pair_networks(ND)
    assert nw1.pdim == nw2.pdim and nw2.dim - nw1.dim == nw1.ne()
    o1, o2 = oracle_for(nw1), oracle_for(nw2)
    nv2 = 2 * nw1.nv()
    d1, d2 = o1.rhs(u1, p), o2.rhs(u2, p)
    np.testing.assert_array_equal(d1[:nv2], d2[:nv2])                  # vertices: same inputs
    np.testing.assert_array_equal(d1[nv2:], d2[nv2:][0::2])            # de = de_dst
    np.testing.assert_array_equal(d1[nv2:], -d2[nv2:][1::2])           # = -de_src
    r1 = o1.rk4(u1, p, 0.0, 0.01, 200)
    r2 = o2.rk4(u2, p, 0.0, 0.01, 200)
    np.testing.assert_array_equal(r1[:nv2], r2[:nv2])
    np.testing.assert_array_equal(r1[nv2:], r2[nv2:][0::2])


@pytest.mark.parametrize("wrapper,src_in", [("anti", -1.0), ("sym", 1.0), ("directed", 0.0)])
def test_wrapper_closed_forms(ND, wrapper, src_in):
    # diffusion vertices sum their inputs (dv = acc): one edge 1 -> 2 with state e hands
    # vertex 2 its dst output e and vertex 1 its src output
    L = ND.Lib
    nw = ND.Network(ND.SimpleDiGraph(2, [1], [2]), L.diffusion_vertex(), L.ode_lag_edge(wrapper))
    u = np.array([0.3, -0.4, 0.7])
    tau = 2.0
    du = oracle_for(nw).rhs(u, np.array([tau]))
    assert du[0] == src_in * 0.7 and du[1] == 0.7
    assert du[2] == 1.0 / tau * (np.sin(0.3 - -0.4) - 0.7)
    o, _ = oracle_for(nw).buffers(u, np.array([tau]))
    _, e = nw.ranges()
    assert o[e[0, 2, 0] - 1] == 0.7 and (wrapper == "directed" or o[e[0, 1, 0] - 1] == src_in * 0.7)


def test_external_inputs_test_edges(ND):
    # test/external_inputs_test.jl:41-53: v_noff reads VIndex(2, :a), e_noff EIndex(2, :a)
    L = ND.Lib
    g = ND.SimpleGraph(3, [1, 2], [2, 3])
    e_noff = L.ext_driven_edge([ND.EIndex(2, "a")])
    nw = ND.Network(g, L.diffusion_vertex(), e_noff)
    v, e = nw.ranges()
    ve, ee = nw.ext_ranges()
    # states: v 1..3, e1.a = 4, e2.a = 5; the edges' external inputs follow the vertices' (none)
    assert [e[0, 0, 0], e[1, 0, 0]] == [4, 5]
    np.testing.assert_array_equal(ee, [[1, 1], [2, 2]])
    np.testing.assert_array_equal(nw.extmap(), [5, 5])
    on = oracle_for(nw)
    u = np.array([0.1, 0.2, 0.3, -1.5, 2.5])
    du = on.rhs(u, np.zeros(0))
    assert du[3] == 2.5 and du[4] == 2.5                              # de = ext = e2.a
    # AntiSymmetric(1) outputs summed per vertex: v1 gets -e1, v2 gets e1 then -e2, v3 e2
    np.testing.assert_array_equal(du[:3], [1.5, -1.5 - 2.5, 2.5])
    # the output names of single-sided wrappers (component_functions.jl:1064-1075):
    # AntiSymmetric src "₋a" is the negated state, an OutBufIdx
    nw2 = ND.Network(g, L.diffusion_vertex(), [L.ode_lag_edge(), L.ext_driven_edge([ND.EIndex(1, "₋e")])])
    _, e2 = nw2.ranges()
    assert nw2.extmap()[0] == -e2[0, 1, 0]
    with pytest.raises(ND.ArgumentError, match="external input"):
        L.ext_driven_edge([])


def _ext_network(ND, n=300, seed=5):
    """Swing vertices; lag edges (AntiSymmetric) and externally driven edges whose input is a
    vertex state, a lag edge's state or the negated src output "₋e" of a lag edge.
    Returns the network, u, p and per driven edge (edge, reference kind, referenced id)."""
    L = ND.Lib
    rng = np.random.default_rng(seed)
    g = ND.watts_strogatz(n, 4, 0.5, seed=seed)
    ne = g.ne()
    lag = L.ode_lag_edge()
    em, refs = [], []
    for i in range(ne):
        if i % 3 == 0:
            em.append(lag)
            continue
        kind = int(rng.integers(0, 3))
        j = 3 * int(rng.integers(0, ne // 3))                          # a lag edge (0-based)
        v = int(rng.integers(0, n))
        ref = [ND.VIndex(v + 1, "omega"), ND.EIndex(j + 1, "e"), ND.EIndex(j + 1, "₋e")][kind]
        em.append(L.ext_driven_edge([ref]))
        refs.append((i, kind, v if kind == 0 else j))
    nw = ND.Network(g, L.kuramoto_second(), em)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2.0, nw.pdim)
    return nw, u, p, refs


def test_ext_network_oracle_invariants(ND):
    nw, u, p, refs = _ext_network(ND)
    m = nw.extmap()
    assert {k for _, k, _ in refs} == {0, 1, 2} and (m > 0).any() and (m < 0).any()
    du = oracle_for(nw).rhs(u, p)
    v, e = nw.ranges()
    for i, kind, j in refs:
        # de of a driven edge is its input: omega of a vertex, a lag state, or minus it
        want = u[v[j, 0, 0]] if kind == 0 else (u[e[j, 0, 0] - 1] if kind == 1 else -u[e[j, 0, 0] - 1])
        assert du[e[i, 0, 0] - 1] == want, (i, kind, j)


def _lines_next_to_static_vertices(ND, n=60, seed=6):
    """Swing vertices and every seventh vertex a static one (output = its parameter),
    joined by one-state AntiSymmetric lag lines: the lines read computed outputs."""
    L = ND.Lib
    g = ND.watts_strogatz(n, 4, 0.5, seed=seed)
    static = np.arange(n) % 7 == 0
    vm = [L.static_vertex() if s else L.kuramoto_second() for s in static]
    nw = ND.Network(g, vm, L.ode_lag_edge())
    rng = np.random.default_rng(seed)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2.0, nw.pdim)
    return nw, g, u, p, static


def test_lag_lines_next_to_static_vertices_equations(ND):
    nw, g, u, p, static = _lines_next_to_static_vertices(ND)
    v, e = nw.ranges()
    n = g.nv()
    theta = np.array([p[v[i, 2, 0] - 1] if static[i] else u[v[i, 0, 0] - 1] for i in range(n)])
    acc = np.zeros(n)
    de = np.zeros(g.ne())
    for k, (a, b) in enumerate(zip(g.src - 1, g.dst - 1)):
        x, tau = u[e[k, 0, 0] - 1], p[e[k, 3, 0] - 1]
        de[k] = 1.0 / tau * (np.sin(theta[a] - theta[b]) - x)       # ode_lag_f
        acc[b] += x
        acc[a] -= x                                                 # AntiSymmetric src side
    du = oracle_for(nw).rhs(u, p)
    for k in range(g.ne()):
        np.testing.assert_allclose(du[e[k, 0, 0] - 1], de[k], rtol=1e-14, atol=1e-14)
    for i in np.flatnonzero(~static):
        x0, p0 = v[i, 0, 0] - 1, v[i, 2, 0] - 1
        M, D, Pm = p[p0: p0 + 3]
        np.testing.assert_allclose(du[x0: x0 + 2], [u[x0 + 1], (Pm - D * u[x0 + 1] + acc[i]) / M],
                                   rtol=1e-13, atol=1e-13)


@pytest.mark.gpu
def test_lag_lines_next_to_static_vertices_on_the_device(ND, dev):
    import torch
    nw, g, u, p, _ = _lines_next_to_static_vertices(ND, n=3000)
    _check_device(ND, dev, nw, u, p)
    on = oracle_for(nw)
    du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    nw(du, torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev), 0.0, perturb=np.zeros(0), perturb_maps={})
    err = scaled_error(du.cpu().numpy(), on.rhs(u, p), on.scale(u, p))
    assert err.max() <= 1e-12, err.max()


def _check_device(ND, dev, nw, u, p, rtol=1e-12):
    import torch
    on = oracle_for(nw)
    ref, S = on.rhs(u, p), on.scale(u, p)
    pd = torch.from_numpy(p if len(p) else np.zeros(1)).to(dev)
    mode = nw.tuning()["gather_mode"]
    for gm in (2, 3):
        nw.tune(gather_mode=gm)
        du = torch.full((nw.dim,), np.nan, dtype=torch.float64, device=dev)
        nw(du, torch.from_numpy(u).to(dev), pd, 0.0)
        err = scaled_error(du.cpu().numpy(), ref, S)
        assert err.max() <= rtol, (gm, err.max(), err.argmax())
    nw.tune(gather_mode=mode)
    # buffers: the wrapped outputs (-e on the AntiSymmetric src side) are copies, bit for
    # bit; feed-forward outputs through sin may differ from the host's in the last ulp
    o, agg = nw.get_buffers(torch.from_numpy(u).to(dev), pd)
    o_ref, agg_ref = on.buffers(u, p)
    np.testing.assert_allclose(o.cpu().numpy(), o_ref, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(agg.cpu().numpy(), agg_ref, rtol=1e-13, atol=1e-13)
    # JVP against finite differences of the oracle
    v = np.random.default_rng(9).normal(size=nw.dim)
    du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    jv = torch.empty_like(du)
    nw.jvp(du, jv, torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev), pd, 0.0)
    np.testing.assert_allclose(jv.cpu().numpy(), fd_jvp(on, u, v, p), rtol=1e-8, atol=1e-8)
    # RK4 stages (ODE-edge epilogues)
    ur = torch.from_numpy(u.copy()).to(dev)
    nw.rk4(ur, pd, 0.0, 0.002, 50)
    np.testing.assert_allclose(ur.cpu().numpy(), on.rk4(u, p, 0.0, 0.002, 50), rtol=0, atol=1e-11)


@pytest.mark.gpu
@pytest.mark.parametrize("wrapper", ["anti", "sym", "directed"])
def test_wrapped_lag_edges_on_the_device(ND, dev, wrapper):
    L = ND.Lib
    g = ND.barabasi_albert(3000, 2, seed=4)
    nw = ND.Network(g, L.kuramoto_second(), L.ode_lag_edge(wrapper))
    rng = np.random.default_rng(4)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2.0, nw.pdim)
    nw.tune(hub_degree=8)                       # hub segments too
    _check_device(ND, dev, nw, u, p)


@pytest.mark.gpu
def test_lag_edges_next_to_other_edge_batches(ND, dev):
    # several edge batches: the kernels read each entry's wrapper at run time
    L = ND.Lib
    g = ND.watts_strogatz(2000, 4, 0.5, seed=8)
    et = [L.ode_lag_edge(), L.ode_lag_edge("sym", name="lag_sym"), L.kuramoto_edge(), L.diffusion_odeedge()]
    rng = np.random.default_rng(8)
    nw = ND.Network(g, L.kuramoto_second(), [et[i] for i in rng.integers(0, len(et), g.ne())])
    _check_device(ND, dev, nw, rng.normal(size=nw.dim), rng.uniform(0.5, 2.0, nw.pdim))


@pytest.mark.gpu
def test_driven_edges_with_external_inputs_on_the_device(ND, dev):
    import torch
    nw, u, p, _ = _ext_network(ND)
    _check_device(ND, dev, nw, u, p)
    # collect_externals on the device, bit for bit with the oracle
    on = oracle_for(nw)
    o, _ = nw.get_buffers(torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev))
    ext = torch.full((nw.info.lastidx_extbuf,), np.nan, dtype=torch.float64, device=dev)
    nw.collect_externals(ext, torch.from_numpy(u).to(dev), o)
    np.testing.assert_array_equal(ext.cpu().numpy(), on.collect_externals(u, o.cpu().numpy()))


@pytest.mark.gpu
def test_vertex_external_input_from_a_negated_edge_output(ND, dev):
    # the per-vertex kernel resolves "₋e" (the AntiSymmetric src output) to minus the state
    L = ND.Lib
    g = ND.SimpleGraph(3, [1, 2], [2, 3])
    src = L.pi_source(extin=[ND.EIndex(1, "₋e")])
    nw = ND.Network(g, [src, L.capacitor_load(), L.capacitor_load()], [L.ode_lag_edge(), L.ode_lag_edge()])
    _, e = nw.ranges()
    assert nw.extmap()[0] == -e[0, 1, 0]
    rng = np.random.default_rng(3)
    u = rng.normal(size=nw.dim)
    p = nw.default_p()
    _check_device(ND, dev, nw, u, p)
