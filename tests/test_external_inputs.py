"""External inputs (src/external_inputs.jl): the ExtMap, collect_externals!, and vertex
models that read them, on the network of docs/src/external_inputs.md:62-99 (a
PI-controlled current source that reads the voltage of a capacitor load through a
resistor) and on random networks of such pairs."""
import numpy as np
import pytest

from conftest import fd_jvp, oracle_for, scaled_error


def _docs_network(ND):
    # Network([source, load], [resistor]): the resistor runs source -> load
    g = ND.SimpleGraph(2, [1], [2])
    src = ND.Lib.pi_source(extin=[ND.VIndex("load", "V")])
    return ND.Network(g, [src, ND.Lib.capacitor_load()], ND.Lib.resistor())


def _random_pairs(ND, npair, seed):
    """npair sources and npair loads on a random graph of resistors; every source reads
    the voltage of a random load."""
    rng = np.random.default_rng(seed)
    n = 2 * npair
    g = ND.erdos_renyi(n, 3 * n, seed=seed)
    srcs = rng.choice(n, npair, replace=False)
    is_src = np.zeros(n, bool)
    is_src[srcs] = True
    loads = np.flatnonzero(~is_src)
    vm = []
    for i in range(n):
        if is_src[i]:
            ref = ND.VIndex(int(rng.choice(loads)) + 1, "V")
            vm.append(ND.Lib.pi_source(extin=[ref]))
        else:
            vm.append(ND.Lib.capacitor_load())
    nw = ND.Network(g, vm, ND.Lib.resistor())
    v, _ = nw.ranges()
    u = rng.normal(0.8, 0.2, nw.dim)
    p = np.empty(nw.pdim)
    for i in range(n):
        a, b = v[i, 2, 0] - 1, v[i, 2, 1]
        p[a:b] = [rng.uniform(0.5, 2.0), 1.0, rng.uniform(0.1, 1.0), rng.uniform(1.0, 10.0)] if is_src[i] \
            else [rng.uniform(0.5, 2.0)]
    _, e = nw.ranges()
    p[e[:, 3, 0] - 1] = rng.uniform(0.05, 0.5, nw.ne())
    return nw, u, p


def test_docs_example_layout_and_extmap(ND, oracle):
    # source (V, xi) registers first, then the load (V): the source reads u[3]
    nw = _docs_network(ND)
    assert nw.dim == 3 and nw.pdim == 4 + 1 + 1
    assert nw.info.lastidx_extbuf == 1
    ve, ee = nw.ext_ranges()
    np.testing.assert_array_equal(ve, [[1, 1], [2, 1]])
    assert ee.shape == (1, 2) and ee[0, 1] == ee[0, 0] - 1
    np.testing.assert_array_equal(nw.extmap(), [3])
    on = oracle_for(nw)
    ov, oe = on.ext_ranges()
    np.testing.assert_array_equal(ov, ve)
    np.testing.assert_array_equal(oe, ee)
    # the symbolic forms: by index, by name, by state position
    src = ND.Lib.pi_source(extin=[ND.VIndex(2, 1)])
    nw2 = ND.Network(ND.SimpleGraph(2, [1], [2]), [src, ND.Lib.capacitor_load()], ND.Lib.resistor())
    np.testing.assert_array_equal(nw2.extmap(), [3])


def test_extmap_outputs_and_errors(ND):
    g = ND.SimpleGraph(3, [1, 2], [2, 3])
    load = ND.Lib.capacitor_load()
    # an output of a non-feed-forward component (an ODE edge's Fiducial StateMask) is an
    # OutBufIdx; an output of a feed-forward edge is refused (external_inputs.jl:38-41)
    ode = ND.EdgeModel(f=ND.components.real_ode_edge, sym=["e_dst", "e_src"], g=ND.Fiducial(dst=1, src=2),
                       outsym=["flow"], name="ode")
    src = ND.Lib.pi_source(extin=[ND.EIndex(2, "flow")])
    nw = ND.Network(g, [src, load, load], [ode, ode])
    _, e = nw.ranges()
    assert nw.extmap()[0] == -e[1, 2, 0]
    src = ND.Lib.pi_source(extin=[ND.EIndex(1, "e_src")])
    nw = ND.Network(g, [src, load, load], [ode, ode])
    assert nw.extmap()[0] == e[0, 0, 0] + 1
    with pytest.raises(ND.ArgumentError, match="feed-forward"):
        ND.Network(g, [ND.Lib.pi_source(extin=[ND.EIndex(1, "P")]), ND.Lib.kuramoto_first(),
                       ND.Lib.kuramoto_first()], ND.Lib.kuramoto_edge())
    with pytest.raises(ND.ArgumentError, match="Cannot resolve"):
        ND.Network(g, [ND.Lib.pi_source(extin=[ND.VIndex(2, "nope")]), load, load], ND.Lib.resistor())
    with pytest.raises(ND.ArgumentError, match="Cannot resolve"):
        ND.Network(g, [ND.Lib.pi_source(extin=[ND.VIndex(9, "V")]), load, load], ND.Lib.resistor())
    with pytest.raises(ND.ArgumentError, match="external input"):
        ND.Lib.pi_source(extin=[])
    with pytest.raises(ND.ArgumentError):
        ND.Lib.pi_source(extin=["load.V"])


@pytest.mark.gpu
def test_docs_example_rhs_and_rk4_match_oracle(ND, dev):
    import torch
    nw = _docs_network(ND)
    on = oracle_for(nw)
    u0 = np.array([0.5, 0.0, 0.5])
    p = nw.default_p()
    assert np.isfinite(p).all()
    rng = np.random.default_rng(5)
    for t in (0.0, 1.3, 17.0):
        for u in (u0, rng.normal(size=3)):
            du = torch.empty(3, dtype=torch.float64, device=dev)
            nw(du, torch.from_numpy(u).to(dev), torch.from_numpy(p).to(dev), t)
            ref = on.rhs(u, p, t)
            err = scaled_error(du.cpu().numpy(), ref, on.scale(u, p, t))
            assert err.max() <= 1e-12, (t, err)
    # docs: solve on (0, 100); fixed-step RK4 here, against the oracle's
    ud = torch.from_numpy(u0.copy()).to(dev)
    nw.rk4(ud, torch.from_numpy(p).to(dev), 0.0, 0.01, 10_000)
    ref = on.rk4(u0.copy(), p, 0.0, 0.01, 10_000)
    np.testing.assert_allclose(ud.cpu().numpy(), ref, rtol=0, atol=1e-10)
    # the controller holds the load voltage near Vref = 1 against the periodic sink
    assert abs(ref[2] - 1.0) < 0.1


@pytest.mark.gpu
@pytest.mark.parametrize("npair,seed", [(40, 1), (3000, 2)])
def test_random_external_input_networks_match_oracle(ND, dev, npair, seed):
    import torch
    nw, u, p = _random_pairs(ND, npair, seed)
    on = oracle_for(nw)
    pd = torch.from_numpy(p).to(dev)
    for gm in (2, 3):          # neighbour outputs read from u / from the compact copy
        nw.tune(gather_mode=gm)
        for t in (0.0, 2.5):
            du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
            nw(du, torch.from_numpy(u).to(dev), pd, t)
            err = scaled_error(du.cpu().numpy(), on.rhs(u, p, t), on.scale(u, p, t))
            assert err.max() <= 1e-12, (gm, t, err.max())
    # the Jacobian-vector product includes the external inputs' columns
    v = np.random.default_rng(seed + 7).normal(size=nw.dim)
    du, jv = (torch.empty(nw.dim, dtype=torch.float64, device=dev) for _ in range(2))
    nw.jvp(du, jv, torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev), pd, 0.0)
    ref = fd_jvp(on, u, v, p)
    np.testing.assert_allclose(jv.cpu().numpy(), ref, rtol=1e-7, atol=1e-7 * np.abs(ref).max())


@pytest.mark.gpu
def test_collect_externals_matches_oracle(ND, dev):
    # the reference's "test extmap performance": a shuffled map over states and outputs,
    # device gather against the CPU one, bit for bit
    import torch
    rng = np.random.default_rng(11)
    n = 400
    g = ND.erdos_renyi(n, 2 * n, seed=11)
    ode = ND.EdgeModel(f=ND.components.real_ode_edge, sym=["e_dst", "e_src"], g=ND.Fiducial(dst=1, src=2),
                       outsym=["flow"], name="ode")
    refs = []
    for i in range(n):
        if rng.random() < 0.5:
            refs.append(ND.VIndex(int(rng.integers(1, n + 1)), "x"))
        else:
            refs.append(ND.EIndex(int(rng.integers(1, g.ne() + 1)), "flow" if rng.random() < 0.5 else "e_src"))
    fext = ND.DeviceFunction("noop_ext", "vertex_f", ND._abi.ND_VM_NOOP, 1, 0)
    vm = [ND.VertexModel(f=fext, dim=1, g=1, extin=[r]) for r in refs]
    nw = ND.Network(g, vm, ode)
    m = nw.extmap()
    assert (m > 0).any() and (m < 0).any()
    u = rng.normal(size=nw.dim)
    ud = torch.from_numpy(u).to(dev)
    o, _ = nw.get_buffers(ud, torch.zeros(1, dtype=torch.float64, device=dev))
    ext = torch.full((nw.info.lastidx_extbuf,), np.nan, dtype=torch.float64, device=dev)
    nw.collect_externals(ext, ud, o)
    on = oracle_for(nw)
    o_ref, _ = on.buffers(u, np.zeros(0))
    np.testing.assert_array_equal(o.cpu().numpy(), o_ref)
    np.testing.assert_array_equal(ext.cpu().numpy(), on.collect_externals(u, o_ref))


def test_oracle_pi_source_network_invariants(ND):
    # Pins the oracle's three models without restating them: the resistor currents cancel
    # in sum C_i dV_i = sum_sources (Kp d + Ki xi) - n_loads (1 + 0.1 sin t), with
    # d = Vref - V(referenced load); and d(xi)/dt reads exactly the referenced load's V
    nw, u, p = _random_pairs(ND, 30, 4)
    on = oracle_for(nw)
    v, _ = nw.ranges()
    m = nw.extmap()
    ve, _ = nw.ext_ranges()
    for t in (0.0, 0.7, 3.1):
        du = on.rhs(u, p, t)
        lhs, rhs, nload = 0.0, 0.0, 0
        for i, model in enumerate(nw.vertexm):
            x0, p0 = v[i, 0, 0] - 1, v[i, 2, 0] - 1
            lhs += p[p0] * du[x0]
            if model.f is ND.components.f_source:
                C_, Vref, Ki, Kp = p[p0:p0 + 4]
                d = Vref - u[m[ve[i, 0] - 1] - 1]
                assert du[x0 + 1] == d
                rhs += Kp * d + Ki * u[x0 + 1]
            else:
                nload += 1
        rhs -= nload * (1 + 0.1 * np.sin(t))
        assert abs(lhs - rhs) <= 1e-12 * (abs(rhs) + nload)


@pytest.mark.gpu
def test_per_vertex_kernel_hub_classes_match_oracle(ND, dev):
    # the per-vertex kernel (here because of the resistor edges) on a Barabasi-Albert graph:
    # big hubs (a block each), mid hubs (a wave each) and the rest, in both gather modes,
    # for the RHS and for fused RK4 stages
    import torch
    g = ND.barabasi_albert(20_000, 2, seed=9)
    nw = ND.Network(g, ND.Lib.diffusion_vertex(), ND.Lib.resistor())
    rng = np.random.default_rng(9)
    u = rng.normal(size=nw.dim)
    p = rng.uniform(0.5, 2.0, nw.pdim)
    on = oracle_for(nw)
    ref, S = on.rhs(u, p), on.scale(u, p)
    pd = torch.from_numpy(p).to(dev)
    for hub, big in ((48, 128), (8, 32)):
        nw.tune(hub_degree=hub, big_hub_degree=big)
        for gm in (2, 3):
            nw.tune(gather_mode=gm)
            du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
            nw(du, torch.from_numpy(u).to(dev), pd, 0.0)
            err = scaled_error(du.cpu().numpy(), ref, S)
            assert err.max() <= 1e-12, (hub, gm, err.max())
            # a step inside RK4's stability region for the hubs' eigenvalues (~2 * 564 / 0.5)
            ud = torch.from_numpy(u).to(dev)
            nw.rk4(ud, pd, 0.0, 2e-4, 5)
            np.testing.assert_allclose(ud.cpu().numpy(), on.rk4(u.copy(), p, 0.0, 2e-4, 5), rtol=0, atol=1e-12)
