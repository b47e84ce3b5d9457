"""Injector nodes and LoopbackConnections (src/post_utils.jl:105-228, coreloop.jl:44-55),
on the network of test/loopback_test.jl: a capacitor hub fed by a voltage source through
a resistor, with a resistor injector and an inductor injector.  Its twin is one CRL
vertex; both must follow the same trajectory (the reference test's own check)."""
import numpy as np
import pytest

from conftest import fd_jvp, oracle_for, scaled_error

R_LINE = 1.0


def _injector_network(ND):
    L = ND.Lib
    g = ND.SimpleDiGraph(4, [1, 3, 4], [2, 2, 2])      # vs->hub, R_injector->hub, L_injector->hub
    vm = [L.voltage_source(), L.capacitor(), L.resistor_injector(), L.inductor_injector()]
    em = [L.resistor(), L.loopback_connection(), L.loopback_connection()]
    nw = ND.Network(g, vm, em)
    p = nw.default_p()
    _, e = nw.ranges()
    p[e[0, 3, 0] - 1] = R_LINE
    return nw, p


def _crl_network(ND):
    L = ND.Lib
    nw = ND.Network(ND.SimpleDiGraph(2, [1], [2]), [L.voltage_source(), L.crl()], L.resistor())
    p = nw.default_p()
    _, e = nw.ranges()
    p[e[0, 3, 0] - 1] = R_LINE
    return nw, p


def _states(nw, which):
    v, _ = nw.ranges()
    return {name: v[i, 0, 0] - 1 + j for i, m in enumerate(nw.vertexm) for j, name in enumerate(m.sym)
            if (i, name) in which}


def test_construction_rules(ND):
    L = ND.Lib
    # a feed-forward vertex must be a leaf with one LoopbackConnection to its hub
    g = ND.SimpleDiGraph(3, [1, 2], [2, 3])
    with pytest.raises(ND.ArgumentError, match="leave nodes"):
        ND.Network(g, [L.capacitor(), L.resistor_injector(), L.capacitor()],
                   [L.loopback_connection(), L.loopback_connection()])
    g = ND.SimpleDiGraph(2, [1], [2])
    with pytest.raises(ND.ArgumentError, match="leave nodes"):
        ND.Network(g, [L.resistor_injector(), L.capacitor()], L.resistor())
    # loopbacks originate at leaves
    g = ND.SimpleDiGraph(3, [1, 1], [2, 3])
    with pytest.raises(ND.ArgumentError, match="originate from leaf"):
        ND.Network(g, [L.capacitor(), L.inductor_injector(), L.capacitor()],
                   [L.loopback_connection(), L.loopback_connection()])
    nw, p = _injector_network(ND)
    assert nw.dim == 2 and np.isfinite(p).all()


def test_oracle_injector_network_equals_crl_vertex(ND):
    # the two formulations of test/loopback_test.jl, integrated by the oracle's RK4
    nwa, pa = _injector_network(ND)
    nwb, pb = _crl_network(ND)
    ia = _states(nwa, {(1, "v"), (3, "i")})
    ib = _states(nwb, {(1, "v"), (1, "iL")})
    ua = oracle_for(nwa).rk4(np.zeros(2), pa, 0.0, 0.001, 10_000)
    ub = oracle_for(nwb).rk4(np.zeros(2), pb, 0.0, 0.001, 10_000)
    np.testing.assert_allclose(ua[ia["v"]], ub[ib["v"]], rtol=0, atol=1e-12)
    np.testing.assert_allclose(ua[ia["i"]], ub[ib["iL"]], rtol=0, atol=1e-12)
    # the hub settles where the source's current equals the injectors' draw: v -> 0
    # through the inductor's short, its current -> V / R_line
    assert abs(ua[ia["i"]] - 1.0 / R_LINE) < 0.05


@pytest.mark.gpu
def test_injector_network_on_the_device(ND, dev):
    import torch
    for nw, p in (_injector_network(ND), _crl_network(ND)):
        on = oracle_for(nw)
        rng = np.random.default_rng(4)
        pd = torch.from_numpy(p).to(dev)
        for u in (np.zeros(nw.dim), rng.normal(size=nw.dim)):
            du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
            nw(du, torch.from_numpy(u).to(dev), pd, 0.0)
            err = scaled_error(du.cpu().numpy(), on.rhs(u, p), on.scale(u, p))
            assert err.max() <= 1e-12, err
        # buffers: the injectors' outputs and the loopback flows into the hub
        u = rng.normal(size=nw.dim)
        o, agg = nw.get_buffers(torch.from_numpy(u).to(dev), pd)
        o_ref, agg_ref = on.buffers(u, p)
        np.testing.assert_allclose(o.cpu().numpy(), o_ref, rtol=1e-15, atol=1e-15)
        np.testing.assert_allclose(agg.cpu().numpy(), agg_ref, rtol=1e-14, atol=1e-15)
        # JVPs through the computed outputs
        v = rng.normal(size=nw.dim)
        du, jv = (torch.empty(nw.dim, dtype=torch.float64, device=dev) for _ in range(2))
        nw.jvp(du, jv, torch.from_numpy(u).to(dev), torch.from_numpy(v).to(dev), pd, 0.0)
        np.testing.assert_allclose(jv.cpu().numpy(), fd_jvp(on, u, v, p), rtol=1e-8, atol=1e-9)
    # RK4 on the device: the injector network against the oracle and against its twin
    nwa, pa = _injector_network(ND)
    nwb, pb = _crl_network(ND)
    ua = torch.zeros(2, dtype=torch.float64, device=dev)
    ub = torch.zeros(2, dtype=torch.float64, device=dev)
    nwa.rk4(ua, torch.from_numpy(pa).to(dev), 0.0, 0.001, 10_000)
    nwb.rk4(ub, torch.from_numpy(pb).to(dev), 0.0, 0.001, 10_000)
    ref = oracle_for(nwa).rk4(np.zeros(2), pa, 0.0, 0.001, 10_000)
    np.testing.assert_allclose(ua.cpu().numpy(), ref, rtol=0, atol=1e-12)
    ia = _states(nwa, {(1, "v"), (3, "i")})
    ib = _states(nwb, {(1, "v"), (1, "iL")})
    a, b = ua.cpu().numpy(), ub.cpu().numpy()
    assert abs(a[ia["v"]] - b[ib["v"]]) <= 1e-12 and abs(a[ia["i"]] - b[ib["iL"]]) <= 1e-12
