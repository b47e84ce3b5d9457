"""Time integration: Tsit5 tableau pinned by its order conditions (CPU), the batched
ensemble integrator against per-instance references (GPU)."""
import itertools

import numpy as np
import pytest

from conftest import oracle_for


def _tableau(ND):
    A, bt, c = ND.ensemble.tsit5_tableau()
    b = A[6].copy()
    return A, b, b - bt, c


def _order_conditions(A, b, c, order):
    """Butcher order conditions up to `order` (trees of order <= 5), as residuals."""
    r = []
    e = np.ones(len(b))
    Ac, Ac2, Ac3 = A @ c, A @ c ** 2, A @ c ** 3
    AAc = A @ Ac
    if order >= 1:
        r += [b @ e - 1]
    if order >= 2:
        r += [b @ c - 1 / 2]
    if order >= 3:
        r += [b @ c ** 2 - 1 / 3, b @ Ac - 1 / 6]
    if order >= 4:
        r += [b @ c ** 3 - 1 / 4, b @ (c * Ac) - 1 / 8, b @ Ac2 - 1 / 12, b @ AAc - 1 / 24]
    if order >= 5:
        r += [b @ c ** 4 - 1 / 5, b @ (c ** 2 * Ac) - 1 / 10, b @ (c * Ac2) - 1 / 15, b @ (c * AAc) - 1 / 30,
              b @ (Ac * Ac) - 1 / 20, b @ Ac3 - 1 / 20, b @ (A @ (c * Ac)) - 1 / 40, b @ (A @ Ac2) - 1 / 60,
              b @ (A @ AAc) - 1 / 120]
    return np.abs(r)


def test_tsit5_tableau_order_conditions(ND):
    A, b, bhat, c = _tableau(ND)
    assert np.allclose(np.tril(A, -1), A)                       # explicit
    assert np.allclose(A[:, 0] + A[:, 1:].sum(axis=1), c)
    assert _order_conditions(A, b, c, 5).max() < 1e-12          # 5th-order solution
    assert _order_conditions(A, bhat, c, 4).max() < 1e-12       # 4th-order embedded pair
    assert _order_conditions(A, bhat, c, 5).max() > 1e-6        # ... which is not 5th order
    assert abs(b[6]) == 0.0 and np.allclose(A[6, :6], b[:6])    # first same as last


def _instances(ND, n_inst, nmin, nmax, seed, m=2):
    rng = np.random.default_rng(seed)
    gs = [ND.barabasi_albert(int(n), m, seed=seed * 1000 + i) for i, n in enumerate(rng.integers(nmin, nmax + 1, n_inst))]
    return gs, rng


def _swing_params(ND, gs, rng):
    ps, us = [], []
    for g in gs:
        M, D, P, K = ND.workloads.swing_params(g.nv(), g.ne(), rng)
        p = np.empty(3 * g.nv() + g.ne())
        p[0:3 * g.nv():3], p[1:3 * g.nv():3], p[2:3 * g.nv():3], p[3 * g.nv():] = M, D, P, K
        u = np.empty(2 * g.nv())
        u[0::2], u[1::2] = rng.uniform(-0.5, 0.5, g.nv()), rng.normal(scale=0.5, size=g.nv())
        ps.append(p)
        us.append(u)
    return ps, us


@pytest.mark.gpu
def test_ensemble_rk4_matches_oracle(ND, dev):
    import torch
    gs, rng = _instances(ND, 24, 60, 2000, seed=3)
    ps, us = _swing_params(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u = torch.from_numpy(np.concatenate(us)).to(dev)
    p = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u, p, (0.0, 0.5), method=ND.ensemble.RK4, dt=0.005))
    torch.cuda.synchronize()
    got = u.cpu().numpy()
    assert (st["retcode"] == 0).all() and (st["naccept"] == 100).all()
    for i, g in enumerate(gs):
        nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        ref = oracle_for(nw).rk4(us[i], ps[i], 0.0, 0.005, 100)
        np.testing.assert_allclose(got[ens.state_slice(i)], ref, rtol=0, atol=1e-11, err_msg=f"instance {i}")


@pytest.mark.gpu
def test_ensemble_launch_shapes_agree(ND, dev):
    # 64..1024 threads per workgroup, 1-4 vertices per thread: RK4 sums every destination
    # list in the same chunks whatever the shape, so the results agree bit for bit; Tsit5's
    # error norm adds per-wave partials, so there the shapes agree to tolerance
    import torch
    gs, rng = _instances(ND, 10, 100, 1000, seed=12)
    ps, us = _swing_params(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u0 = torch.from_numpy(np.concatenate(us)).to(dev)
    p = torch.from_numpy(np.concatenate(ps)).to(dev)
    shapes = [[(512, 2)], [(1024, 1)], [(512, 4)], [(1024, 2)], [(256, 4)]]
    rk, ts = [], []
    for sh in shapes:
        ens.set_classes(sh)
        u = u0.clone()
        st = ND.Ensemble.stats_to_numpy(ens.solve(u, p, (0.0, 0.5), method=ND.ensemble.RK4, dt=0.005))
        assert (st["retcode"] == 0).all(), sh
        rk.append(u.cpu().numpy())
        u = u0.clone()
        st = ND.Ensemble.stats_to_numpy(ens.solve(u, p, (0.0, 1.0), reltol=1e-8, abstol=1e-8))
        assert (st["retcode"] == 0).all(), sh
        ts.append(u.cpu().numpy())
    for sh, r, t in zip(shapes[1:], rk[1:], ts[1:]):
        assert np.array_equal(r, rk[0]), sh
        np.testing.assert_allclose(t, ts[0], rtol=0, atol=1e-6, err_msg=str(sh))
    with pytest.raises(ND.ArgumentError):
        ens.set_class(0, 1024, 3)


@pytest.mark.gpu
def test_network_tsit5_paths(ND, dev):
    # Network.tsit5 (nd_tsit5 through the C ABI): a single swing network runs in one
    # workgroup and agrees with the host-stepped path to tolerance, and bit for bit with
    # the one-instance Ensemble; a network outside the kernel's models (load vertices next
    # to swing ones, static lines) gets ND_ERR_UNSUPPORTED and takes the host-stepped path
    import torch
    gs, rng = _instances(ND, 1, 600, 600, seed=21)
    ps, us = _swing_params(ND, gs, rng)
    nw = ND.Network(gs[0], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    p = torch.from_numpy(ps[0]).to(dev)
    u = torch.from_numpy(us[0]).to(dev)
    info = nw.tsit5(u, p, (0.0, 2.0))
    assert info["path"] == "one workgroup" and info["retcode"] == 0 and info["t"] == 2.0, info
    uh = torch.from_numpy(us[0]).to(dev)
    ref = ND.tsit5_per_instance(nw, uh, p, (0.0, 2.0))
    assert abs(ref["naccept"] - info["naccept"]) <= 2 + 0.02 * ref["naccept"]
    np.testing.assert_allclose(u.cpu().numpy(), uh.cpu().numpy(), rtol=0, atol=1e-6)
    ue = torch.from_numpy(us[0]).to(dev)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    st = ND.Ensemble.stats_to_numpy(ens.solve(ue, p, (0.0, 2.0), reltol=1e-8, abstol=1e-8))
    assert torch.equal(u, ue) and int(st["naccept"][0]) == info["naccept"]
    g = gs[0]
    vm = [ND.Lib.kuramoto_second() if i % 3 else ND.Lib.load_vertex() for i in range(g.nv())]
    nw2 = ND.Network(g, vm, ND.Lib.static_line_edge())
    p2 = rng.uniform(0.5, 2.0, nw2.pdim)
    u2 = rng.normal(size=nw2.dim) * 0.1
    a = torch.from_numpy(u2).to(dev)
    info2 = nw2.tsit5(a, torch.from_numpy(p2).to(dev), (0.0, 0.5))
    assert info2["path"] == "host steps" and info2["retcode"] == 0
    b = torch.from_numpy(u2).to(dev)
    ND.tsit5_per_instance(nw2, b, torch.from_numpy(p2).to(dev), (0.0, 0.5))
    assert torch.equal(a, b)
    # arguments are checked before either path is chosen
    for bad in (dict(tspan=(1.0, 0.0)), dict(tspan=(0.0, 1.0), reltol=0.0)):
        for net, uu, pp in ((nw, u, p), (nw2, a, torch.from_numpy(p2).to(dev))):
            with pytest.raises(ND.ArgumentError):
                net.tsit5(uu, pp, **bad)


@pytest.mark.gpu
def test_ensemble_tsit5_matches_per_instance(ND, dev):
    import torch
    gs, rng = _instances(ND, 12, 50, 600, seed=4)
    ps, us = _swing_params(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u = torch.from_numpy(np.concatenate(us)).to(dev)
    p = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u, p, (0.0, 2.0), reltol=1e-8, abstol=1e-8))
    torch.cuda.synchronize()
    got = u.cpu().numpy()
    assert (st["retcode"] == 0).all() and np.allclose(st["t"], 2.0)
    for i, g in enumerate(gs):
        nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        ui = torch.from_numpy(us[i]).to(dev)
        info = ND.tsit5_per_instance(nw, ui, torch.from_numpy(ps[i]).to(dev), (0.0, 2.0), reltol=1e-8, abstol=1e-8)
        # same method and controller: step counts agree closely, solutions to tolerance level
        assert abs(info["naccept"] - st["naccept"][i]) <= 2 + 0.02 * info["naccept"]
        np.testing.assert_allclose(got[ens.state_slice(i)], ui.cpu().numpy(), rtol=0, atol=1e-6)


def test_autotune_choice_logic(ND):
    from networkdynamics_jl_amd.ensemble_tune import choose
    # three bins; the small bins are cheapest on narrow shapes, the big bin fits only 512x4
    bin_max = [200, 900, 2000]
    weight = [10.0, 5.0, 1.0]
    cost = [{(128, 2): 1.0, (256, 1): 2.0, (256, 4): 3.0, (512, 4): 9.0},
            {(256, 4): 2.0, (512, 2): 1.5, (512, 4): 4.0},
            {(512, 4): 1.0}]
    shapes, pred = choose(bin_max, weight, cost)
    assert shapes == [(128, 2), (512, 2), (512, 4)]
    assert pred == 10.0 * 1.0 + 5.0 * 1.5 + 1.0
    # a single class must hold the largest bin
    shapes, _ = choose([2000], [1.0], [{(512, 4): 1.0}])
    assert shapes == [(512, 4)]
    # the dynamic programme agrees with brute force over every admissible list
    import itertools
    from networkdynamics_jl_amd.ensemble import NCLASS
    rng = np.random.default_rng(3)
    pool = [(64, 2), (128, 2), (256, 1), (128, 4), (256, 2), (512, 1), (256, 4), (512, 2), (512, 3), (512, 4)]
    for trial in range(20):
        bins = sorted(rng.choice(np.arange(50, 2049), size=4, replace=False).tolist())
        weight = rng.uniform(0.5, 3, 4).tolist()
        cost = [{sh: float(rng.uniform(0.5, 2)) for sh in pool if sh[0] * sh[1] >= m} for m in bins]
        shapes = sorted(pool, key=lambda x: (x[0] * x[1], x))
        best = float("inf")
        for k in range(1, NCLASS + 1):
            for combo in itertools.combinations(shapes, k):
                caps = [t * v for t, v in combo]
                if any(caps[i] >= caps[i + 1] for i in range(k - 1)) or caps[-1] < bins[-1]:
                    continue
                tot = sum(w * cost[b].get(next(x for x, cp in zip(combo, caps) if cp >= m), float("inf"))
                          for b, (m, w) in enumerate(zip(bins, weight)))
                best = min(best, tot)
        got, pred = choose(bins, weight, cost)
        assert abs(pred - best) <= 1e-12 * best


@pytest.mark.gpu
def test_ensemble_graph_replays_match_eager(ND, dev):
    import torch
    gs, rng = _instances(ND, 40, 30, 1500, seed=5)
    ps, us = _swing_params(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u0 = torch.from_numpy(np.concatenate(us)).to(dev)
    p = torch.from_numpy(np.concatenate(ps)).to(dev)
    ue = u0.clone()
    ste = ND.Ensemble.stats_to_numpy(ens.solve(ue, p, (0.0, 1.0), reltol=1e-8, abstol=1e-8))
    torch.cuda.synchronize()
    ug = u0.clone()
    graph = ens.capture(ug, p, (0.0, 1.0), reltol=1e-8, abstol=1e-8)
    for rep in range(3):
        ug.copy_(u0)
        junk = torch.randn(1000, device=dev).sum()          # eager work between replays
        stg = ND.Ensemble.stats_to_numpy(graph.launch())
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        assert torch.equal(ug, ue), f"replay {rep}"
        assert (stg == ste).all(), f"replay {rep}"
    ens.set_class(0, 64, 2)
    with pytest.raises(ND.ArgumentError):
        graph.launch()


@pytest.mark.gpu
def test_ensemble_autotune_and_cache(ND, dev, tmp_path):
    import torch
    from networkdynamics_jl_amd.ensemble_tune import autotune
    gs, rng = _instances(ND, 48, 40, 1200, seed=6)
    ps, us = _swing_params(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u0 = torch.from_numpy(np.concatenate(us)).to(dev)
    p = torch.from_numpy(np.concatenate(ps)).to(dev)
    cache = str(tmp_path / "tune.json")
    res = autotune(ens, u0, p, probe_t=0.2, nbins=3, fill=256, cache=cache)
    assert not res["cached"]
    caps = [t * v for t, v in res["classes"]]
    assert caps == sorted(caps) and caps[-1] >= max(g.nv() for g in gs)
    assert [tuple(c) for c in ens.classes[:len(res["classes"])]] == [tuple(c) for c in res["classes"]]
    again = autotune(ens, u0, p, probe_t=0.2, nbins=3, fill=256, cache=cache)
    assert again["cached"] and again["classes"] == res["classes"]
    # tuned launch shapes integrate to the same answer as the per-instance path
    u = u0.clone()
    st = ND.Ensemble.stats_to_numpy(ens.solve(u, p, (0.0, 1.0), reltol=1e-8, abstol=1e-8))
    torch.cuda.synchronize()
    assert (st["retcode"] == 0).all()
    for i in (0, 17, 47):
        nw = ND.Network(gs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        ui = torch.from_numpy(us[i]).to(dev)
        ND.tsit5_per_instance(nw, ui, torch.from_numpy(ps[i]).to(dev), (0.0, 1.0), reltol=1e-8, abstol=1e-8)
        np.testing.assert_allclose(u[ens.state_slice(i)].cpu().numpy(), ui.cpu().numpy(), rtol=0, atol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["swing_static_line", "kuramoto_first"])
def test_ensemble_rk4_other_models_match_oracle(ND, dev, kind):
    # the per-edge stage values are shared by both endpoints; check two more model pairs,
    # including the two-parameter static line (B, phi), against the oracle's RK4
    import torch
    gs, rng = _instances(ND, 10, 30, 1500, seed=7)
    if kind == "swing_static_line":
        vm, em = ND.Lib.kuramoto_second(), ND.Lib.static_line_edge()
    else:
        vm, em = ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge()
    us, ps = [], []
    for g in gs:
        n, m = g.nv(), g.ne()
        if kind == "swing_static_line":
            pv = np.stack([rng.uniform(1, 5, n), rng.uniform(0.1, 1, n), rng.normal(size=n)], 1).ravel()
            pe = np.stack([rng.uniform(1, 20, m), rng.uniform(0, 0.1, m)], 1).ravel()
            u = np.empty(2 * n)
            u[0::2], u[1::2] = rng.uniform(-1, 1, n), rng.normal(scale=0.5, size=n)
        else:
            pv, pe = rng.normal(size=n), np.full(m, 3.0)
            u = rng.uniform(-np.pi, np.pi, n)
        us.append(u)
        ps.append(np.r_[pv, pe])
    ens = ND.Ensemble(gs, vm, em)
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 0.2), method=ND.ensemble.RK4, dt=0.002))
    torch.cuda.synchronize()
    assert (st["retcode"] == 0).all() and (st["naccept"] == 100).all()
    got = u_d.cpu().numpy()
    for i, g in enumerate(gs):
        ref = oracle_for(ND.Network(g, vm, em)).rk4(us[i], ps[i], 0.0, 0.002, 100)
        np.testing.assert_allclose(got[ens.state_slice(i)], ref, rtol=0, atol=1e-11, err_msg=f"instance {i}")


@pytest.mark.gpu
def test_ensemble_line_tripping_matches_oracle_stepping(ND, dev):
    # cascading-failure callback (docs/examples/cascading_failure.jl): after every RK4 step
    # a line whose |K sin(theta_s - theta_d)| exceeds its limit gets K = 0.  Reference:
    # the oracle's RK4 one step at a time with the same check on the host.
    import torch
    gs, rng = _instances(ND, 8, 30, 400, seed=8)
    ps, us = _swing_params(ND, gs, rng)
    for u in us:
        u[0::2] = rng.uniform(-1.5, 1.5, len(u) // 2)       # far from synchrony: lines overload
    lims = [rng.uniform(0.85, 1.0, g.ne()) * p[3 * g.nv():] for g, p in zip(gs, ps)]
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    lim_d = torch.from_numpy(np.concatenate(lims)).to(dev)
    dt, nsteps = 0.005, 60
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, dt * nsteps), method=ND.ensemble.RK4, dt=dt,
                                              line_limit=lim_d))
    torch.cuda.synchronize()
    got_u, got_p = u_d.cpu().numpy(), p_d.cpu().numpy()
    assert st["ntrip"].sum() > 0
    for i, g in enumerate(gs):
        nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        on = oracle_for(nw)
        u, p = us[i].copy(), ps[i].copy()
        n = g.nv()
        s, d = g.src - 1, g.dst - 1
        trips = 0
        for _ in range(nsteps):
            u = on.rk4(u, p, 0.0, dt, 1)
            K = p[3 * n:]
            flow = K * np.sin(u[2 * s] - u[2 * d])
            hit = (K != 0) & (np.abs(flow) > lims[i])
            K[hit] = 0.0
            trips += int(hit.sum())
        assert st["ntrip"][i] == trips, i
        np.testing.assert_array_equal(got_p[ens.param_slice(i)], p)
        np.testing.assert_allclose(got_u[ens.state_slice(i)], u, rtol=0, atol=1e-10, err_msg=f"instance {i}")


@pytest.mark.gpu
def test_ensemble_tsit5_line_tripping_consistent(ND, dev):
    # adaptive Tsit5 with tripping: every tripped line had K > 0 before and ends at K = 0,
    # no other parameter changes, and the counts add up
    import torch
    gs, rng = _instances(ND, 16, 50, 800, seed=9)
    ps, us = _swing_params(ND, gs, rng)
    for u in us:
        u[0::2] = rng.uniform(-1.5, 1.5, len(u) // 2)
    lims = np.concatenate([rng.uniform(0.85, 1.0, g.ne()) * p[3 * g.nv():] for g, p in zip(gs, ps)])
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    p0 = np.concatenate(ps)
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(p0.copy()).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 2.0), line_limit=torch.from_numpy(lims).to(dev)))
    torch.cuda.synchronize()
    p1 = p_d.cpu().numpy()
    assert (st["retcode"] == 0).all() and st["ntrip"].sum() > 0
    changed = p1 != p0
    assert (p1[changed] == 0).all()
    for i, g in enumerate(gs):
        sl = ens.param_slice(i)
        assert changed[sl][:3 * g.nv()].sum() == 0
        assert changed[sl].sum() == st["ntrip"][i]


def _damped_near_sync(ND, gs, rng):
    # M = 1, D = 2, small balanced injections, start at theta = 0: every instance settles
    # to its synchronous state within a few time units
    ps, us = [], []
    for g in gs:
        n = g.nv()
        P = rng.normal(scale=0.1, size=n)
        P -= P.mean()
        p = np.empty(3 * n + g.ne())
        p[0:3 * n:3], p[1:3 * n:3], p[2:3 * n:3] = 1.0, 2.0, P
        p[3 * n:] = rng.uniform(5.0, 10.0, g.ne())
        u = np.zeros(2 * n)
        u[1::2] = rng.normal(scale=0.05, size=n)
        ps.append(p)
        us.append(u)
    return ps, us


def _ss_ratio(du, u, abstol, reltol):
    """max_i of how far |du_i| is from passing (<= 1 passes): the allDerivPass test."""
    a = np.abs(du)
    return float(np.max(np.minimum(a / abstol, np.where(reltol * np.abs(u) > 0, a / (reltol * np.abs(u) + 1e-300),
                                                           np.inf))))


@pytest.mark.gpu
def test_ensemble_rk4_steady_state_termination_matches_oracle(ND, dev):
    # TerminateSteadyState semantics: after every RK4 step, stop at the first state where
    # every |du_i| <= abstol or <= reltol*|u_i|.  Reference: the oracle stepped one RK4
    # step at a time with the same test on the host.
    import torch
    gs, rng = _instances(ND, 12, 40, 400, seed=21)
    ps, us = _damped_near_sync(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    dt, nsteps, tols = 0.01, 4000, (1e-6, 1e-9)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, dt * nsteps), method=ND.ensemble.RK4, dt=dt,
                                              steady_state=tols))
    torch.cuda.synchronize()
    got = u_d.cpu().numpy()
    assert (st["retcode"] == ND.ensemble.TERMINATED).all(), st["retcode"]
    checked = 0
    for i, g in enumerate(gs):
        on = oracle_for(ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge()))
        u = us[i].copy()
        for k in range(1, nsteps + 1):
            u = on.rk4(u, ps[i], 0.0, dt, 1)
            r = _ss_ratio(on.rhs(u, ps[i]), u, *tols)
            if r <= 1.0:
                break
        steps = int(round(st["t"][i] / dt))
        assert st["naccept"][i] == steps
        if abs(r - 1.0) > 1e-6:                     # a decision this close to the boundary may round either way
            assert steps == k, (i, steps, k)
            np.testing.assert_allclose(got[ens.state_slice(i)], u, rtol=0, atol=1e-10)
            checked += 1
    assert checked >= 10


@pytest.mark.gpu
def test_ensemble_tsit5_steady_state_termination(ND, dev):
    # adaptive steps: every instance stops before t1 with retcode TERMINATED at a state that
    # passes the test by the oracle's right-hand side; off, the same instances run to t1.
    # Near the equilibrium the step size settles at the stability limit of the fastest
    # modes, where their residual oscillation is held at the integration tolerance: the
    # steady-state tolerance sits well above it.
    import torch
    gs, rng = _instances(ND, 16, 50, 800, seed=22)
    ps, us = _damped_near_sync(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u0 = np.concatenate(us)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    tols = (1e-5, 1e-9)
    u_d = torch.from_numpy(u0).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 200.0), steady_state=tols))
    u_full = torch.from_numpy(u0).to(dev)
    st_full = ND.Ensemble.stats_to_numpy(ens.solve(u_full, p_d, (0.0, 200.0)))
    torch.cuda.synchronize()
    got, full = u_d.cpu().numpy(), u_full.cpu().numpy()
    final = [_ss_ratio(oracle_for(ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())).rhs(
        full[ens.state_slice(i)], ps[i]), full[ens.state_slice(i)], *tols) for i, g in enumerate(gs)]
    assert (st["retcode"] == ND.ensemble.TERMINATED).all() and (st["t"] < 200.0).all(), (st["retcode"], final)
    assert (st_full["retcode"] == 0).all() and (st_full["t"] == 200.0).all()
    assert (st["nf"] < st_full["nf"]).all()
    for i, g in enumerate(gs):
        on = oracle_for(ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge()))
        u = got[ens.state_slice(i)]
        assert _ss_ratio(on.rhs(u, ps[i]), u, *tols) <= 1.0 + 1e-9
        # both runs sit at the same synchronous state (angles up to the common drift)
        th, thf = u[0::2], full[ens.state_slice(i)][0::2]
        np.testing.assert_allclose(th - th.mean(), thf - thf.mean(), atol=1e-4)
    with pytest.raises(ND.ArgumentError):
        ens.solve(u_d, p_d, (0.0, 1.0), steady_state=(-1.0, 0.0))


@pytest.mark.gpu
def test_ensemble_rk4_preset_trip_matches_oracle_stepping(ND, dev):
    # PresetTimeComponentCallback (docs/examples/cascading_failure.jl:125-131): the line
    # given per instance loses its K at t = 0.1, the end of step 20 of dt = 0.005.
    # Reference: the oracle's RK4 one step at a time, K set to 0 after step 20.
    import torch
    gs, rng = _instances(ND, 6, 30, 300, seed=12)
    ps, us = _swing_params(ND, gs, rng)
    edges = np.array([int(rng.integers(0, g.ne())) for g in gs], np.int32)
    edges[3] = -1                                            # no event for this instance
    times = np.full(len(gs), 0.1)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    dt, nsteps = 0.005, 60
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, dt * nsteps), method=ND.ensemble.RK4, dt=dt,
                                              preset_trip=(times, edges)))
    torch.cuda.synchronize()
    got_u, got_p = u_d.cpu().numpy(), p_d.cpu().numpy()
    for i, g in enumerate(gs):
        nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        on = oracle_for(nw)
        u, p = us[i].copy(), ps[i].copy()
        for k in range(nsteps):
            u = on.rk4(u, p, 0.0, dt, 1)
            if k + 1 == 20 and edges[i] >= 0:
                p[3 * g.nv() + edges[i]] = 0.0
        assert st["ntrip"][i] == (1 if edges[i] >= 0 else 0)
        np.testing.assert_array_equal(got_p[ens.param_slice(i)], p)
        np.testing.assert_allclose(got_u[ens.state_slice(i)], u, rtol=0, atol=1e-10, err_msg=f"instance {i}")
    with pytest.raises(ND.ArgumentError, match="step grid"):
        ens.solve(u_d, p_d, (0.0, dt * nsteps), method=ND.ensemble.RK4, dt=dt,
                  preset_trip=(np.full(len(gs), 0.1013), edges))
    with pytest.raises(ND.ArgumentError, match="one time and one edge"):
        ens.solve(u_d, p_d, (0.0, 1.0), preset_trip=(times[:2], edges[:2]))


@pytest.mark.gpu
def test_ensemble_tsit5_preset_trip_is_a_tstop(ND, dev):
    # Tsit5 lands on the trip time and restarts there: the same as two solves split at
    # t* with K set to 0 in between (to the integration tolerance), and clearly unlike a
    # trip at a slightly later time
    import torch
    gs, rng = _instances(ND, 8, 50, 400, seed=13)
    ps, us = _swing_params(ND, gs, rng)
    edges = np.array([int(rng.integers(0, g.ne())) for g in gs], np.int32)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    tol = dict(reltol=1e-11, abstol=1e-11)

    def run(t_event):
        u_d = torch.from_numpy(np.concatenate(us)).to(dev)
        p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
        st = ens.solve(u_d, p_d, (0.0, 2.0), preset_trip=(np.full(len(gs), t_event), edges), **tol)
        torch.cuda.synchronize()
        return u_d.cpu().numpy(), p_d.cpu().numpy(), ND.Ensemble.stats_to_numpy(st)

    u1, p1, st = run(0.7)
    assert (st["retcode"] == 0).all() and (st["ntrip"] == 1).all() and np.all(st["t"] == 2.0)
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    ens.solve(u_d, p_d, (0.0, 0.7), **tol)
    p_h = p_d.cpu().numpy()
    for i, g in enumerate(gs):
        p_h[ens.param_slice(i).start + 3 * g.nv() + edges[i]] = 0.0
    np.testing.assert_array_equal(p_h, p1)
    p_d = torch.from_numpy(p_h).to(dev)
    ens.solve(u_d, p_d, (0.7, 2.0), **tol)
    u2 = u_d.cpu().numpy()
    np.testing.assert_allclose(u1, u2, rtol=0, atol=1e-7)
    u3, _, _ = run(0.71)
    assert np.abs(u3 - u2).max() > 100 * np.abs(u1 - u2).max()
