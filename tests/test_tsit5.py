"""Tsit5: the oracle's restatement pinned by the order conditions of its coefficients,
fifth-order convergence and exact solutions (CPU); the device steppers -- the ensemble
kernel at the basin-stability workload's start states and size, and Network.tsit5's
host-stepped path on a time-dependent network -- against it (GPU).

The reference solves with OrdinaryDiffEqTsit5's Tsit5() (docs/examples/cascading_failure.jl:145,
test/Project.toml:36); its step sequence cannot be run here, so the oracle restates the
published method and the tests below pin that restatement."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import oracle_for


def _order_residuals(A, b, c, order):
    """Butcher order conditions (all rooted trees up to `order` <= 5) as residuals."""
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


def _check_tableau(A, bt, c):
    b = A[6].copy()
    bhat = b - bt
    assert np.array_equal(np.triu(A), np.zeros_like(A))              # explicit
    assert np.allclose(A.sum(axis=1), c, atol=1e-15, rtol=0)        # c = A 1
    assert _order_residuals(A, b, c, 5).max() < 1e-12               # 5th-order solution
    assert _order_residuals(A, bhat, c, 4).max() < 1e-12            # 4th-order embedded
    assert _order_residuals(A, bhat, c, 5).max() > 1e-6             # ... which is not 5th order
    assert b[6] == 0.0                                               # first same as last


def test_device_and_oracle_tableaus(ND, oracle):
    """The kernel's own coefficients (nd_tsit5_tableau, the table the kernel reads) and the
    oracle's separately typed copy both satisfy the order conditions, and are equal."""
    Ad, btd, cd = ND.ensemble.tsit5_tableau()
    Ao, bto, co = oracle.tsit5_tableau()
    _check_tableau(Ad, btd, cd)
    _check_tableau(Ao, bto, co)
    assert np.array_equal(Ad, Ao) and np.array_equal(btd, bto) and np.array_equal(cd, co)
    # a mistyped coefficient breaks an order condition
    Ab = Ao.copy()
    Ab[4, 2] *= 1 + 1e-9
    assert _order_residuals(Ab, Ab[6], Ab.sum(axis=1), 5).max() > 1e-10


def _capacitor_pair(ND, C=1.0, R=1.0):
    """Two capacitor loads joined by a resistor (docs/src/external_inputs.md:62-79):
    V' = (i - (1 + 0.1 sin t)) / C.  The currents cancel in the sum S = V1 + V2, so
    S(t) = S0 - (2/C)(t + 0.1 (1 - cos t)); the difference D decays as D0 exp(-2t/(RC))."""
    g = ND.SimpleGraph(2, [1], [2])
    nw = ND.Network(g, ND.Lib.capacitor_load(), ND.Lib.resistor())
    p = np.array([C, C, R])
    return nw, p


def _capacitor_exact(u0, t, C, R):
    S0, D0 = u0[0] + u0[1], u0[0] - u0[1]
    S = S0 - 2.0 / C * (t + 0.1 * (1 - np.cos(t)))
    D = D0 * np.exp(-2 * t / (R * C))
    return np.array([(S + D) / 2, (S - D) / 2])


def _fixed_step(on, u0, p, t1, dt):
    """The oracle's Tsit5 at a fixed step: every step accepted (huge tolerances), dt
    capped at dtmax = dt."""
    u, _, st = on.tsit5(u0, p, (0.0, t1), reltol=1e30, abstol=1e30, dt0=dt, dtmax=dt)
    assert st["retcode"] == 0 and st["nreject"] == 0
    return u, st


def test_oracle_tsit5_fifth_order_time_dependent(ND):
    """Fixed steps on the time-dependent capacitor pair: the error falls 2^5-fold per
    halving.  Stages evaluated at the wrong times (t + h instead of t + c_i h) would
    integrate the sin(t) forcing to first order only."""
    nw, p = _capacitor_pair(ND, C=1.0, R=1.0)
    on = oracle_for(nw)
    u0 = np.array([0.7, -0.2])
    errs = []
    for dt in (0.2, 0.1, 0.05):
        u, st = _fixed_step(on, u0, p, 2.0, dt)
        assert st["naccept"] == round(2.0 / dt)
        errs.append(np.abs(u - _capacitor_exact(u0, 2.0, 1.0, 1.0)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates.min() > 4.6, (errs, rates)


def test_oracle_tsit5_fifth_order_on_diffusion(ND):
    """Fixed steps on a diffusion network du = -L x (test/diffusion_test.jl:80-90), against
    the exact exp(-t L) x0."""
    import scipy.linalg
    g = ND.watts_strogatz(12, 4, 0.3, seed=3)
    nw = ND.Network(g, ND.Lib.diffusion_vertex(), ND.Lib.diffusion_edge())
    on = oracle_for(nw)
    n = g.nv()
    L = np.zeros((n, n))
    for s, d in zip(g.src - 1, g.dst - 1):
        L[s, s] += 1
        L[d, d] += 1
        L[s, d] -= 1
        L[d, s] -= 1
    x0 = np.random.default_rng(4).normal(size=n)
    exact = scipy.linalg.expm(-1.0 * L) @ x0
    w = np.ones(nw.pdim)                    # unit edge weights: du = -L x
    errs = [np.abs(_fixed_step(on, x0, w, 1.0, dt)[0] - exact).max() for dt in (0.1, 0.05, 0.025)]
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert rates.min() > 4.6, (errs, rates)


def test_oracle_tsit5_adaptive_error_follows_tolerance(ND):
    """Adaptive runs on the capacitor pair: the global error stays within a small
    multiple of the tolerance and shrinks with it, and tighter tolerances take more steps
    at about the fifth-root rate; the solve ends exactly at t1."""
    nw, p = _capacitor_pair(ND, C=1.0, R=0.5)
    on = oracle_for(nw)
    u0 = np.array([1.5, -0.5])
    exact = _capacitor_exact(u0, 10.0, 1.0, 0.5)
    steps = []
    for tol in (1e-6, 1e-9, 1e-12):
        u, _, st = on.tsit5(u0, p, (0.0, 10.0), reltol=tol, abstol=tol)
        assert st["retcode"] == 0 and st["t"] == 10.0
        err = np.abs(u - exact).max()
        assert err < 50 * tol * max(1.0, np.abs(exact).max()), (tol, err)
        steps.append(st["naccept"])
    r1, r2 = steps[1] / steps[0], steps[2] / steps[1]
    # 1000x tighter tolerance: about 10**(3/5) = 4x the steps
    assert 2.5 < r1 < 6.5 and 2.5 < r2 < 6.5, steps


def test_oracle_tsit5_retcodes_and_empty_span(ND):
    nw, p = _capacitor_pair(ND)
    on = oracle_for(nw)
    u0 = np.array([1.0, 0.0])
    u, _, st = on.tsit5(u0, p, (0.0, 5.0), maxiters=3)
    assert st["retcode"] == 1 and st["naccept"] + st["nreject"] == 3
    u, _, st = on.tsit5(u0, p, (2.0, 2.0))
    assert st["retcode"] == 0 and st["naccept"] == 0 and np.array_equal(u, u0)
    u, _, st = on.tsit5(np.array([np.nan, 0.0]), p, (0.0, 1.0))
    assert st["retcode"] in (2, 3)


def test_oracle_tsit5_preset_trip_is_a_tstop(ND):
    """The preset trip lands a step on its time and zeroes the edge parameter there.  Two
    first-order Kuramoto oscillators on one line: after K goes to 0 they run free, so the
    end state is the state at the trip time plus omega * (t1 - t_trip)."""
    g = ND.SimpleGraph(2, [1], [2])
    nw = ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
    on = oracle_for(nw)
    p = np.array([0.3, -0.4, 2.0])          # omega1, omega2, K
    u0 = np.array([0.1, 1.2])
    u, pt, st = on.tsit5(u0, p, (0.0, 3.0), reltol=1e-10, abstol=1e-10, preset_trip=(1.25, 0))
    assert st["ntrip"] == 1 and pt[2] == 0.0 and st["retcode"] == 0
    ua, _, sa = on.tsit5(u0, p, (0.0, 1.25), reltol=1e-10, abstol=1e-10)
    assert sa["t"] == 1.25
    # after the trip the oscillators run free at their natural frequencies
    expect = ua + np.array([0.3, -0.4]) * (3.0 - 1.25)
    assert np.allclose(u, expect, atol=1e-9, rtol=0)


# ---- device against the oracle -------------------------------------------------------

def _scaled_state_error(u, ref):
    return np.abs(u - ref) / np.maximum(1.0, np.abs(ref))


@pytest.mark.gpu
def test_ensemble_tsit5_matches_oracle_at_c4_starts(ND, dev):
    """The basin-stability workload's start states (uneven BA swing networks, N in
    100..2000, batched-Newton fixpoints plus one perturbed node), Tsit5 at reltol 1e-8 to
    t = 5 in the ensemble kernel and in the oracle: accepted steps within 2, states within
    1e-8 scaled."""
    import torch
    from networkdynamics_jl_amd import experiments as X
    graphs, us, ps, _, _, _ = X.c4_start(dev, ninst=12, seed=11)
    assert len(graphs) >= 10
    ens = ND.ensemble.Ensemble(graphs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ens.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 5.0), reltol=1e-8, abstol=1e-8))
    uf = u_d.cpu().numpy()

    def run(i):
        nw = ND.Network(graphs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        return oracle_for(nw).tsit5(us[i], ps[i], (0.0, 5.0), reltol=1e-8, abstol=1e-8)

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(run, range(len(graphs))))
    for i, (ur, _, sr) in enumerate(refs):
        assert st["retcode"][i] == 0 and sr["retcode"] == 0
        assert abs(int(st["naccept"][i]) - sr["naccept"]) <= 2, (i, st["naccept"][i], sr["naccept"])
        err = _scaled_state_error(uf[ens.state_slice(i)], ur).max()
        assert err <= 1e-8, (i, graphs[i].nv(), err)


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["crossing", "step_end"])
def test_ensemble_tsit5_trips_match_oracle(ND, dev, mode):
    """Line tripping (crossings located in the step, or limits checked at step ends) and
    a preset trip in the ensemble kernel against the oracle's Tsit5 with the same
    callbacks: the same lines trip (parameters bitwise equal), the retcodes agree.  States
    agree within 1e-5 scaled: a step-end trip near its threshold may fire one step earlier
    or later in one of the two, and step sequences that differ by a step already differ
    by O(reltol)."""
    import torch
    from networkdynamics_jl_amd import experiments as X
    graphs, us, ps, _, _, rng = X.c4_start(dev, ninst=8, seed=23, nmin=60, nmax=400)
    ens = ND.ensemble.Ensemble(graphs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    n_inst = len(graphs)
    # line limits at 1.4x each line's flow at the start plus 0.5: the kicked node's lines trip
    lims = []
    for g, u, p in zip(graphs, us, ps):
        n = g.nv()
        th = u[0::2]
        flow = p[3 * n:] * np.sin(th[g.src - 1] - th[g.dst - 1])
        lims.append(1.4 * np.abs(flow) + 0.5)
    lim_h = np.concatenate(lims)
    times = np.full(n_inst, 0.7)
    edges = np.array([int(rng.integers(g.ne())) for g in graphs], np.int32)
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ens.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 20.0), reltol=1e-8, abstol=1e-8,
                                      line_limit=torch.from_numpy(lim_h).to(dev), preset_trip=(times, edges),
                                      line_events=mode))
    uf, pf = u_d.cpu().numpy(), p_d.cpu().numpy()

    def run(i):
        nw = ND.Network(graphs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        return oracle_for(nw).tsit5(us[i], ps[i], (0.0, 20.0), reltol=1e-8, abstol=1e-8, line_limit=lims[i],
                                    preset_trip=(0.7, int(edges[i])), line_events=mode)

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(run, range(n_inst)))
    trips = 0
    for i, (ur, pr, sr) in enumerate(refs):
        assert st["retcode"][i] == sr["retcode"] == 0, (i, st["retcode"][i], sr["retcode"])
        assert st["ntrip"][i] == sr["ntrip"], (i, st["ntrip"][i], sr["ntrip"])
        assert np.array_equal(pf[ens.param_slice(i)], pr)
        assert abs(int(st["naccept"][i]) - sr["naccept"]) <= 2
        assert _scaled_state_error(uf[ens.state_slice(i)], ur).max() <= 1e-5
        trips += sr["ntrip"]
    assert trips > n_inst                      # more than the preset trips fired


@pytest.mark.gpu
def test_ensemble_tsit5_steady_state_matches_oracle(ND, dev):
    """Steady-state termination (TerminateSteadyState's test) in the ensemble kernel and
    in the oracle's Tsit5 on damped networks near sync: both stop with retcode 4, at the
    same accepted step within 2 and the same time within a step."""
    import torch
    from test_integrators import _damped_near_sync, _instances
    gs, rng = _instances(ND, 8, 50, 600, seed=27)
    ps, us = _damped_near_sync(ND, gs, rng)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    tols = (1e-5, 1e-9)
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 200.0), steady_state=tols))

    def run(i):
        nw = ND.Network(gs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        return oracle_for(nw).tsit5(us[i], ps[i], (0.0, 200.0), steady_state=tols)

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(run, range(len(gs))))
    for i, (_, _, sr) in enumerate(refs):
        assert st["retcode"][i] == sr["retcode"] == 4, (i, st["retcode"][i], sr["retcode"])
        assert abs(int(st["naccept"][i]) - sr["naccept"]) <= 2, (i, st["naccept"][i], sr["naccept"])
        assert abs(st["t"][i] - sr["t"]) <= 2 * max(st["dt"][i], sr["dt"]), (i, st["t"][i], sr["t"])


@pytest.mark.gpu
def test_ensemble_c4_full_size(ND, dev):
    """The basin-stability workload at its full size: 4096 uneven BA swing networks from
    their fixpoints plus a one-node kick, Tsit5 to t = 50 at reltol 1e-8.  Every instance
    reaches t1, the captured hipGraph's replay equals the eager solve bit for bit, instances
    solved in a smaller ensemble of their own end in the same bits, and on 16 sampled
    instances the oracle's Tsit5 gives the same return-to-sync class, accepted steps
    within 2 and states within 1e-6 scaled."""
    import torch
    from networkdynamics_jl_amd import experiments as X
    graphs, us, ps, dropped, _, rng = X.c4_start(dev, ninst=4096, seed=1)
    assert len(graphs) + dropped == 4096 and len(graphs) >= 4000
    ens = ND.ensemble.Ensemble(graphs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    u_g = u_d.clone()
    st = ens.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 50.0), reltol=1e-8, abstol=1e-8))
    assert (st["retcode"] == 0).all() and np.all(st["t"] == 50.0)
    graph = ens.capture(u_g, p_d, (0.0, 50.0), reltol=1e-8, abstol=1e-8)
    sg = ens.stats_to_numpy(graph.launch())
    torch.cuda.synchronize()
    assert torch.equal(u_g, u_d)
    assert np.array_equal(sg["naccept"], st["naccept"])
    del graph
    uf = u_d.cpu().numpy()
    # instances are independent: 64 of them from the smallest and the largest size
    # class, solved in an ensemble of their own, end in the same bits
    nv = np.array([g.nv() for g in graphs])
    sub = list(np.flatnonzero(nv <= 256)[:32]) + list(np.flatnonzero(nv > 1536)[:32])
    ens_s = ND.ensemble.Ensemble([graphs[i] for i in sub], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_s = torch.from_numpy(np.concatenate([us[i] for i in sub])).to(dev)
    p_s = torch.from_numpy(np.concatenate([ps[i] for i in sub])).to(dev)
    ens_s.solve(u_s, p_s, (0.0, 50.0), reltol=1e-8, abstol=1e-8)
    us_f = u_s.cpu().numpy()
    for k, i in enumerate(sub):
        assert np.array_equal(us_f[ens_s.state_slice(k)], uf[ens.state_slice(i)]), i
    sync = np.array([X.returned_to_sync(uf[ens.state_slice(i)]) for i in range(len(graphs))])
    assert 0.5 < sync.mean() < 1.0              # both classes occur
    # sample both classes, small and large instances
    pick = list(rng.choice(np.flatnonzero(sync), 12, replace=False)) + \
        list(rng.choice(np.flatnonzero(~sync), 4, replace=False))

    def run(i):
        nw = ND.Network(graphs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        return oracle_for(nw).tsit5(us[i], ps[i], (0.0, 50.0), reltol=1e-8, abstol=1e-8)

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(run, pick))
    for i, (ur, _, sr) in zip(pick, refs):
        assert sr["retcode"] == 0
        assert X.returned_to_sync(ur) == sync[i], i
        assert abs(int(st["naccept"][i]) - sr["naccept"]) <= 2, (i, st["naccept"][i], sr["naccept"])
        err = _scaled_state_error(uf[ens.state_slice(i)], ur)
        # the rotating (unsynchronised) nodes carry their phase error along: compare
        # frequencies and phase differences to the first vertex for them
        if sync[i]:
            assert err.max() <= 1e-6, (i, err.max())
        else:
            w, wr = uf[ens.state_slice(i)][1::2], ur[1::2]
            assert np.abs(w - wr).max() <= 1e-6 * max(1.0, np.abs(wr).max()), i


@pytest.mark.gpu
def test_network_tsit5_host_path_time_dependent(ND, dev):
    """Network.tsit5 on the time-dependent capacitor pair takes the host-stepped path
    (its models are outside the one-workgroup kernel): it matches the closed form and the
    oracle's Tsit5 (same accepted steps within 2); a NaN start stops with retcode 3;
    an empty span returns at once."""
    import torch
    nw, p = _capacitor_pair(ND, C=1.0, R=0.5)
    on = oracle_for(nw)
    u0 = np.array([1.5, -0.5])
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    info = nw.tsit5(u, pd, (0.0, 10.0), reltol=1e-9, abstol=1e-9)
    assert info["path"] == "host steps" and info["retcode"] == 0 and info["t"] == 10.0
    ur, _, sr = on.tsit5(u0, p, (0.0, 10.0), reltol=1e-9, abstol=1e-9)
    assert abs(info["naccept"] - sr["naccept"]) <= 2, (info, sr)
    exact = _capacitor_exact(u0, 10.0, 1.0, 0.5)
    assert np.abs(u.cpu().numpy() - exact).max() < 1e-7
    assert np.abs(u.cpu().numpy() - ur).max() < 1e-10
    un = torch.tensor([np.nan, 0.0], dtype=torch.float64, device=dev)
    assert nw.tsit5(un, pd, (0.0, 1.0))["retcode"] in (2, 3)
    ue = torch.from_numpy(u0).to(dev)
    info = nw.tsit5(ue, pd, (3.0, 3.0))
    assert info["retcode"] == 0 and info["naccept"] == 0 and torch.equal(ue, torch.from_numpy(u0).to(dev))
