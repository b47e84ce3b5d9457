"""Line trips located at the crossing of |P| - limit inside a Tsit5 step.

The reference trips lines with a ContinuousComponentCallback (src/callbacks.jl:113-140;
docs/examples/cascading_failure.jl:104-111): the condition |P| - limit, the affect K = 0
on up- and down-crossings, run by DiffEqBase's VectorContinuousCallback on Tsit5's dense
output.  Pinned here:
  * the dense-output table (Tsitouras' continuous extension) by the continuous order
    conditions, on the oracle's and the device library's copies (CPU);
  * the oracle's event location by closed forms of the Adler equation of two
    Kuramoto oscillators on one line, delta' = domega - 2K sin(delta), delta = theta_1 -
    theta_2, whose line flow is P = K sin(delta): the crossing time, the state frozen
    (or drifting at the natural frequencies) after the trip, up- and down-crossings,
    and a line above its limit at t0 that waits for its crossing (CPU);
  * the ensemble kernel against those closed forms and against the oracle on cascades
    in swing networks (GPU).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from conftest import oracle_for


def _continuous_order_residuals(A, c, r, s):
    """Residuals of the order conditions of the dense output b(s) up to order 4:
    sum_i b_i(s) Phi_i(tree) = s^rho / gamma(tree)."""
    bs = r @ np.array([s, s ** 2, s ** 3, s ** 4])
    Ac = A @ c
    conds = [(np.ones(7), s, 1), (c, s ** 2, 2), (c ** 2, s ** 3, 3), (Ac, s ** 3, 6), (c ** 3, s ** 4, 4),
             (c * Ac, s ** 4, 8), (A @ c ** 2, s ** 4, 12), (A @ Ac, s ** 4, 24)]
    return np.array([bs @ phi - num / den for phi, num, den in conds])


def _check_dense(A, c, r):
    b = A[6]
    assert np.array_equal(r[1:, 0], np.zeros(6))                # only stage 1 has a linear term
    np.testing.assert_allclose(r.sum(axis=1), b, rtol=0, atol=2e-14)   # b_i(1) = b_i
    worst = max(np.abs(_continuous_order_residuals(A, c, r, s)).max() for s in np.linspace(0.0, 1.0, 11))
    assert worst < 1e-13, worst
    # not of order 5 inside the step (sum b_i(s) c_i^4 != s^5/5 at s = 1/2)
    assert abs(r @ np.array([0.5, 0.25, 0.125, 0.0625]) @ c ** 4 - 0.5 ** 5 / 5) > 1e-4


def test_dense_output_order_conditions(ND, oracle):
    A, _, c = oracle.tsit5_tableau()
    r_or = oracle.tsit5_dense()
    r_dev = ND.ensemble.tsit5_dense()
    assert np.array_equal(r_or, r_dev)
    _check_dense(A, c, r_or)
    # a perturbed table fails: each coefficient enters some order condition
    for i in range(7):
        for j in range(1, 4):
            bad = r_or.copy()
            bad[i, j] += 1e-6
            with pytest.raises(AssertionError):
                _check_dense(A, c, bad)


# ---- the Adler equation of two oscillators on one line ------------------------------

def _pair(ND):
    g = ND.SimpleGraph(2, [1], [2])
    return ND.Network(g, ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge()), g


def _drift_time(delta_end, dw, K):
    """t at which delta' = dw - 2K sin(delta), dw > 2K, from delta = 0 reaches delta_end
    (in (-pi, pi)): the integral of 1/(a - b sin x) = 2/W atan((a tan(x/2) - b)/W)."""
    a, b = dw, 2 * K
    W = np.sqrt(a * a - b * b)
    F = lambda x: 2 / W * np.arctan((a * np.tan(x / 2) - b) / W)
    return F(delta_end) - F(0.0)


def _relax_time(delta0, delta_end, K):
    """t at which delta' = -2K sin(delta) goes from delta0 to delta_end (same sign):
    tan(delta/2) = tan(delta0/2) exp(-2Kt)."""
    return np.log(np.tan(delta0 / 2) / np.tan(delta_end / 2)) / (2 * K)


# (name, omega_1, omega_2, K, limit, delta0, crossing delta, crossing time)
def _cases():
    cases = []
    d_up = np.pi - np.arcsin(0.6)
    cases.append(("up from rest", 0.0, 0.0, 1.0, 0.6, 3.0, d_up, _relax_time(3.0, d_up, 1.0)))
    d_dn = np.arcsin(0.5)
    cases.append(("down from above the limit", 0.0, 0.0, 1.0, 0.5, 1.2, d_dn, _relax_time(1.2, d_dn, 1.0)))
    d_dr = np.arcsin(0.9)
    cases.append(("up while drifting", 1.5, -1.5, 1.0, 0.9, 0.0, d_dr, _drift_time(d_dr, 3.0, 1.0)))
    return cases


def _closed_form_end(case, T):
    """State at T of a case whose line tripped at its crossing: theta_1 + theta_2 = 0
    throughout (omega_1 + omega_2 = 0), after the trip each phase runs at its omega."""
    _, w1, w2, _, _, _, dc, tc = case
    return np.array([dc / 2 + w1 * (T - tc), -dc / 2 + w2 * (T - tc)])


def _oracle_pair(ND, case, T, line_events="crossing", tol=1e-11):
    nw, _ = _pair(ND)
    _, w1, w2, K, L, d0, _, _ = case
    return oracle_for(nw).tsit5(np.array([d0 / 2, -d0 / 2]), np.array([w1, w2, K]), (0.0, T), reltol=tol,
                                abstol=tol, line_limit=np.array([L]), line_events=line_events)


@pytest.mark.parametrize("k", range(3))
def test_oracle_crossing_time_and_state(ND, k):
    case = _cases()[k]
    tc = case[7]
    for T in (tc + 0.05, tc + 1.7):
        u, p, st = _oracle_pair(ND, case, T)
        assert st["retcode"] == 0 and st["ntrip"] == 1 and p[2] == 0.0, (case[0], st)
        np.testing.assert_allclose(u, _closed_form_end(case, T), rtol=0, atol=2e-9, err_msg=case[0])
    # just before the crossing nothing trips
    u, p, st = _oracle_pair(ND, case, tc - 1e-3)
    assert st["ntrip"] == 0 and p[2] == case[3], case[0]


def test_oracle_line_above_limit_waits_for_its_crossing(ND):
    case = _cases()[1]                              # |P| = sin(1.2) = 0.93 > 0.5 at t0
    T = case[7] + 1.0
    u_c, _, st_c = _oracle_pair(ND, case, T, "crossing")
    u_s, _, st_s = _oracle_pair(ND, case, T, "step_end")
    assert st_c["ntrip"] == st_s["ntrip"] == 1
    np.testing.assert_allclose(u_c[0] - u_c[1], case[6], rtol=0, atol=2e-9)     # tripped at delta = pi/6
    assert u_s[0] - u_s[1] > 1.15                  # tripped at the end of the first step


def test_oracle_step_end_trips_late_by_up_to_a_step(ND):
    case = _cases()[2]
    T = case[7] + 1.0
    u_c, _, _ = _oracle_pair(ND, case, T, "crossing", tol=1e-6)
    u_s, _, st_s = _oracle_pair(ND, case, T, "step_end", tol=1e-6)
    ref = _closed_form_end(case, T)
    err_c = np.abs(u_c - ref).max()
    err_s = np.abs(u_s - ref).max()
    assert err_c < 1e-5 and err_s > 10 * err_c, (err_c, err_s)


def _swing_cascade(ND, n_inst, seed, nmin=40, nmax=300):
    """Swing networks with lines loaded near their limits: a kick at one vertex starts
    cascades of crossings."""
    rng = np.random.default_rng(seed)
    gs, us, ps, lims = [], [], [], []
    for i in range(n_inst):
        n = int(rng.integers(nmin, nmax + 1))
        g = ND.barabasi_albert(n, 2, seed=seed * 1000 + i)
        P = rng.normal(scale=0.5, size=n)
        P -= P.mean()
        K = rng.uniform(4.0, 8.0, g.ne())
        pv = np.empty(3 * n)
        pv[0::3], pv[1::3], pv[2::3] = 1.0, 0.5, P
        u = np.zeros(2 * n)
        u[0::2] = rng.normal(scale=0.05, size=n)
        u[2 * int(rng.integers(n)) + 1] = 8.0                         # frequency kick
        th = u[0::2]
        flow = K * np.sin(th[g.src - 1] - th[g.dst - 1])
        gs.append(g)
        us.append(u)
        ps.append(np.r_[pv, K])
        lims.append(np.abs(flow) + rng.uniform(0.3, 1.5, g.ne()))
    return gs, us, ps, lims


def test_oracle_cascade_consistency(ND):
    """On kicked swing networks both modes trip lines, only lines (K > 0 before, 0
    after), and the counts add up."""
    gs, us, ps, lims = _swing_cascade(ND, 2, seed=31)
    for g, u, p, lim in zip(gs, us, ps, lims):
        nw = ND.Network(g, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        on = oracle_for(nw)
        n = g.nv()
        for mode in ("crossing", "step_end"):
            _, pm, sm = on.tsit5(u, p, (0.0, 5.0), line_limit=lim, line_events=mode)
            assert sm["retcode"] == 0 and sm["ntrip"] > 0, mode
            assert np.array_equal(pm[:3 * n], p[:3 * n])
            changed = pm[3 * n:] != p[3 * n:]
            assert (pm[3 * n:][changed] == 0).all() and changed.sum() == sm["ntrip"], mode


# ---- the ensemble kernel --------------------------------------------------------------

@pytest.mark.gpu
def test_ensemble_crossings_match_closed_forms(ND, dev):
    """Each case of the Adler pair as ensemble instances, stopped at several times after
    the crossing, against the closed forms; and stopped just before it, untripped."""
    import torch
    cases = _cases()
    rows = []
    for case in cases:
        tc = case[7]
        for T in (tc - 1e-3, tc + 0.05, tc + 1.7):
            rows.append((case, T))
    # one ensemble per stopping time (tspan is shared by the instances of a solve)
    for case, T in rows:
        nw, g = _pair(ND)
        ens = ND.Ensemble([g], ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
        _, w1, w2, K, L, d0, _, tc = case
        u_d = torch.tensor([d0 / 2, -d0 / 2], dtype=torch.float64, device=dev)
        p_d = torch.tensor([w1, w2, K], dtype=torch.float64, device=dev)
        lim = torch.tensor([L], dtype=torch.float64, device=dev)
        st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, T), reltol=1e-11, abstol=1e-11, line_limit=lim,
                                                  line_events="crossing"))
        u, p = u_d.cpu().numpy(), p_d.cpu().numpy()
        if T < tc:
            assert st["ntrip"][0] == 0 and p[2] == K, case[0]
        else:
            assert st["retcode"][0] == 0 and st["ntrip"][0] == 1 and p[2] == 0.0, (case[0], T)
            np.testing.assert_allclose(u, _closed_form_end(case, T), rtol=0, atol=2e-9, err_msg=f"{case[0]} T={T}")


@pytest.mark.gpu
def test_ensemble_line_above_limit_waits_for_its_crossing(ND, dev):
    import torch
    case = _cases()[1]
    _, w1, w2, K, L, d0, dc, tc = case
    out = {}
    for mode in ("crossing", "step_end"):
        _, g = _pair(ND)
        ens = ND.Ensemble([g], ND.Lib.kuramoto_first(), ND.Lib.kuramoto_edge())
        u_d = torch.tensor([d0 / 2, -d0 / 2], dtype=torch.float64, device=dev)
        p_d = torch.tensor([w1, w2, K], dtype=torch.float64, device=dev)
        lim = torch.tensor([L], dtype=torch.float64, device=dev)
        ens.solve(u_d, p_d, (0.0, tc + 1.0), reltol=1e-11, abstol=1e-11, line_limit=lim, line_events=mode)
        out[mode] = u_d.cpu().numpy()
    assert abs(out["crossing"][0] - out["crossing"][1] - dc) < 2e-9
    assert out["step_end"][0] - out["step_end"][1] > 1.15


@pytest.mark.gpu
def test_ensemble_cascades_match_oracle(ND, dev):
    """Kicked swing networks with lines near their limits: the kernel trips the same
    lines as the oracle (parameters bitwise equal), counts the same trips, and ends in
    the same state within 1e-6 scaled (a root located a few ulps apart moves the rest of
    the run by rounding only, but the step sequences then differ)."""
    import torch
    gs, us, ps, lims = _swing_cascade(ND, 12, seed=41)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    lim_d = torch.from_numpy(np.concatenate(lims)).to(dev)
    st = ND.Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, 5.0), reltol=1e-8, abstol=1e-8, line_limit=lim_d,
                                              line_events="crossing"))
    uf, pf = u_d.cpu().numpy(), p_d.cpu().numpy()

    def run(i):
        nw = ND.Network(gs[i], ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
        return oracle_for(nw).tsit5(us[i], ps[i], (0.0, 5.0), reltol=1e-8, abstol=1e-8, line_limit=lims[i],
                                    line_events="crossing")

    with ThreadPoolExecutor(8) as ex:
        refs = list(ex.map(run, range(len(gs))))
    trips = 0
    for i, (ur, pr, sr) in enumerate(refs):
        assert st["retcode"][i] == sr["retcode"] == 0, i
        assert st["ntrip"][i] == sr["ntrip"], (i, st["ntrip"][i], sr["ntrip"])
        assert np.array_equal(pf[ens.param_slice(i)], pr), i
        assert abs(int(st["naccept"][i]) - sr["naccept"]) <= 2, (i, st["naccept"][i], sr["naccept"])
        ui = uf[ens.state_slice(i)]
        err = np.abs(ui - ur) / (1e-8 + 1e-8 * np.maximum(np.abs(ui), np.abs(ur)))
        assert err.max() <= 100.0, (i, err.max())          # 1e-6 relative at these tolerances
        trips += sr["ntrip"]
    assert trips >= 2 * len(gs)


@pytest.mark.gpu
def test_ensemble_crossing_graph_replay_and_errors(ND, dev):
    import torch
    gs, us, ps, lims = _swing_cascade(ND, 6, seed=43)
    ens = ND.Ensemble(gs, ND.Lib.kuramoto_second(), ND.Lib.kuramoto_edge())
    lim_d = torch.from_numpy(np.concatenate(lims)).to(dev)
    u_e = torch.from_numpy(np.concatenate(us)).to(dev)
    p_e = torch.from_numpy(np.concatenate(ps)).to(dev)
    ens.solve(u_e, p_e, (0.0, 3.0), line_limit=lim_d)             # default: crossing for Tsit5
    u_g = torch.from_numpy(np.concatenate(us)).to(dev)
    p_g = torch.from_numpy(np.concatenate(ps)).to(dev)
    graph = ens.capture(u_g, p_g, (0.0, 3.0), line_limit=lim_d, line_events="crossing")
    graph.launch()
    torch.cuda.synchronize()
    assert torch.equal(u_g, u_e) and torch.equal(p_g, p_e)
    del graph
    with pytest.raises(ND.UnsupportedError):
        ens.solve(u_e, p_e, (0.0, 0.1), method=ND.ensemble.RK4, dt=0.01, line_limit=lim_d, line_events="crossing")
    with pytest.raises(ND.ArgumentError):
        ens.solve(u_e, p_e, (0.0, 0.1), line_limit=lim_d, line_events="sometimes")
