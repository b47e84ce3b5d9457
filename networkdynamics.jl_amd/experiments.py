"""Workload runs beyond the single-network RHS: each returns a dict of measurements.

c1_rk4        first-order Kuramoto on G(1000, 4000): RK4 dt = 0.01 to t = 10 through the
              fused per-stage RHS launches and through the one-launch ensemble kernel
c3_rhs        heterogeneous swing/load network on a random geometric graph, N = 2e5
c4_ensemble   4096 uneven Barabasi-Albert swing networks (N ~ U{100..2000}) from their
              fixpoints plus a one-node perturbation, Tsit5 to t = 50 (reltol 1e-8):
              the fused ensemble kernel against the per-instance GPU path
c5_warmstart  swing/load grids: Newton from flat starts vs from the learned warm start
extras        find_fixpoint, colored Jacobians, the DAE integrator, and the per-vertex
              kernel against the fused ones
"""
import sys
import time

import numpy as np

from . import graphs as G
from . import library as Lib
from . import workloads
from .ensemble import RK4, TSIT5, Ensemble, tsit5_per_instance
from .fixpoint import FixpointSolver, isfixpoint_residual, swing_load_grid
from .network import Network


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def _timeit(fn, reps, warm=2):
    import torch
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps * 1e-3


def c1_rk4(dev):
    import torch
    nw, p, u0 = workloads.kuramoto_er(1000, 4000, 3.0, seed=1)
    pd = torch.from_numpy(p).to(dev)
    u = torch.from_numpy(u0).to(dev)
    du = torch.empty_like(u)
    t_rhs = _timeit(lambda: nw(du, u, pd, 0.0), 200)
    uu = u.clone()
    nw.tune(rk4_path=2)
    t_plan = _timeit(lambda: (uu.copy_(u), nw.rk4(uu, pd, 0.0, 0.01, 1000)), 3, warm=1)
    nw.tune(rk4_path=1)
    ue = u.clone()
    t_ens = _timeit(lambda: (ue.copy_(u), nw.rk4(ue, pd, 0.0, 0.01, 1000)), 3, warm=1)
    return dict(workload="Kuramoto ER N=1000 E=4000, RK4 dt=0.01 to t=10", rhs_us=t_rhs * 1e6,
                rk4_fused_stage_launches_ms=t_plan * 1e3, rk4_one_launch_ms=t_ens * 1e3,
                max_state_difference=float((uu - ue).abs().max().item()))


def c3_rhs(dev):
    import torch
    nw, p, u0 = workloads.heterogeneous_rgg(200_000, seed=1)
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    du = torch.empty_like(u)
    t = _timeit(lambda: nw(du, u, pd, 0.0), 200)
    algo = 8 * nw.dim * 2 + 8 * nw.pdim + 8 * nw.ne()
    return dict(workload=f"heterogeneous RGG N={nw.nv()} E={nw.ne()} (2 vertex batches)", rhs_us=t * 1e6,
                rhs_per_s=1.0 / t, algorithmic_bytes=algo, achieved_GBs=algo / t / 1e9)


def c4_instances(ninst, seed=1, nmin=100, nmax=2000):
    """Uneven BA swing networks with C2 parameters whose synchronous fixpoint exists."""
    rng = np.random.default_rng(seed)
    out = []
    k = 0
    while len(out) < ninst:
        cand = []
        for _ in range(ninst - len(out)):
            n = int(rng.integers(nmin, nmax + 1))
            g = G.barabasi_albert(n, 2, seed=seed * 100000 + k)
            k += 1
            M, D, P, K = workloads.swing_params(n, g.ne(), rng)
            cand.append((g, M, D, P, K))
        out += cand
    return out, rng


def c4_start(dev, ninst=4096, seed=1, nmin=100, nmax=2000):
    """The C4 starting states: uneven BA swing networks at their synchronous fixpoints
    (batched Newton on the device) plus one perturbed node each (dtheta ~ U(-pi, pi),
    domega ~ U(-10, 10)).  Returns graphs, per-instance states u and parameters p
    (flat layout [M, D, P] per vertex then K per edge), the number of instances dropped
    for lack of a fixpoint, and the Newton time."""
    import torch
    inst, rng = c4_instances(ninst, seed, nmin, nmax)
    fs = FixpointSolver([x[0] for x in inst])
    th = torch.zeros(int(fs.voff[-1]), dtype=torch.float64, device=dev)
    P = torch.from_numpy(np.concatenate([x[3] for x in inst])).to(dev)
    S = torch.from_numpy(np.concatenate([x[1] for x in inst])).to(dev)
    K = torch.from_numpy(np.concatenate([x[4] for x in inst])).to(dev)
    torch.cuda.synchronize()
    tn = time.perf_counter()
    st = FixpointSolver.stats_to_numpy(fs.solve(th, P, S, K))
    torch.cuda.synchronize()
    t_newton = time.perf_counter() - tn
    ok = st["retcode"] == 0
    _log(f"fixpoints: {int(ok.sum())} of {len(inst)} in {t_newton:.2f} s")
    thh = th.cpu().numpy()
    graphs, us, ps = [], [], []
    for i in range(len(inst)):
        if not ok[i]:
            continue
        g, M, D, Pi, Ki = inst[i]
        n = g.nv()
        u = np.zeros(2 * n)
        u[0::2] = thh[fs.voff[i]:fs.voff[i + 1]]
        j = int(rng.integers(n))
        u[2 * j] += rng.uniform(-np.pi, np.pi)
        u[2 * j + 1] += rng.uniform(-10.0, 10.0)
        p = np.empty(3 * n + g.ne())
        p[0:3 * n:3], p[1:3 * n:3], p[2:3 * n:3], p[3 * n:] = M, D, Pi, Ki
        graphs.append(g)
        us.append(u)
        ps.append(p)
    return graphs, us, ps, int((~ok).sum()), t_newton, rng


def returned_to_sync(u):
    """Sync classification of a swing state (theta, omega interleaved): every |omega| < 0.5."""
    return bool(np.abs(np.asarray(u)[1::2]).max() < 0.5)


def c4_ensemble(dev, ninst=4096, t1=50.0, reltol=1e-8, abstol=1e-8, baseline_seconds=30.0, seed=1, tune=0,
                steady_state_tol=1e-5, one_workgroup_seconds=10.0, tune_cache=None):
    """tune: pick the size classes with ensemble_tune.autotune first (tune_cache: its JSON
    cache file, so that a later run, e.g. under a profiler, reuses the choice without
    probing)."""
    import torch
    t0 = time.perf_counter()
    graphs, us, ps, dropped, t_newton, rng = c4_start(dev, ninst, seed)
    t_setup = time.perf_counter() - t0
    ens = Ensemble(graphs, Lib.kuramoto_second(), Lib.kuramoto_edge())
    u_d = torch.from_numpy(np.concatenate(us)).to(dev)
    p_d = torch.from_numpy(np.concatenate(ps)).to(dev)
    tuning, t_tune = None, 0.0
    if tune:
        from .ensemble_tune import autotune
        tt = time.perf_counter()
        tuning = autotune(ens, u_d, p_d, probe_t=0.5, reltol=reltol, abstol=abstol, cache=tune_cache)
        t_tune = time.perf_counter() - tt
        _log(f"autotune: {tuning['classes']} in {t_tune:.1f} s")
    torch.cuda.synchronize()
    ts = time.perf_counter()
    est = Ensemble.stats_to_numpy(ens.solve(u_d, p_d, (0.0, t1), method=TSIT5, reltol=reltol, abstol=abstol))
    torch.cuda.synchronize()
    t_ens = time.perf_counter() - ts
    uf = u_d.cpu().numpy()
    sync = np.array([returned_to_sync(uf[ens.state_slice(i)]) for i in range(len(graphs))])
    # the same solve captured once as a hipGraph and replayed
    u_g = torch.from_numpy(np.concatenate(us)).to(dev)
    graph = ens.capture(u_g, p_d, (0.0, t1), method=TSIT5, reltol=reltol, abstol=abstol)
    torch.cuda.synchronize()
    tg = time.perf_counter()
    graph.launch()
    torch.cuda.synchronize()
    t_graph = time.perf_counter() - tg
    graph_equal = bool(torch.equal(u_g, u_d))
    del graph
    # basin-stability mode: each instance stops once settled (every |du_i| <= tol); the
    # ones that do not return to sync run to t1
    ss = None
    if steady_state_tol > 0:
        u_s = torch.from_numpy(np.concatenate(us)).to(dev)
        torch.cuda.synchronize()
        tss = time.perf_counter()
        sst = Ensemble.stats_to_numpy(ens.solve(u_s, p_d, (0.0, t1), method=TSIT5, reltol=reltol, abstol=abstol,
                                                steady_state=(steady_state_tol, 0.0)))
        torch.cuda.synchronize()
        t_ss = time.perf_counter() - tss
        term = sst["retcode"] == 4
        ss = dict(tol=steady_state_tol, seconds=t_ss, traj_per_s=len(graphs) / t_ss, terminated=int(term.sum()),
                  terminated_and_synced=int((term & sync).sum()), mean_stop_t=float(sst["t"][term].mean())
                  if term.any() else None, steps_mean=float(sst["naccept"].mean()))
        _log(f"steady-state termination: {int(term.sum())} stopped early, {t_ss:.2f} s")
    # plain per-instance GPU path on a time-bounded sample (largest-first order not used: random sample)
    _log(f"ensemble: {len(graphs)} trajectories in {t_ens:.2f} s")
    order = rng.permutation(len(graphs))
    done, t_base = 0, 0.0
    steps = []
    for i in order:
        nw = Network(graphs[i], Lib.kuramoto_second(), Lib.kuramoto_edge())      # construction not timed
        ui = torch.from_numpy(us[i]).to(dev)
        pi = torch.from_numpy(ps[i]).to(dev)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        info = tsit5_per_instance(nw, ui, pi, (0.0, t1), reltol=reltol, abstol=abstol)
        torch.cuda.synchronize()
        t_base += time.perf_counter() - tb
        steps.append(info["naccept"])
        done += 1
        if t_base > baseline_seconds:
            break
    _log(f"per-instance path: {done} trajectories in {t_base:.2f} s")
    # one network at a time through nd_tsit5 (all steps of a network in one workgroup), on
    # a time-bounded sample; each network's one-instance setup is inside the timing
    done1, t_one, steps1, paths = 0, 0.0, [], set()
    for i in order:
        if one_workgroup_seconds <= 0:
            break
        nw = Network(graphs[i], Lib.kuramoto_second(), Lib.kuramoto_edge())
        nw.plan                                                                  # plan construction not timed
        ui = torch.from_numpy(us[i]).to(dev)
        pi = torch.from_numpy(ps[i]).to(dev)
        torch.cuda.synchronize()
        tb = time.perf_counter()
        info = nw.tsit5(ui, pi, (0.0, t1), reltol=reltol, abstol=abstol)
        torch.cuda.synchronize()
        t_one += time.perf_counter() - tb
        steps1.append(info["naccept"])
        paths.add(info["path"])
        done1 += 1
        if t_one > one_workgroup_seconds:
            break
    if done1:
        _log(f"one-workgroup per-instance path: {done1} trajectories in {t_one:.2f} s")
    one_wg = None if not done1 else dict(
        instances=done1, seconds=t_one, traj_per_s=done1 / t_one, steps_mean=float(np.mean(steps1)),
        paths=sorted(paths), ensemble_speedup=(len(graphs) / t_ens) / (done1 / t_one))
    traj_ens = len(graphs) / t_ens
    traj_base = done / t_base
    return dict(workload=f"{len(graphs)} BA(m=2) swing networks N~U{{100..2000}}, Tsit5 t=0..{t1}, "
                         f"reltol={reltol}, abstol={abstol}",
                instances=len(graphs), dropped_without_fixpoint=dropped, newton_s=t_newton,
                setup_s=t_setup, ensemble_s=t_ens, ensemble_traj_per_s=traj_ens,
                graph_replay_s=t_graph, graph_replay_traj_per_s=len(graphs) / t_graph,
                graph_replay_equals_eager=graph_equal,
                ensemble_steps_mean=float(est["naccept"].mean()), ensemble_rejects_mean=float(est["nreject"].mean()),
                ensemble_retcodes=np.bincount(est["retcode"], minlength=5).tolist(),
                returned_to_sync=int(sync.sum()), baseline_instances=done, baseline_s=t_base,
                baseline_traj_per_s=traj_base, baseline_steps_mean=float(np.mean(steps)),
                speedup=traj_ens / traj_base, one_workgroup_per_instance=one_wg,
                classes=[list(c) for c in ens.classes], steady_state=ss,
                tuning=None if tuning is None else dict(predicted_s=tuning["predicted_s"], bins=tuning["bins"],
                                                        seconds=t_tune))


def _grid_job(args):
    n, s = args
    return swing_load_grid(int(n), seed=int(s))


def c5_grids(count, seed, nmin=500, nmax=5000, workers=16):
    """Seeded swing/load grids; generated in a process pool while the GPU is untouched."""
    rng = np.random.default_rng(seed)
    jobs = [(int(n), seed * 1000003 + i) for i, n in enumerate(rng.integers(nmin, nmax + 1, count))]
    try:
        import torch
        gpu_started = torch.cuda.is_initialized()
    except ImportError:
        gpu_started = False
    if count < 64 or workers < 2 or gpu_started:
        return [_grid_job(j) for j in jobs]
    import multiprocessing as mp
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_grid_job, jobs, chunksize=32)


def c5_warmstart(dev, ntrain=20000, neval=1000, epochs=30, batch_grids=50, lr=2e-3, seed=7, warm_tol=1e-6,
                 workers=16):
    """Train on fixpoints of ntrain grids, evaluate Newton from flat / DC / physics / learned
    starts on neval held-out grids (grids without a flat-start fixpoint are dropped).
    workers: processes generating the grids (1: in this process, e.g. under a profiler
    that would initialise the GPU in each forked worker)."""
    import torch
    from . import warmstart as W
    out = {}
    t0 = time.perf_counter()
    train = c5_grids(ntrain, seed, workers=workers)
    evalg = c5_grids(neval, seed + 1, workers=workers)
    out["generate_s"] = time.perf_counter() - t0
    _log(f"generated {len(train)} + {len(evalg)} grids in {out['generate_s']:.1f} s")
    th, st, tsol = W.solve_grids(train, dev)
    ok = st["retcode"] == 0
    out.update(train_grids=int(ok.sum()), train_dropped=int((~ok).sum()), train_newton_s=tsol)
    _log(f"training fixpoints: {int(ok.sum())} converged, {tsol:.1f} s")
    train = [g for g, o in zip(train, ok) if o]
    th = [W.center(t) for t, o in zip(th, ok) if o]
    dc, s0, X, _ = W.physics_inputs(train, dev)
    tgt = [W.source(g, t) - a for g, t, a in zip(train, th, s0)]
    net = W.WarmStartNet(dev, seed=0)
    Xall = np.concatenate(X)
    net.xmean, net.xstd = Xall.mean(axis=0), Xall.std(axis=0) + 1e-12
    net.tscale = 1.0 / np.concatenate(tgt).std()
    batches = []
    for b in range(0, len(train), batch_grids):
        gs = train[b:b + batch_grids]
        yb = np.concatenate(tgt[b:b + batch_grids]) * net.tscale
        batches.append((W.GraphBatch(gs, dev), net.normalise(np.concatenate(X[b:b + batch_grids])),
                        torch.from_numpy(yb.astype(np.float32)).to(dev)))
    torch.cuda.synchronize()
    tt = time.perf_counter()
    hist = []
    for ep in range(epochs):
        tot = 0.0
        for bt, Xd, yd in batches:
            tot += net.train_step(bt, Xd, yd, lr=lr * (0.1 if ep >= int(0.8 * epochs) else 1.0)).item()
        hist.append(tot / len(batches))
        _log(f"epoch {ep}: loss {hist[-1]:.4e}")
    torch.cuda.synchronize()
    out.update(train_s=time.perf_counter() - tt, loss_first=hist[0], loss_last=hist[-1], epochs=epochs,
               adam_steps=epochs * len(batches))
    # held-out grids: the static per-grid data is built once, every per-solve step runs
    # on the GPU (DC solve, features, network, warm-start solve, Newton)
    tg = time.perf_counter()
    gb0 = W.GridBatch(evalg, dev)
    t_static0 = time.perf_counter() - tg
    fth, fst, ft = gb0.solve()
    keep = fst["retcode"] == 0
    evalg = [g for g, o in zip(evalg, keep) if o]
    tg = time.perf_counter()
    gb = W.GridBatch(evalg, dev)
    out["static_prep_s"] = time.perf_counter() - tg
    fth, fst, ft = gb.solve()
    _, fx, ftx = gb.solve(inexact=False)
    acc_flat = sum(isfixpoint_residual(g, t, dev) <= 1e-10 for g, t in
                   zip(evalg, np.split(fth.cpu().numpy(), gb.voff[1:-1])))
    out.update(eval_grids=len(evalg), eval_dropped=int((~keep).sum()), flat_newton_mean=float(fst["newton_iters"].mean()),
               flat_newton_mean_exact_cg=float(fx["newton_iters"].mean()), flat_s=ft, flat_exact_cg_s=ftx,
               accepted_flat=int(acc_flat))
    # target fit on the held-out grids
    s0_e, _ = gb.features(gb.center(gb.solve(linearized=True)[0]))
    t_true = gb.source(gb.center(fth)) - s0_e
    starts = {}
    th_dc, _, t_dc = gb.solve(linearized=True, tol=warm_tol)
    starts["dc"] = (gb.center(th_dc), dict(dc_solve=t_dc))
    starts["physics"] = W.warm_start_device(gb, None, tol=warm_tol)
    starts["learned"] = W.warm_start_device(gb, net, tol=warm_tol)
    # the network's prediction of t on the held-out grids
    import torch
    xm, xs = torch.from_numpy(net.xmean).to(dev), torch.from_numpy(net.xstd).to(dev)
    _, Xe = gb.features(starts["dc"][0])
    pred = gb.center(net.forward(gb.graph, ((Xe - xm) / xs).to(torch.float32).contiguous()).double() / net.tscale)
    tt = gb.center(t_true)
    out["target_r2"] = float(1.0 - ((pred - tt) ** 2).sum().item() / (tt ** 2).sum().item())
    for name, (th0, times) in starts.items():
        sth, sst, stt = gb.solve(theta0=th0)
        _, xst, xstt = gb.solve(theta0=th0, inexact=False)
        acc = sum(isfixpoint_residual(g, t, dev) <= 1e-10 for g, t in
                  zip(evalg, np.split(sth.cpu().numpy(), gb.voff[1:-1])))
        prep = sum(times.values())
        out[f"{name}_newton_mean"] = float(sst["newton_iters"].mean())
        out[f"{name}_newton_mean_exact_cg"] = float(xst["newton_iters"].mean())
        out[f"{name}_retcodes"] = np.bincount(sst["retcode"], minlength=4).tolist()
        out[f"{name}_accepted"] = int(acc)
        out[f"{name}_prep_s"] = prep
        out[f"{name}_prep_detail_s"] = times
        out[f"{name}_newton_s"] = stt
        out[f"{name}_total_s"] = prep + stt
        out[f"{name}_total_exact_cg_s"] = prep + xstt
    # the pipeline as a user runs it: learned start (loose pre-solves, capped Newton),
    # flat-start fallback for the grids it fails on, every result checked by its RHS residual
    for rep in range(2):           # the first call also compiles nothing; keep the second
        t0 = time.perf_counter()
        pth, pst, fell, ptimes = W.learned_solve(gb, net)
        torch.cuda.synchronize()
        ptot = time.perf_counter() - t0
    acc = sum(isfixpoint_residual(g, t, dev) <= 1e-10 for g, t in
              zip(evalg, np.split(pth.cpu().numpy(), gb.voff[1:-1])))
    out.update(pipeline_total_s=ptot, pipeline_detail_s=ptimes, pipeline_fell_back=int(fell.sum()),
               pipeline_accepted=int(acc), pipeline_newton_mean=float(pst["newton_iters"].mean()),
               time_ratio_pipeline=out["flat_s"] / ptot)
    out.update(newton_ratio_learned=out["flat_newton_mean"] / max(out["learned_newton_mean"], 1e-9),
               newton_ratio_learned_exact_cg=out["flat_newton_mean_exact_cg"] / max(out["learned_newton_mean_exact_cg"], 1e-9),
               time_ratio_learned=out["flat_s"] / out["learned_total_s"],
               time_ratio_learned_exact_cg=out["flat_exact_cg_s"] / out["learned_total_exact_cg_s"])
    return out


def extras(dev, n_grid=4000, n_ba=1_000_000):
    """Timings of the paths around the RHS kernels: find_fixpoint (sparse LU of the colored
    Jacobian) on a DAE grid, one colored Jacobian, the DAE integrator per step, and the
    one-thread-per-vertex kernel against the fused kernels on the same BA graph (resistor
    edges y = (v_s - v_d)/R against weighted diffusion edges y = w (v_s - v_d))."""
    import torch
    from .initialization import find_fixpoint
    from .sparsity import color_columns, jac_prototype, sparse_jacobian
    from . import dae
    out = {}
    grid = swing_load_grid(n_grid, seed=53)
    nw = grid.network(algebraic_loads=True)
    p = grid.flat_p()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, info = find_fixpoint(nw, np.zeros(nw.dim), p)
    torch.cuda.synchronize()
    out["find_fixpoint"] = dict(states=nw.dim, seconds=time.perf_counter() - t0, **info)
    P = jac_prototype(nw)
    colors, ncolor = color_columns(P)
    pd = torch.from_numpy(p).to(dev)
    sparse_jacobian(nw, x, pd, 0.0, P, colors)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        sparse_jacobian(nw, x, pd, 0.0, P, colors)
    torch.cuda.synchronize()
    out["sparse_jacobian"] = dict(states=nw.dim, nnz=int(P.nnz), colors=int(ncolor),
                                  ms=(time.perf_counter() - t0) / 10 * 1e3)
    rng = np.random.default_rng(3)
    u0 = grid.flat_u(rng.normal(scale=0.05, size=grid.g.nv()), rng.normal(scale=0.1, size=grid.g.nv()))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, dinfo = dae.rk4(nw, u0, p, 0.0, 0.01, 100)
    torch.cuda.synchronize()
    out["dae_rk4"] = dict(states=nw.dim, steps=100, ms_per_step=(time.perf_counter() - t0) / 100 * 1e3, **dinfo)
    # generic vs fused kernel on one graph
    g = G.barabasi_albert(n_ba, 2, seed=1)
    res = {}
    for name, em in (("fused (weighted diffusion)", Lib.diffusion_edge()), ("per-vertex (resistor)", Lib.resistor())):
        net = Network(g, Lib.diffusion_vertex(), em)
        u = torch.from_numpy(rng.normal(size=net.dim)).to(dev)
        pp = torch.from_numpy(rng.uniform(0.5, 2.0, net.pdim)).to(dev)
        du = torch.empty_like(u)
        res[name] = _timeit(lambda: net(du, u, pp, 0.0), 200) * 1e6
    out["rhs_us_ba_1e6_diffusion"] = res
    return out
