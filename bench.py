"""Benchmark: single network RHS nw(du, u, p, t) of the N = 1e6 swing network.

Workload: second-order Kuramoto (swing) vertices on a Barabasi-Albert graph
N = 1,000,000, m = 2 (E = 1,999,996), Kuramoto edges K ~ U(5,10); FP64.
A step is one RHS evaluation with u, p, du resident in HBM.

Prints one JSON line: RHS evals/s, the roofline of one RHS (algorithmic bytes
56 N + 16 E over the HIP-event-timed average duration of the RHS's launches,
against 8 TB/s), the CPU baseline (OpenMP restatement of the reference's
ThreadedExecution + ThreadedAggregator path, timed on this host), the achievable
bandwidth of a STREAM-like device copy, `extended.c4`: the ensemble's
trajectories/s on the basin-stability workload (4096 uneven swing networks, Tsit5
to t = 50) against the per-instance paths, and `extended.c5`: Newton iterations and
device seconds of the learned warm start against the flat start on held-out grids,
with the default inexact and with exact CG linear solves.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "swing_ba_1e6_traffic.txt")


def neighbour_gather(dev, g, reps=20):
    """The RHS's neighbour-read pattern on its own: torch.index_select of an N-vector of
    doubles (the compact vertex outputs) at every list entry's neighbour, entries in list
    order (each vertex's neighbours sorted, as the destination lists are), int32 indices;
    and the same number of reads at sequential indices.  Microseconds per gather."""
    import numpy as np
    import torch
    n = g.nv()
    s, d = g.src - 1, g.dst - 1
    owner = np.concatenate([d, s])
    nbr = np.concatenate([s, d])
    order = np.lexsort((nbr, owner))
    idx = torch.from_numpy(nbr[order].astype(np.int32)).to(dev)
    seq = torch.arange(idx.numel(), dtype=torch.int32, device=dev) % n
    vo = torch.rand(n, dtype=torch.float64, device=dev)
    out = {}
    for name, ix in (("random", idx), ("sequential", seq)):
        for _ in range(3):
            torch.index_select(vo, 0, ix)
        torch.cuda.synchronize()
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            torch.index_select(vo, 0, ix)
        b.record()
        torch.cuda.synchronize()
        out[name + "_us"] = a.elapsed_time(b) * 1e3 / reps
    out["reads"] = int(idx.numel())
    return out


def stream_copy(dev, mib, reps=20):
    """Achievable bandwidth of a device copy of `mib` MiB (read + write counted), GB/s."""
    import torch
    n = mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        b.copy_(a)
    e.record()
    torch.cuda.synchronize()
    t = s.elapsed_time(e) * 1e-3 / reps
    del a, b
    return 2 * n * 8 / t / 1e9


def _traffic_from_profile():
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if present."""
    try:
        for line in open(PROFILE_TRAFFIC):
            if line.startswith("traffic_bytes_per_rhs:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def cpu_baseline(nw, p, u0, seconds=10.0, what="full N=1e6 swing network"):
    """ThreadedExecution{true} + ThreadedAggregator(+) restated in C/OpenMP, all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    on = oracle.OracleNet(nw.nv(), nw.g.src, nw.g.dst, nw.kinds, nw.vertex_kind, nw.edge_kind)
    du = np.empty(nw.dim)
    on.rhs_threaded(u0, p, nthreads=cores, du=du)     # warm-up
    n, t0 = 0, time.perf_counter()
    best = float("inf")
    while True:
        t1 = time.perf_counter()
        on.rhs_threaded(u0, p, nthreads=cores, du=du)
        best = min(best, time.perf_counter() - t1)
        n += 1
        if time.perf_counter() - t0 > seconds and n >= 3:
            break
    el = time.perf_counter() - t0
    return dict(value=n / el, unit="RHS evals/s", cores=cores, kind="port", label="restatement",
                best_ms=best * 1e3,
                sample=f"{what}, {n} calls in {el:.1f} s, {cores} OpenMP threads "
                       "(C restatement of ThreadedExecution{true}+ThreadedAggregator(+))")


def cpu_baseline_c1(ND, seconds=10.0):
    """The C1 workload on the host, SequentialExecution restated in C (one thread): RHS
    calls for about half the budget, then whole RK4 runs (dt = 0.01 to t = 10)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    nw, p, u0 = ND.workloads.kuramoto_er(1000, 4000, 3.0, seed=1)
    on = oracle.OracleNet(nw.nv(), nw.g.src, nw.g.dst, nw.kinds, nw.vertex_kind, nw.edge_kind)
    on.rhs(u0, p)
    n, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds / 2 or n < 3:
        on.rhs(u0, p)
        n += 1
    t_rhs = (time.perf_counter() - t0) / n
    m, t1 = 0, time.perf_counter()
    while time.perf_counter() - t1 < seconds / 2 or m < 1:
        on.rk4(u0, p, 0.0, 0.01, 1000)
        m += 1
    t_rk4 = (time.perf_counter() - t1) / m
    return dict(rhs_us=t_rhs * 1e6, rhs_us_inside_rk4=t_rk4 / 4000 * 1e6, rk4_ms=t_rk4 * 1e3, cores=1,
                kind="port", label="restatement",
                sample=f"the C1 network: {n} RHS calls through the ctypes binding (its call overhead "
                       f"included), then {m} RK4 runs of 1000 steps in C, one thread (C restatement of "
                       "SequentialExecution)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="leave out the extended.c4 ensemble measurement")
    ap.add_argument("--no-c5", action="store_true", help="leave out the extended.c5 warm-start measurement")
    ap.add_argument("--c4-baseline-seconds", type=float, default=10.0,
                    help="time budget of each per-instance C4 baseline sample")
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "extras"],
                    help="c2 (default): the N=1e6 swing RHS line; c1/c3/c4/c5: the other workloads; "
                         "extras: timings of the paths around the RHS kernels")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="keyword arguments of the c1/c3/c4/c5 workload runs (numbers)")
    ap.add_argument("--calibrate-copy-mb", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the single-GPU "
              "kernel is the per-GPU share of a sharded network", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import __graft_entry__ as entry

    ND = entry.load()
    dev = torch.device("cuda:0")
    if args.config != "c2":
        from importlib import import_module
        ex = import_module(ND.__name__ + ".experiments")
        fn = dict(c1=ex.c1_rk4, c3=ex.c3_rhs, c4=ex.c4_ensemble, c5=ex.c5_warmstart, extras=ex.extras)[args.config]
        def value(v):
            for conv in (int, float):
                try:
                    return conv(v)
                except ValueError:
                    pass
            return v                            # a string, such as a file path
        kw = {k: value(v) for k, v in (x.split("=", 1) for x in args.set)}
        res = fn(dev, **kw)
        if not args.no_cpu and args.config == "c1":
            res["cpu_baseline"] = cpu_baseline_c1(ND, args.cpu_seconds)
        elif not args.no_cpu and args.config == "c3":
            nw3, p3, u3 = ND.workloads.heterogeneous_rgg(200_000, seed=1)
            res["cpu_baseline"] = cpu_baseline(nw3, p3, u3, args.cpu_seconds, "the C3 network (N=2e5 RGG)")
        print(json.dumps(dict(config=args.config, **res)))
        return
    nw, p, u0 = ND.workloads.swing_ba(args.n, 2, seed=1)
    N, E = nw.nv(), nw.ne()
    u = torch.from_numpy(u0).to(dev)
    pd = torch.from_numpy(p).to(dev)
    du = torch.empty(nw.dim, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)

    if args.calibrate_copy_mb:
        a = torch.ones(args.calibrate_copy_mb * 2**20 // 8, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    for _ in range(args.warmup):
        nw(du, u, pd, 0.0)
    torch.cuda.synchronize()

    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record(stream)
    for _ in range(args.steps):
        nw(du, u, pd, 0.0)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = start.elapsed_time(stop)
    per_rhs_s = ev_ms * 1e-3 / args.steps           # all launches of one RHS, HIP events on its stream
    value = args.steps / wall

    algo_bytes = ND.workloads.algorithmic_bytes_swing(N, E)
    achieved = algo_bytes / per_rhs_s / 1e9
    traffic = _traffic_from_profile()
    copy_hbm = stream_copy(dev, 2048)               # 4 GiB moved: beyond the 256 MB Infinity Cache
    copy_small = stream_copy(dev, 64)               # 128 MiB moved: cache-resident
    gather = neighbour_gather(dev, nw.g)            # the neighbour reads alone, through torch

    cpu = None if args.no_cpu else cpu_baseline(nw, p, u0, args.cpu_seconds)
    c4 = None
    if not args.no_c4:
        del u, pd, du
        torch.cuda.empty_cache()
        from importlib import import_module
        ex = import_module(ND.__name__ + ".experiments")
        # size classes picked by the autotuner (its short probing solves run before the
        # timed solve, like plan construction)
        r = ex.c4_ensemble(dev, baseline_seconds=args.c4_baseline_seconds,
                           one_workgroup_seconds=args.c4_baseline_seconds / 2, steady_state_tol=0.0, tune=1)
        ow = r["one_workgroup_per_instance"] or {}
        c4 = {"workload": r["workload"], "unit": "trajectories/s", "instances": r["instances"],
              "ensemble_traj_per_s": r["ensemble_traj_per_s"],
              "graph_replay_traj_per_s": r["graph_replay_traj_per_s"],
              "graph_replay_equals_eager": r["graph_replay_equals_eager"],
              "one_workgroup_per_instance_traj_per_s": ow.get("traj_per_s"),
              "one_workgroup_sample": ow.get("instances"),
              "host_stepped_per_instance_traj_per_s": r["baseline_traj_per_s"],
              "host_stepped_sample": r["baseline_instances"],
              "ratio_vs_one_workgroup": ow.get("ensemble_speedup"),
              "ratio_vs_host_stepped": r["speedup"],
              "steps_mean": r["ensemble_steps_mean"], "retcodes": r["ensemble_retcodes"],
              "returned_to_sync": r["returned_to_sync"], "classes": r["classes"],
              "autotune_s": r["tuning"]["seconds"], "target_ratio": 20.0}
    c5 = None
    if not args.no_c5:
        # in a child process: its grid generation forks a process pool, which needs a
        # process that has not initialised the GPU yet; the child's progress lines pass through
        import subprocess
        torch.cuda.empty_cache()
        proc = subprocess.run([sys.executable, os.path.abspath(__file__), "--config", "c5", "--no-cpu"],
                              stdout=subprocess.PIPE, text=True, timeout=1500)
        if proc.returncode != 0:
            raise RuntimeError(f"extended.c5 child exited with status {proc.returncode}")
        r = json.loads(proc.stdout.strip().splitlines()[-1])
        c5 = {"workload": f"{r['eval_grids']} held-out RGG swing/load grids N~U{{500..5000}}, Newton to residual 1e-10; "
                          f"network trained on {r['train_grids']} grids, {r['epochs']} epochs",
              "unit": "Newton iterations (mean) and device seconds",
              "flat_newton_mean": r["flat_newton_mean"], "learned_newton_mean": r["learned_newton_mean"],
              "newton_ratio_inexact_cg": r["newton_ratio_learned"],
              "newton_ratio_exact_cg": r["newton_ratio_learned_exact_cg"],
              "flat_s": r["flat_s"], "learned_pipeline_s": r["pipeline_total_s"],
              "time_ratio_inexact_cg": r["time_ratio_pipeline"],
              "flat_exact_cg_s": r["flat_exact_cg_s"], "learned_exact_cg_s": r["learned_total_exact_cg_s"],
              "time_ratio_exact_cg": r["time_ratio_learned_exact_cg"],
              "pipeline_fell_back": r["pipeline_fell_back"], "pipeline_accepted": r["pipeline_accepted"],
              "target_r2": r["target_r2"], "target_newton_ratio": 3.0}
    out = {
        "metric": "RHS evals/s + % HBM peak at N=1e6 swing",
        "value": value,
        "unit": "RHS evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Barabasi-Albert graph and parameters)",
        "config": {"workload": f"swing (2nd-order Kuramoto) RHS, Barabasi-Albert N={N}, m=2, E={E}, FP64",
                   "n": N, "e": E, "hub_degree": nw.tuning()["hub_degree"]},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/swing_ba_1e6_traffic.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
                     "algorithmic_bytes": algo_bytes, "kernel_us": per_rhs_s * 1e6,
                     "stream_copy_GBs": {"4GiB_moved": copy_hbm, "128MiB_moved": copy_small},
                     "neighbour_gather_torch": gather},
        "cpu_baseline": cpu,
        "extended": {"c4": c4, "c5": c5},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
