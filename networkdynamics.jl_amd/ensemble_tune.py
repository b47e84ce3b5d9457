"""Autotuner for the ensemble kernel's size classes.

A size class is a launch shape: `threads` per workgroup (the tile) and vertices per
thread `vper` (the segment each thread owns), holding every instance with
nv <= threads * vper that no earlier class holds.  The tuner
  1. splits the ensemble's instances into size bins (quantiles of nv);
  2. for every bin and every launch shape that fits the bin's largest instance, runs
     a short solve of a sample of that bin (replicated to fill the GPU) and measures
     the time per instance and right-hand-side evaluation;
  3. picks the list of at most six shapes (increasing capacity) that minimises the
     predicted total time, each bin going to the first shape that holds it.
Results are cached in a JSON file keyed by device, models, method and bins.
"""
import itertools
import json
import os

import numpy as np

from .ensemble import NCLASS, TSIT5, Ensemble

SHAPES = [(t, v) for t in (64, 128, 192, 256, 320, 384, 448, 512) for v in (1, 2, 3, 4)]


def _bins(nv, nbins):
    qs = np.unique(np.quantile(nv, np.linspace(0, 1, nbins + 1)[1:], method="higher").astype(np.int64))
    idx = np.searchsorted(qs, nv)               # bin of each instance: first edge >= nv
    return qs, idx


def choose(bin_max, bin_weight, cost):
    """Best list of <= NCLASS shapes with increasing capacity.  bin_max: increasing bin
    upper sizes; cost[b][shape] = seconds per instance-evaluation (missing = does not fit).
    A bin goes to the first listed shape whose capacity holds it, so a list splits the
    bins into consecutive groups; dynamic programming over (shapes used, last shape).
    Returns (shapes, predicted cost)."""
    shapes = sorted({s for c in cost for s in c}, key=lambda s: (s[0] * s[1], s))
    caps = [t * v for t, v in shapes]
    nb = len(bin_max)

    def group_cost(lo_cap, j):
        tot = 0.0
        for b in range(nb):
            if lo_cap < bin_max[b] <= caps[j]:
                c = cost[b].get(shapes[j])
                if c is None:
                    return float("inf")
                tot += bin_weight[b] * c
        return tot

    inf = float("inf")
    f = [[(inf, None)] * len(shapes) for _ in range(NCLASS + 1)]
    for j in range(len(shapes)):
        f[1][j] = (group_cost(-1, j), None)
    for k in range(2, NCLASS + 1):
        for j in range(len(shapes)):
            best = (inf, None)
            for i in range(len(shapes)):
                if caps[i] < caps[j] and f[k - 1][i][0] < inf:
                    c = f[k - 1][i][0] + group_cost(caps[i], j)
                    if c < best[0]:
                        best = (c, i)
            f[k][j] = best
    best = (inf, None, None)
    for k in range(1, NCLASS + 1):
        for j in range(len(shapes)):
            if caps[j] >= bin_max[-1] and f[k][j][0] < best[0]:
                best = (f[k][j][0], k, j)
    if best[1] is None:
        return None, inf
    out, k, j = [], best[1], best[2]
    while j is not None:
        out.append(shapes[j])
        j = f[k][j][1]
        k -= 1
    return out[::-1], best[0]


def _signature(ens, method, reltol, qs):
    import torch
    return json.dumps(dict(device=torch.cuda.get_device_name(), vertex=ens.vertexm.name, edge=ens.edgem.name,
                           method=int(method), reltol=float(reltol), bins=[int(x) for x in qs]), sort_keys=True)


def autotune(ens, u, p, probe_t=0.5, method=TSIT5, reltol=1e-8, abstol=1e-8, nbins=6, fill=1024,
             cache=None, apply=True):
    """Tunes `ens`'s classes on its own data u, p (device, not modified).  probe_t: length
    of the short solves.  fill: instances per timed run (a bin's sample, repeated)."""
    import torch
    qs, idx = _bins(ens.nv, nbins)
    key = _signature(ens, method, reltol, qs)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            shapes = [tuple(s) for s in db[key]["classes"]]
            if apply:
                ens.set_classes(shapes)
            return dict(db[key], cached=True)
    uh, ph = u.cpu().numpy(), p.cpu().numpy()
    rng = np.random.default_rng(0)
    cost, weight = [], []
    for b, m in enumerate(qs):
        members = np.flatnonzero(idx == b)
        pick = rng.choice(members, size=min(len(members), 64), replace=False)
        pick = np.resize(pick, fill)
        sub = Ensemble([ens.graphs[i] for i in pick], ens.vertexm, ens.edgem)
        su = torch.from_numpy(np.concatenate([uh[ens.state_slice(i)] for i in pick])).to(u.device)
        sp = torch.from_numpy(np.concatenate([ph[ens.param_slice(i)] for i in pick])).to(u.device)
        work = torch.empty_like(su)
        cb = {}
        for shape in SHAPES:
            if shape[0] * shape[1] < m:
                continue
            sub.set_classes([shape])
            times = []
            for rep in range(2):
                work.copy_(su)
                torch.cuda.synchronize()
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record()
                st = sub.solve(work, sp, (0.0, probe_t), method=method, reltol=reltol, abstol=abstol)
                e1.record()
                torch.cuda.synchronize()
                times.append(e0.elapsed_time(e1) * 1e-3)
            nf = Ensemble.stats_to_numpy(st)["nf"].sum()
            cb[shape] = min(times) / max(nf, 1)
        cost.append(cb)
        nf_mean = float(nf) / fill
        weight.append(float(len(members)) * nf_mean)
    shapes, pred = choose(list(qs), weight, cost)
    out = dict(classes=[list(s) for s in shapes], predicted_s=pred, bins=[int(x) for x in qs],
               cost_ns={f"{int(m)}": {f"{t}x{v}": c * 1e9 for (t, v), c in cb.items()} for m, cb in zip(qs, cost)},
               cached=False)
    if apply:
        ens.set_classes(shapes)
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = {k: v for k, v in out.items() if k != "cached"}
        os.makedirs(os.path.dirname(os.path.abspath(cache)), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1, sort_keys=True)
    return out
