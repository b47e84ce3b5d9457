"""Learned warm start for fixpoint solves.

At a fixpoint theta* of  P_i + sum_j K_ij sin(theta_j - theta_i) = 0  the weighted
Laplacian L_K satisfies  L_K theta* = P + s*,  with the per-vertex source
    s_i = sum_j K_ij (sin(D_ij) - D_ij),   D_ij = theta_j - theta_i.
So theta* is one linear solve away once s* is known.  The source is a sum over the
vertex's own lines, small (cubic in the angle differences) and much more local than
theta* itself, whose correction from the DC solution spreads over the whole grid.

Pipeline for a batch of swing/load grids:
  1. DC power flow theta_dc: L_K theta_dc = P (`FixpointSolver(linearized=True)`);
  2. physics source s0 = s(theta_dc) and per-vertex features;
  3. a 4-layer message-passing network (hidden 64, FP32, nd_mp_* kernels)
     predicts the remaining source correction t = s* - s0;
  4. warm start: L_K theta_w = P + s0 + t_hat (a second linear solve);
  5. Newton (`FixpointSolver`) from theta_w; its residual <= 1e-10 is the acceptance test.
Training targets come from fixpoints the batched Newton solver produced from flat
starts.  The method without the network is Newton from theta = 0; steps 1, 2, 4
with t_hat = 0 are the physics-only ablation.
"""
import ctypes as C
import time

import numpy as np

from . import _abi as A
from .fixpoint import FixpointSolver

NFEAT = 8


class GraphBatch:
    """Stacked vertices of several grids with the row-normalised K-weighted adjacency."""

    def __init__(self, grids, dev):
        import scipy.sparse as sp
        import torch
        self.grids = list(grids)
        nv = np.array([gr.g.nv() for gr in self.grids], np.int64)
        self.voff = np.r_[0, np.cumsum(nv)]
        n = int(self.voff[-1])
        ne = np.array([gr.g.ne() for gr in self.grids], np.int64)
        shift = np.repeat(self.voff[:-1], ne)
        s = np.concatenate([gr.g.src for gr in self.grids]) - 1 + shift
        d = np.concatenate([gr.g.dst for gr in self.grids]) - 1 + shift
        K = np.concatenate([gr.K for gr in self.grids])
        # CSR of the symmetric K-weighted adjacency (scipy's O(nnz) COO -> CSR conversion)
        A_ = sp.csr_matrix((np.r_[K, K], (np.r_[s, d], np.r_[d, s])), shape=(n, n))
        rowptr = A_.indptr.astype(np.int32)
        cols = A_.indices
        w = A_.data
        deg_w = np.add.reduceat(w, rowptr[:-1])            # every vertex of a grid's largest component has a line
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        a_fwd = (w / deg_w[rows]).astype(np.float32)
        a_bwd = (w / deg_w[cols]).astype(np.float32)      # weight of the transposed entry (col -> row)
        self.n = n
        self.rowptr = torch.from_numpy(rowptr).to(dev)
        self.nbr = torch.from_numpy(cols.astype(np.int32)).to(dev)
        self.a_fwd = torch.from_numpy(a_fwd).to(dev)
        self.a_bwd = torch.from_numpy(a_bwd).to(dev)
        self.host = dict(rowptr=rowptr, nbr=cols, a_fwd=a_fwd, a_bwd=a_bwd)
        self.w = np.ascontiguousarray(w, np.float64)


def source(grid, theta):
    """s_i = sum_j K_ij (sin(D_ij) - D_ij), D_ij = theta_j - theta_i (per vertex)."""
    g = grid.g
    s, d = g.src - 1, g.dst - 1
    D = theta[d] - theta[s]
    q = grid.K * (np.sin(D) - D)
    n = g.nv()
    return np.bincount(s, weights=q, minlength=n) - np.bincount(d, weights=q, minlength=n)


def features(grid, theta_dc, s0=None):
    """Per-vertex inputs: P, generator flag, log weighted degree, degree, the physics
    source s0 and its K-weighted neighbour mean, sum K D^2 and max |D| of the DC flows.
    A vertex without lines gets 0 for the log degree and the neighbour mean."""
    g = grid.g
    s, d = g.src - 1, g.dst - 1
    n = g.nv()
    x = theta_dc[d] - theta_dc[s]
    K = grid.K
    if s0 is None:
        s0 = source(grid, theta_dc)
    kdeg = np.bincount(s, weights=K, minlength=n) + np.bincount(d, weights=K, minlength=n)
    deg = np.bincount(s, minlength=n) + np.bincount(d, minlength=n)
    ns0 = np.bincount(s, weights=K * s0[d], minlength=n) + np.bincount(d, weights=K * s0[s], minlength=n)
    ns0 = np.divide(ns0, kdeg, out=np.zeros(n), where=kdeg > 0)
    sq = np.bincount(s, weights=K * x * x, minlength=n) + np.bincount(d, weights=K * x * x, minlength=n)
    mx = np.zeros(n)
    np.maximum.at(mx, s, np.abs(x))
    np.maximum.at(mx, d, np.abs(x))
    return np.stack([grid.P, grid.is_gen.astype(float), np.log(kdeg, out=np.zeros(n), where=kdeg > 0), deg / 6.0, s0, ns0, sq, mx], axis=1)


class WarmStartNet:
    def __init__(self, dev, seed=0):
        import torch
        self.dev = dev
        npar = A.lib().nd_mp_param_count(NFEAT)
        rng = np.random.default_rng(seed)
        p = np.zeros(npar, np.float32)
        k = 0

        def glorot(fan_in, fan_out):
            return rng.normal(0, np.sqrt(2.0 / (fan_in + fan_out)), (fan_in, fan_out)).astype(np.float32)

        p[k:k + NFEAT * 64] = glorot(NFEAT, 64).ravel(); k += NFEAT * 64 + 64
        for _ in range(4):
            p[k:k + 64 * 128] = np.concatenate([glorot(64, 64), glorot(64, 64)], axis=1).ravel(); k += 64 * 128 + 64
        p[k:k + 64] = glorot(64, 1).ravel(); k += 65
        assert k == npar
        self.params = torch.from_numpy(p).to(dev)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grads = torch.zeros_like(self.params)
        self.step = 0
        self.xmean = np.zeros(NFEAT)
        self.xstd = np.ones(NFEAT)
        self.tscale = 1.0
        self._work = None

    def _workspace(self, n):
        import torch
        need = int(A.lib().nd_mp_workspace_floats(n))
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.float32, device=self.dev)
        return self._work

    def normalise(self, X):
        import torch
        Xn = (X - self.xmean) / self.xstd
        return torch.from_numpy(np.ascontiguousarray(Xn, np.float32)).to(self.dev)

    def forward(self, batch, Xd):
        import torch
        out = torch.empty(batch.n, dtype=torch.float32, device=self.dev)
        work = self._workspace(batch.n)
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        A.check(A.lib().nd_mp_forward(batch.n, batch.rowptr.data_ptr(), batch.nbr.data_ptr(), batch.a_fwd.data_ptr(),
                                      Xd.data_ptr(), NFEAT, self.params.data_ptr(), work.data_ptr(), out.data_ptr(),
                                      stream))
        return out

    def gradient(self, batch, Xd, target):
        """Forward + backward; returns the loss (device scalar) and fills self.grads."""
        import torch
        self.forward(batch, Xd)
        lp = torch.zeros(batch.n // 4 + 2, dtype=torch.float32, device=self.dev)
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        A.check(A.lib().nd_mp_backward(batch.n, batch.rowptr.data_ptr(), batch.nbr.data_ptr(), batch.a_bwd.data_ptr(),
                                       Xd.data_ptr(), NFEAT, self.params.data_ptr(), target.data_ptr(),
                                       self._workspace(batch.n).data_ptr(), self.grads.data_ptr(), lp.data_ptr(),
                                       stream))
        return lp.sum() / batch.n

    def train_step(self, batch, Xd, target, lr=1e-3):
        import torch
        loss = self.gradient(batch, Xd, target)
        self.step += 1
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        A.check(A.lib().nd_mp_adam(self.params.numel(), self.params.data_ptr(), self.grads.data_ptr(),
                                   self.m.data_ptr(), self.v.data_ptr(), float(lr), 0.9, 0.999, 1e-8, self.step,
                                   stream))
        return loss

    def save(self, path):
        np.savez(path, params=self.params.cpu().numpy(), xmean=self.xmean, xstd=self.xstd, tscale=self.tscale)

    def load(self, path):
        import torch
        z = np.load(path)
        self.params = torch.from_numpy(z["params"]).to(self.dev)
        self.xmean, self.xstd, self.tscale = z["xmean"], z["xstd"], float(z["tscale"])


def _concat(grids, attr):
    return np.concatenate([getattr(gr, attr) if not callable(getattr(gr, attr)) else getattr(gr, attr)()
                           for gr in grids])


def solve_grids(grids, dev, theta0=None, linearized=False, P=None, **kw):
    """Batched Newton (or, linearized, L_K theta = P) solve of several grids; returns
    (theta per grid, stats, seconds).  P: per-grid source override (linearized solves)."""
    import torch
    fs = FixpointSolver([gr.g for gr in grids])
    n = int(fs.voff[-1])
    th = torch.zeros(n, dtype=torch.float64, device=dev) if theta0 is None else \
        torch.from_numpy(np.ascontiguousarray(theta0, np.float64)).to(dev)
    Pd = torch.from_numpy(np.concatenate([gr.P for gr in grids] if P is None else P)).to(dev)
    S = torch.from_numpy(np.concatenate([gr.scale for gr in grids])).to(dev)
    K = torch.from_numpy(np.concatenate([gr.K for gr in grids])).to(dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = fs.solve(th, Pd, S, K, linearized=linearized, **kw)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    t = th.cpu().numpy()
    return [t[fs.voff[i]:fs.voff[i + 1]] for i in range(len(grids))], FixpointSolver.stats_to_numpy(st), el


def center(x):
    return x - x.mean()


def physics_inputs(grids, dev):
    """DC solutions, physics sources s0 and features of a batch of grids."""
    dc, _, t_dc = solve_grids(grids, dev, linearized=True)
    dc = [center(t) for t in dc]
    s0 = [source(gr, t) for gr, t in zip(grids, dc)]
    X = [features(gr, t, s) for gr, t, s in zip(grids, dc, s0)]
    return dc, s0, X, t_dc


def predict(net, grids, X, dev):
    """Network source corrections t_hat per grid (host float64), each centred: a source
    must sum to zero over a grid for L_K theta = P + s0 + t_hat to have a solution."""
    batch = GraphBatch(grids, dev)
    out = net.forward(batch, net.normalise(np.concatenate(X))).double().cpu().numpy() / net.tscale
    return [center(out[batch.voff[i]:batch.voff[i + 1]]) for i in range(len(grids))]


def warm_start(grids, s0, t_hat, dev):
    """theta_w from L_K theta_w = P + s0 + t_hat (t_hat None: physics only); (theta list, seconds)."""
    src = [gr.P + a + (0.0 if t_hat is None else b) for gr, a, b in
           zip(grids, s0, t_hat if t_hat is not None else [None] * len(grids))]
    th, _, el = solve_grids(grids, dev, linearized=True, P=src)
    return [center(t) for t in th], el


class GridBatch:
    """Device-resident data of a set of grids for repeated solves: the Newton solver's
    topology, the network's CSR (GraphBatch), and P, scale, K.  Built once per grid set;
    every per-solve step below runs on the GPU."""

    def __init__(self, grids, dev):
        import torch
        self.grids = list(grids)
        self.dev = dev
        self.fs = FixpointSolver([gr.g for gr in self.grids])
        self.graph = GraphBatch(self.grids, dev)
        self.n = self.graph.n
        self.voff = self.graph.voff
        self.w = torch.from_numpy(self.graph.w).to(dev)                     # K of each CSR entry
        self.glen = torch.from_numpy(np.diff(self.voff).astype(np.int64)).to(dev)
        cat = lambda a: torch.from_numpy(np.ascontiguousarray(np.concatenate(a))).to(dev)
        self.P = cat([gr.P for gr in self.grids])
        self.S = cat([gr.scale for gr in self.grids])
        self.K = cat([gr.K for gr in self.grids])
        self.is_gen = cat([gr.is_gen.astype(float) for gr in self.grids])
        # the residual of the flat start theta = 0 (F = P) per grid
        self.flat_residual = np.array([np.max(np.abs(gr.P / gr.scale)) for gr in self.grids])

    def start_residual(self, theta):
        """Per grid max_i |F_i / scale_i| at theta (nd_ws_residual and a per-grid max): the
        residual the Newton solver starts from.  Host array; NaN where theta gives a NaN."""
        import torch
        th = theta.contiguous()
        r = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        A.check(A.lib().nd_ws_residual(self.n, self.graph.rowptr.data_ptr(), self.graph.nbr.data_ptr(),
                                       self.w.data_ptr(), th.data_ptr(), self.P.data_ptr(), self.S.data_ptr(),
                                       r.data_ptr(), stream))
        return torch.segment_reduce(r, "max", lengths=self.glen).cpu().numpy()

    def solve(self, theta0=None, P=None, linearized=False, **kw):
        """Batched Newton / linear solve; theta0, P: device tensors (defaults 0, the grids' P).
        Returns (theta device tensor, stats, seconds)."""
        import torch
        th = torch.zeros(self.n, dtype=torch.float64, device=self.dev) if theta0 is None else theta0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st = self.fs.solve(th, self.P if P is None else P, self.S, self.K, linearized=linearized, **kw)
        torch.cuda.synchronize()
        return th, FixpointSolver.stats_to_numpy(st), time.perf_counter() - t0

    def center(self, x):
        """x minus its mean over each grid."""
        import torch
        mean = torch.segment_reduce(x, "sum", lengths=self.glen) / self.glen.to(x.dtype)
        return x - torch.repeat_interleave(mean, self.glen)

    def source(self, theta):
        """s_i = sum_j K_ij (sin(D_ij) - D_ij) per vertex (the s0 output of nd_ws_features)."""
        return self.features(theta)[0]

    def features(self, theta_dc):
        """(s0, X): the physics source and the per-vertex inputs of `features`, on the GPU
        (nd_ws_features: two kernels, sums in CSR order)."""
        import torch
        th = theta_dc.contiguous()
        s0 = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        X = torch.empty((self.n, 8), dtype=torch.float64, device=self.dev)
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        A.check(A.lib().nd_ws_features(self.n, self.graph.rowptr.data_ptr(), self.graph.nbr.data_ptr(),
                                       self.w.data_ptr(), th.data_ptr(), self.P.data_ptr(), self.is_gen.data_ptr(),
                                       s0.data_ptr(), X.data_ptr(), stream))
        return s0, X


def warm_start_device(gb, net=None, tol=1e-6, dc_tol=None, warm_eta=1e-2):
    """Warm start of every grid in a GridBatch, entirely on the GPU: DC solve, physics
    source, network correction (net None: physics only), second linear solve from the DC
    solution.  The linear solves stop at residual dc_tol (default tol) and tol: they only
    provide a starting point, Newton reaches the 1e-10 residual from it.  warm_eta caps
    the second solve's CG forcing term (the solver's eta_max).  Returns (theta_w, dict of
    seconds)."""
    import torch
    times = {}
    th_dc, _, times["dc_solve"] = gb.solve(linearized=True, tol=tol if dc_tol is None else dc_tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    th_dc = gb.center(th_dc)
    s0, X = gb.features(th_dc)
    src = gb.P + s0
    if net is not None:
        xm = torch.from_numpy(net.xmean).to(gb.dev)
        xs = torch.from_numpy(net.xstd).to(gb.dev)
        Xn = ((X - xm) / xs).to(torch.float32).contiguous()
        out = net.forward(gb.graph, Xn).double() / net.tscale
        src = src + gb.center(out)
    torch.cuda.synchronize()
    times["features_and_network"] = time.perf_counter() - t0
    th_w, _, times["warm_solve"] = gb.solve(theta0=th_dc, P=src, linearized=True, tol=tol, eta_max=warm_eta)
    return gb.center(th_w), times


def solve_with_fallback(gb, theta0, tol=1e-10, first_max_newton=None, times=None, **kw):
    """Newton on every grid of a GridBatch from the starts theta0 (a device vector, e.g.
    the learned warm starts), then Newton from the flat start theta = 0 -- the method
    without the network -- on every grid whose first solve did not end with a residual
    <= tol (no convergence, a stalled line search, a NaN).  Grids the flat start does not
    bring to tol either raise NetworkInitError, as find_fixpoint does on a failed solve
    (src/initialization.jl:127-141).

    first_max_newton caps the Newton steps of the first solve (a good start converges in a
    few quadratic steps; a bad one then falls back early instead of holding the batch for
    the full budget); the flat solve keeps max_newton.  The second solve runs the whole
    batch with theta = 0 on the failed grids only: the other grids start at their
    solutions, already at tol, and return unchanged without a Newton step.  Returns
    (theta, stats per grid -- the accepted solve's --, fell_back per grid); times, a dict,
    receives the seconds of both solves."""
    import torch
    from .initialization import NetworkInitError
    first = dict(kw)
    if first_max_newton is not None:
        first["max_newton"] = int(first_max_newton)
    th, st, t1 = gb.solve(theta0=theta0, tol=tol, **first)

    def bad(s):
        return (s["retcode"] != 0) | ~(s["residual"] <= tol)

    failed = bad(st)
    t2 = 0.0
    if failed.any():
        mask = torch.from_numpy(np.repeat(failed, np.diff(gb.voff))).to(gb.dev)
        th2, st2, t2 = gb.solve(theta0=torch.where(mask, torch.zeros_like(th), th), tol=tol, **kw)
        still = failed & bad(st2)
        if still.any():
            idx = np.flatnonzero(still)
            raise NetworkInitError(
                f"Initialization failed: {len(idx)} grid(s) reach no fixpoint with residual <= {tol:g} from the "
                f"learned start or from the flat start (grid indices {idx[:10].tolist()}"
                f"{' ...' if len(idx) > 10 else ''}; flat-start retcodes {st2['retcode'][idx[:10]].tolist()})")
        th = torch.where(mask, th2, th)
        st = st.copy()
        st[failed] = st2[failed]
    if times is not None:
        times.update(newton=t1, fallback_newton=t2)
    return th, st, failed


def learned_solve(gb, net, tol=1e-10, warm_tol=1e-1, dc_tol=3e-1, warm_eta=0.3, first_max_newton=20, **kw):
    """Fixpoints of every grid of a GridBatch from the learned warm start, with the flat
    start as the fallback (solve_with_fallback).  net None: the physics-only start.
    The two linear solves of the warm start stop at residuals 0.3 (DC) and 0.1 (warm
    start) by default, one inexact Newton step each: the features need a rough DC
    solution only, and Newton takes the start to tol anyway.  Tighter pre-solves leave
    the Newton iterations from the start nearly unchanged (2.97-3.09 on 884 held-out
    grids from (1e-2, 1e-2) to (0.3, 0.1)) and cost 20 ms more device time.  The
    warm-start solve's CG stops at a relative 0.3 instead of 1e-2 (warm_eta): 9 ms less,
    with a Newton mean from the start of 3.12 instead of 3.08; the DC solve keeps 1e-2,
    since the features are computed from it.  The first solve takes up to 20 Newton
    steps, as many as the flat start needs on its hardest held-out grids: a start that
    lands far from the fixpoint can need 14, and a grid the cap stops falls back to a
    whole second solve from flat.  The start's residual (GridBatch.start_residual) does
    not tell such starts apart: on the held-out grids a learned start above the flat
    start's residual needs 4 steps where flat needs 19.
    Returns (theta, stats, fell_back, dict of seconds)."""
    theta_w, times = warm_start_device(gb, net, tol=warm_tol, dc_tol=dc_tol, warm_eta=warm_eta)
    th, st, fell = solve_with_fallback(gb, theta_w, tol=tol, first_max_newton=first_max_newton, times=times, **kw)
    return th, st, fell, times
