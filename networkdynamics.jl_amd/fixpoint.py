"""Batched synchronous fixpoints (find_fixpoint for Kuramoto-type power grids).

The reference solves `find_fixpoint(nw, x0, p)` (src/initialization.jl:73-144) with
NonlinearSolve's Newton and ForwardDiff Jacobians, one network at a time.  Here
`FixpointSolver(graphs)` runs damped Newton + conjugate-gradient linear solves for
many grids in one launch (nd_fixpoint_* in include/nd_hip.h), and `isfixpoint`
checks the result with the network right-hand side itself
(src/linear_analysis.jl:8-15: all |du| < tol).
"""
import ctypes as C

import numpy as np

from . import _abi as A
from . import graphs as G
from . import library as Lib
from .network import Network

STAT_DTYPE = np.dtype([("residual", "f8"), ("newton_iters", "i4"), ("cg_iters", "i4"), ("retcode", "i4"),
                       ("pad", "i4")])


class _Opts(C.Structure):
    _fields_ = [("tol", C.c_double), ("cg_rtol", C.c_double), ("max_newton", C.c_int32),
                ("cg_maxiter", C.c_int32), ("linearized", C.c_int32), ("inexact", C.c_int32), ("eta_max", C.c_double)]


class FixpointSolver:
    """Newton solver for F_i = P_i + sum_j K_ij sin(theta_j - theta_i) = 0 on many grids."""

    def __init__(self, graph_list):
        self.graphs = list(graph_list)
        self.ngrid = len(self.graphs)
        nv = np.array([g.nv() for g in self.graphs], np.int64)
        ne = np.array([g.ne() for g in self.graphs], np.int64)
        src = np.ascontiguousarray(np.concatenate([g.src for g in self.graphs]))
        dst = np.ascontiguousarray(np.concatenate([g.dst for g in self.graphs]))
        h = C.c_void_p()
        A.check(A.lib().nd_fixpoint_create(self.ngrid, nv.ctypes.data, ne.ctypes.data, src.ctypes.data,
                                           dst.ctypes.data, C.byref(h)))
        self._h = h
        self.nv, self.ne = nv, ne
        self.voff = np.r_[0, np.cumsum(nv)]
        self.eoff = np.r_[0, np.cumsum(ne)]

    def __del__(self):
        L = getattr(A, "_lib", None) if A is not None else None
        if L is not None and getattr(self, "_h", None):
            L.nd_fixpoint_destroy(self._h)
            self._h = None

    def solve(self, theta, P, scale, K, tol=1e-10, cg_rtol=1e-12, max_newton=50, cg_maxiter=20000, linearized=False,
              inexact=True, eta_max=1e-2):
        """theta (in/out), P, scale: concatenated per vertex; K per edge; float64 device tensors.
        inexact: loose CG tolerances while the Newton residual is large (forcing term
        min(eta_max, residual)), cg_rtol once it is small."""
        import torch
        for x, n in ((theta, self.voff[-1]), (P, self.voff[-1]), (scale, self.voff[-1]), (K, self.eoff[-1])):
            if x.dtype != torch.float64 or x.device.type != "cuda" or x.numel() != n:
                raise A.ArgumentError("theta/P/scale/K must be float64 device vectors of the right size")
        stats = torch.zeros(self.ngrid * STAT_DTYPE.itemsize // 8, dtype=torch.float64, device=theta.device)
        o = _Opts(float(tol), float(cg_rtol), int(max_newton), int(cg_maxiter), int(bool(linearized)),
                  int(bool(inexact)), float(eta_max))
        stream = torch.cuda.current_stream(theta.device).cuda_stream
        A.check(A.lib().nd_fixpoint_solve(self._h, theta.data_ptr(), P.data_ptr(), scale.data_ptr(), K.data_ptr(),
                                          C.byref(o), stats.data_ptr(), stream))
        return stats

    @staticmethod
    def stats_to_numpy(stats):
        return stats.cpu().numpy().view(STAT_DTYPE)


class SwingLoadGrid:
    """A power grid of swing generators and first-order loads joined by K*sin lines
    (ComponentLibrary's kuramoto_second / kuramoto_edge plus the load vertex)."""

    def __init__(self, g, is_gen, M, D, tau, P, K, xy=None):
        self.g, self.is_gen = g, np.asarray(is_gen, bool)
        self.M, self.D, self.tau, self.P, self.K = M, D, tau, P, K
        self.xy = xy
        self._nw = {}

    @property
    def scale(self):
        return np.where(self.is_gen, self.M, self.tau)

    def network(self, algebraic_loads=False):
        """The grid as a Network; algebraic_loads gives the loads mass 0 (a DAE with the
        same layout and the same fixpoints)."""
        key = bool(algebraic_loads)
        if key not in self._nw:
            swing = Lib.kuramoto_second()
            load = Lib.load_vertex(mass_matrix=0 if key else None)
            self._nw[key] = Network(self.g, [swing if x else load for x in self.is_gen], Lib.kuramoto_edge())
        return self._nw[key]

    def flat_p(self):
        """The network's flat parameter vector (vertex batches, then edges)."""
        nw = self.network()
        p = np.empty(nw.pdim)
        v, e = nw.ranges()
        pf = v[:, 2, 0] - 1
        gi = np.where(self.is_gen)[0]
        li = np.where(~self.is_gen)[0]
        p[pf[gi]], p[pf[gi] + 1], p[pf[gi] + 2] = self.M[gi], self.D[gi], self.P[gi]
        p[pf[li]], p[pf[li] + 1] = self.tau[li], self.P[li]
        p[e[:, 3, 0] - 1] = self.K
        return p

    def flat_u(self, theta, omega=None):
        nw = self.network()
        u = np.zeros(nw.dim)
        v, _ = nw.ranges()
        u[v[:, 0, 0] - 1] = theta
        if omega is not None:
            gi = np.where(self.is_gen)[0]
            u[v[gi, 0, 0]] = omega[gi]
        return u


def swing_load_grid(n, seed, mean_degree=6.0, gen_fraction=0.3, p_std=0.5, k_range=(5.0, 15.0)):
    """Random geometric grid in the unit square (largest connected component), 30 %
    swing generators, injections P ~ N(0, p_std) balanced to sum 0, K ~ U(k_range)."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components
    rng = np.random.default_rng(seed)
    g0, xy = G.random_geometric(n, np.sqrt(mean_degree / (np.pi * n)), seed=seed, return_points=True)
    A_ = coo_matrix((np.ones(g0.ne()), (g0.src - 1, g0.dst - 1)), shape=(n, n))
    _, lab = connected_components(A_, directed=False)
    big = np.argmax(np.bincount(lab))
    keep = np.where(lab == big)[0]
    new = -np.ones(n, np.int64)
    new[keep] = np.arange(len(keep))
    m = (new[g0.src - 1] >= 0) & (new[g0.dst - 1] >= 0)
    g = G.SimpleGraph._from_sorted(len(keep), new[g0.src[m] - 1] + 1, new[g0.dst[m] - 1] + 1)
    N = g.nv()
    is_gen = np.zeros(N, bool)
    is_gen[rng.choice(N, int(round(gen_fraction * N)), replace=False)] = True
    M = rng.uniform(1.0, 5.0, N)
    D = rng.uniform(0.1, 1.0, N)
    tau = rng.uniform(0.5, 2.0, N)
    P = rng.normal(0.0, p_std, N)
    P -= P.mean()
    K = rng.uniform(k_range[0], k_range[1], g.ne())
    return SwingLoadGrid(g, is_gen, M, D, tau, P, K, xy[keep])


def isfixpoint_residual(grid, theta, dev):
    """max |du| of the network RHS at (theta, omega = 0): the acceptance test."""
    import torch
    nw = grid.network()
    u = torch.from_numpy(grid.flat_u(theta)).to(dev)
    p = torch.from_numpy(grid.flat_p()).to(dev)
    du = torch.empty_like(u)
    nw(du, u, p, 0.0)
    return float(du.abs().max().item())
