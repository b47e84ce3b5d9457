"""Graphs with Graphs.jl's edge order, and seeded generators.

`SimpleGraph.edges()` yields (src, dst) with src <= dst in lexicographic order,
`SimpleDiGraph.edges()` yields (src, dst) lexicographically -- the order of
Graphs.jl's `edges(g)`, which fixes the reference's edge registration order
(IndexManager.edgevec, src/network_structure.jl:37).
"""
import ctypes as C

import numpy as np

from . import _abi as A


class SimpleGraph:
    directed = False

    def __init__(self, nv, src=(), dst=()):
        self._nv = int(nv)
        s = np.asarray(src, dtype=np.int64).reshape(-1)
        d = np.asarray(dst, dtype=np.int64).reshape(-1)
        if len(s) != len(d):
            raise A.ArgumentError("src and dst lengths differ")
        if len(s) and (min(s.min(), d.min()) < 1 or max(s.max(), d.max()) > self._nv):
            raise A.ArgumentError("vertex index outside 1:nv")
        if not self.directed:
            s, d = np.minimum(s, d), np.maximum(s, d)
        key = s * (self._nv + 1) + d
        key, first = np.unique(key, return_index=True)
        self.src = np.ascontiguousarray(s[first])
        self.dst = np.ascontiguousarray(d[first])

    @classmethod
    def _from_sorted(cls, nv, src, dst):
        g = cls.__new__(cls)
        g._nv = int(nv)
        g.src = np.ascontiguousarray(src, dtype=np.int64)
        g.dst = np.ascontiguousarray(dst, dtype=np.int64)
        return g

    def nv(self):
        return self._nv

    def ne(self):
        return len(self.src)

    def edges(self):
        return list(zip(self.src.tolist(), self.dst.tolist()))

    def degree(self):
        deg = np.bincount(self.src - 1, minlength=self._nv) + np.bincount(self.dst - 1, minlength=self._nv)
        return deg

    def __repr__(self):
        return f"{{{self._nv}, {self.ne()}}} {'directed' if self.directed else 'undirected'} simple Int64 graph"


class SimpleDiGraph(SimpleGraph):
    directed = True


def complete_graph(n):
    s, d = np.triu_indices(n, 1)
    return SimpleGraph._from_sorted(n, s + 1, d + 1)


def star_digraph(n):
    """Edges 1 -> i for i in 2:n (Graphs.jl star_digraph)."""
    return SimpleDiGraph(n, np.ones(n - 1, np.int64), np.arange(2, n + 1))


def path_graph(n):
    return SimpleGraph(n, np.arange(1, n), np.arange(2, n + 1))


def cycle_graph(n):
    return SimpleGraph(n, np.arange(1, n + 1), np.r_[np.arange(2, n + 1), 1])


def _take(el):
    L = A.lib()
    n = L.nd_edgelist_size(el)
    s = np.empty(n, np.int64)
    d = np.empty(n, np.int64)
    A.check(L.nd_edgelist_copy(el, s.ctypes.data, d.ctypes.data))
    L.nd_edgelist_free(el)
    return s, d


def erdos_renyi(n, m, seed=0):
    """G(n, m): m distinct undirected edges drawn uniformly."""
    el = C.c_void_p()
    A.check(A.lib().nd_graph_erdos_renyi(n, m, seed, C.byref(el)))
    return SimpleGraph._from_sorted(n, *_take(el))


def barabasi_albert(n, m, seed=0):
    """Preferential attachment: m initial isolated vertices, m edges per new vertex."""
    el = C.c_void_p()
    A.check(A.lib().nd_graph_barabasi_albert(n, m, seed, C.byref(el)))
    return SimpleGraph._from_sorted(n, *_take(el))


def random_geometric(n, radius, seed=0, return_points=False):
    """n uniform points in the unit square; edges between points closer than radius."""
    el = C.c_void_p()
    xy = np.empty((n, 2), np.float64)
    A.check(A.lib().nd_graph_random_geometric(n, radius, seed, xy.ctypes.data, C.byref(el)))
    g = SimpleGraph._from_sorted(n, *_take(el))
    return (g, xy) if return_points else g


def watts_strogatz(n, k, beta, seed=0):
    """Ring lattice with k neighbours, each edge rewired with probability beta."""
    el = C.c_void_p()
    A.check(A.lib().nd_graph_watts_strogatz(n, k, beta, seed, C.byref(el)))
    return SimpleGraph._from_sorted(n, *_take(el))
