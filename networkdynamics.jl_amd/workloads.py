"""Synthetic workloads (seeded): graph, component models, parameters, initial state.

kuramoto_er        first-order Kuramoto on G(n, m), omega ~ N(0,1), K = 3, theta0 ~ U(-pi, pi)
swing_ba           second-order Kuramoto (swing) on Barabasi-Albert(m=2): M ~ U(1,5),
                   D ~ U(0.1,1), P ~ N(0,1) re-centred to sum 0, K ~ U(5,10)
heterogeneous_rgg  random geometric graph (mean degree 4), 30 % swing / 70 % first-order
                   load vertices (tau ~ U(0.5,2)), static lines B ~ U(1,20), phi ~ U(0,0.1)

Parameter layout follows the network's flat p (vertex batches first, then edge
batches, each in component order), read from the network's own batches.
"""
import numpy as np

from . import graphs
from . import library as Lib
from .network import Network


def _fill_batch_params(nw, rng, vfill, efill):
    p = np.empty(nw.pdim)
    for is_edge, fill in ((False, vfill), (True, efill)):
        for b in nw.batches(is_edge):
            k = nw.kinds[b["kind"]]
            n, pd = b["count"], k["pdim"]
            if pd == 0:
                continue
            blk = fill(k["model"], n, rng, b["indices"])
            p[b["param_first"] - 1: b["param_first"] - 1 + n * pd] = np.asarray(blk).reshape(n, pd).reshape(-1)
    return p


def kuramoto_er(n=1000, m=4000, K=3.0, seed=1):
    g = graphs.erdos_renyi(n, m, seed=seed)
    nw = Network(g, Lib.kuramoto_first(), Lib.kuramoto_edge())
    rng = np.random.default_rng(seed)
    omega = rng.normal(size=n)
    p = np.concatenate([omega, np.full(g.ne(), K)])
    u0 = rng.uniform(-np.pi, np.pi, n)
    return nw, p, u0


def swing_params(N, E, rng):
    M = rng.uniform(1.0, 5.0, N)
    D = rng.uniform(0.1, 1.0, N)
    P = rng.normal(size=N)
    P -= P.mean()
    K = rng.uniform(5.0, 10.0, E)
    return M, D, P, K


def swing_ba(n=1_000_000, m=2, seed=1, graph=None):
    g = graph if graph is not None else graphs.barabasi_albert(n, m, seed=seed)
    n = g.nv()
    nw = Network(g, Lib.kuramoto_second(), Lib.kuramoto_edge())
    rng = np.random.default_rng(seed)
    M, D, P, K = swing_params(n, g.ne(), rng)
    p = np.empty(nw.pdim)
    p[0:3 * n:3], p[1:3 * n:3], p[2:3 * n:3] = M, D, P
    p[3 * n:] = K
    u0 = np.empty(2 * n)
    u0[0::2] = rng.uniform(-np.pi, np.pi, n)
    u0[1::2] = rng.normal(size=n)
    return nw, p, u0


def heterogeneous_rgg(n=200_000, seed=1, mean_degree=4.0, gen_fraction=0.3):
    r = np.sqrt(mean_degree / (np.pi * n))
    g = graphs.random_geometric(n, r, seed=seed)
    rng = np.random.default_rng(seed)
    is_gen = np.zeros(n, bool)
    is_gen[rng.choice(n, int(round(gen_fraction * n)), replace=False)] = True
    swing, load = Lib.kuramoto_second(), Lib.load_vertex()
    vm = [swing if x else load for x in is_gen]
    nw = Network(g, vm, Lib.static_line_edge())
    P = rng.normal(size=n)
    P -= P.mean()

    def vfill(model, cnt, rng_, idx):
        if model == swing.model:
            return np.stack([rng_.uniform(1, 5, cnt), rng_.uniform(0.1, 1, cnt), P[idx - 1]], axis=1)
        return np.stack([rng_.uniform(0.5, 2.0, cnt), P[idx - 1]], axis=1)

    def efill(model, cnt, rng_, idx):
        return np.stack([rng_.uniform(1, 20, cnt), rng_.uniform(0, 0.1, cnt)], axis=1)

    p = _fill_batch_params(nw, rng, vfill, efill)
    v, _ = nw.ranges()
    # states: theta ~ U(-pi, pi) for every vertex, omega ~ N(0,1) for generators
    u0 = rng.normal(size=nw.dim)
    theta_idx = v[:, 0, 0] - 1
    u0[theta_idx] = rng.uniform(-np.pi, np.pi, n)
    return nw, p, u0


def algorithmic_bytes_swing(n, e):
    """Compulsory bytes of one swing RHS: u (16N) + vertex p (24N) + K (8E) + int32 src/dst (8E) + du (16N)."""
    return 56 * n + 16 * e
