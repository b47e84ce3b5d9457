"""Jacobian sparsity and colored sparse Jacobians on the device JVP.

The reference keeps an optional `jac_prototype` (src/network_structure.jl:101,
`set_jac_prototype!` :344-357) that `get_jac_prototype`
(ext/NetworkDynamicsSparsityExt.jl:56-135) fills by tracing the RHS with
SparseConnectivityTracer; implicit solvers and `find_fixpoint` then color it and take
one ForwardDiff pass per color.  Here the pattern is read off the network structure
instead of traced, so it is structural: every state of a component depends on all of
its own states, its aggregated inputs depend on the output states of both endpoints of
each incident edge (or on the ODE edge's output state), and external inputs add their
source states.  It holds every entry a trace would find, plus possibly some that
vanish numerically.  Columns are colored greedily so that no two columns of one color
share a row; `sparse_jacobian` then runs one device JVP per color and scatters each
compressed column back.
"""
import numpy as np

from . import _abi as A


def _output_states(nw):
    """u-index (0-based) of every output-buffer entry that is a StateMask of a state, -1
    elsewhere (feed-forward edge outputs)."""
    st = np.full(int(nw.info.lastidx_out), -1, np.int64)
    for b in nw.batches(False):
        k = nw.kinds[b["kind"]]
        q = np.arange(b["count"])
        for j, m in enumerate(k["mask_dst"]):
            st[b["out_first"] - 1 + q * k["outdim_dst"] + j] = b["state_first"] - 1 + q * k["dim"] + m - 1
    for b in nw.batches(True):
        k = nw.kinds[b["kind"]]
        if k["ff"] != A.ND_FF_PURESTATEMAP:
            continue
        q = np.arange(b["count"])
        stride = k["outdim_src"] + k["outdim_dst"]
        for j, m in enumerate(tuple(k["mask_src"]) + tuple(k["mask_dst"])):
            st[b["out_first"] - 1 + q * stride + j] = b["state_first"] - 1 + q * k["dim"] + m - 1
    return st


def _output_edges(nw):
    """edge id (0-based) owning every output-buffer entry, -1 for vertex outputs."""
    own = np.full(int(nw.info.lastidx_out), -1, np.int64)
    for b in nw.batches(True):
        k = nw.kinds[b["kind"]]
        stride = k["outdim_src"] + k["outdim_dst"]
        if stride == 0:
            continue
        q = np.arange(b["count"])
        for j in range(stride):
            own[b["out_first"] - 1 + q * stride + j] = b["indices"] - 1
    return own


def jac_prototype(nw):
    """Structural sparsity of d(du)/du as a scipy CSR matrix of booleans (dim x dim)."""
    from scipy.sparse import coo_matrix
    v, e = nw.ranges()
    ost = _output_states(nw)
    oedge = _output_edges(nw)
    # states behind each vertex's output components (-1: none or not a StateMask)
    depth = max(int(nw.info.vdepth), 1)
    vout_states = np.full((nw.nv(), depth), -1, np.int64)
    for j in range(depth):
        has = v[:, 1, 1] >= v[:, 1, 0] + j
        vout_states[has, j] = ost[v[has, 1, 0] - 1 + j]
    vout_state = vout_states[:, 0]
    rows, cols = [], []

    def block(r0, n, c):
        """rows r0..r0+n-1 of a component against the column list c (per component)"""
        rr = np.repeat(r0[:, None] + np.arange(n)[None, :], c.shape[1], axis=1).reshape(len(r0), -1)
        cc = np.tile(c, (1, n))
        ok = cc.reshape(-1) >= 0
        rows.append(rr.reshape(-1)[ok])
        cols.append(cc.reshape(-1)[ok])

    rp, ent = nw.csr()
    xm = nw.extmap() if nw.info.lastidx_extbuf > 0 else None
    ve = nw.ext_ranges()[0] if xm is not None else None
    for b in nw.batches(False):
        k = nw.kinds[b["kind"]]
        if k["model"] == A.ND_VM_NOOP or k["dim"] == 0:
            continue
        idx = b["indices"] - 1
        d = k["dim"]
        r0 = v[idx, 0, 0] - 1
        block(r0, d, r0[:, None] + np.arange(d)[None, :])               # own states
        # aggregated inputs: each entry's edge, through both endpoint outputs or its state
        if nw.info.edepth > 0:
            a0 = v[idx, 3, 0] - 1
            lo, hi = rp[a0], rp[a0 + nw.info.edepth]
            cnt = hi - lo
            owner = np.repeat(np.arange(len(idx)), cnt)
            pos = np.concatenate([np.arange(x, y) for x, y in zip(lo, hi)]) if cnt.sum() else np.zeros(0, np.int64)
            k_out = ent[pos] - 1
            eid = oedge[k_out]
            st = ost[k_out]
            ff = st < 0
            nb = np.where(ff, -1, st)
            cols_ = [nb]
            for j in range(depth):              # a feed-forward entry reads every output component
                cols_.append(np.where(ff, vout_states[nw.g.src[eid] - 1, j], -1))
                cols_.append(np.where(ff, vout_states[nw.g.dst[eid] - 1, j], -1))
            for c in cols_:
                rr = r0[owner][:, None] + np.arange(d)[None, :]
                cc = np.repeat(c[:, None], d, axis=1)
                ok = cc.reshape(-1) >= 0
                rows.append(rr.reshape(-1)[ok])
                cols.append(cc.reshape(-1)[ok])
        if xm is not None and k.get("extdim", 0) > 0:
            nx = k["extdim"]
            srcs = xm[(ve[idx, 0] - 1)[:, None] + np.arange(nx)[None, :]]
            srcs = np.where(srcs > 0, srcs - 1, ost[np.abs(srcs) - 1])
            block(r0, d, srcs)
    for b in nw.batches(True):
        k = nw.kinds[b["kind"]]
        if k["dim"] == 0 or k["model"] == A.ND_EM_NOOP:
            continue
        idx = b["indices"] - 1
        d = k["dim"]
        r0 = e[idx, 0, 0] - 1
        ends = np.stack([vout_state[nw.g.src[idx] - 1], vout_state[nw.g.dst[idx] - 1]], 1)
        block(r0, d, np.concatenate([r0[:, None] + np.arange(d)[None, :], ends], 1))
    r = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    c = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    P = coo_matrix((np.ones(len(r), bool), (r, c)), shape=(nw.dim, nw.dim)).tocsr()
    P.sum_duplicates()
    P.data[:] = True
    return P


def color_columns(P):
    """Greedy distance-1 coloring of the column intersection graph: columns that share
    a row get different colors.  Returns (colors per column, number of colors)."""
    Pc = P.tocsc()
    Pr = P.tocsr()
    n = P.shape[1]
    color = np.full(n, -1, np.int64)
    mark = np.full(n + 1, -1, np.int64)
    ncolor = 0
    for j in range(n):
        rows = Pc.indices[Pc.indptr[j]:Pc.indptr[j + 1]]
        for r in rows:
            for jj in Pr.indices[Pr.indptr[r]:Pr.indptr[r + 1]]:
                cj = color[jj]
                if cj >= 0:
                    mark[cj] = j
        c = 0
        while mark[c] == j:
            c += 1
        color[j] = c
        ncolor = max(ncolor, c + 1)
    return color, ncolor


def sparse_jacobian(nw, u, p, t=0.0, pattern=None, colors=None, chunk=8):
    """d(du)/du at (u, p, t) as a scipy CSR matrix (host), from device JVPs of the color
    seeds: `chunk` (<= 8) colors per launch through nw.jvp_multi where the network's
    kernels have the K-tangent form, one color per nw.jvp otherwise.  pattern / colors:
    from jac_prototype / color_columns (computed when None)."""
    import torch
    from scipy.sparse import csr_matrix
    P = jac_prototype(nw) if pattern is None else pattern
    if colors is None:
        colors, ncolor = color_columns(P)
    else:
        ncolor = int(colors.max()) + 1 if len(colors) else 0
    Pc = P.tocoo()
    rows, cols = Pc.row, Pc.col
    n = nw.dim
    dev = u.device
    du = torch.empty(n, dtype=torch.float64, device=dev)
    jall = torch.empty((max(ncolor, 1), n), dtype=torch.float64, device=dev)
    colors_d = torch.from_numpy(np.asarray(colors, np.int64)).to(dev)
    multi = chunk > 1 and hasattr(nw, "jvp_multi")
    c0 = 0
    while c0 < ncolor:
        k = min(chunk, ncolor - c0) if multi else 1
        ids = torch.arange(c0, c0 + k, device=dev)
        seeds = (colors_d[None, :] == ids[:, None]).to(torch.float64)
        if multi:
            try:
                nw.jvp_multi(du, jall[c0:c0 + k], u, seeds, p, t)
            except A.UnsupportedError:
                multi = False
                continue
        else:
            nw.jvp(du, jall[c0], u, seeds[0], p, t)
        c0 += k
    vals = jall.cpu().numpy()[colors[cols], rows] if len(rows) else np.zeros(0)
    return csr_matrix((vals, (rows, cols)), shape=(n, n))
