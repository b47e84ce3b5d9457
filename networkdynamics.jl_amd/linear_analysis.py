"""Fixpoint check and linear analysis on the device Jacobian (src/linear_analysis.jl,
src/linear_analysis_base.jl).

isfixpoint              every |du_i| < tol at (u, p, t)                       (linear_analysis.jl:8-15)
linearize_network       descriptor system M x' = A x, A = d(du)/du from one
                        dual-number JVP per column on the device              (linear_analysis.jl:74-85)
reduce_dae              eliminate the algebraic states (M_ii = 0) by the Schur
                        complement A_s = f_x - f_y g_y^-1 g_x                 (linear_analysis.jl:130-214)
jacobian_eigenvals      eigenvalues of the reduced Jacobian                   (linear_analysis.jl:230-237)
is_linear_stable        all eigenvalues in the open left half plane           (linear_analysis.jl:252-261)
participation_factors   |W[i,k] V[k,i]| of the reduced system                 (linear_analysis.jl:286-309)

The Jacobian comes from the device (`Network.jacobian`); the reductions and
eigen-decompositions of the resulting dense matrices run in numpy on the host.
"""
import warnings

import numpy as np

from . import _abi
from . import components as comp


def isfixpoint(nw, u, p, t=0.0, tol=1e-10):
    import torch
    du = torch.empty(nw.dim, dtype=torch.float64, device=u.device)
    nw(du, u, p, t)
    return bool((du.abs() < tol).all().item())


class NetworkDescriptorSystem:
    """M x' = A x + B u, y = C x + D u (linear_analysis_base.jl:21-60).  M is `I`, a
    `Diagonal` or a dense matrix.  Called with a complex frequency s it returns the
    transfer function C (M s - A)^-1 B + D (linear_analysis_base.jl:79-89)."""

    def __init__(self, A, M=comp.I, B=None, C=None, D=None, sym=None, insym=None, outsym=None):
        A_ = np.asarray(A)
        dim = A_.shape[0]
        if A_.ndim != 2 or A_.shape[1] != dim:
            raise _abi.ArgumentError("A must be square")
        if sym is not None and len(sym) != dim:
            raise _abi.ArgumentError("Length of sym must match number of rows in A")
        io = (B, C, D, insym, outsym)
        if any(x is None for x in io) and not all(x is None for x in io):
            raise _abi.ArgumentError("B, C, D, insym, and outsym must all be provided or all be nothing")
        if B is not None:
            B, C, D = np.atleast_2d(B), np.atleast_2d(C), np.atleast_2d(D)
            if B.shape[0] != dim:
                raise _abi.ArgumentError("Number of rows in B must match number of rows in A")
            if C.shape[1] != dim:
                raise _abi.ArgumentError("Number of columns in C must match number of rows in A")
            if D.shape != (C.shape[0], B.shape[1]):
                raise _abi.ArgumentError("D must have same number of rows as C and same number of columns as B")
        self.M, self.A, self.B, self.C, self.D = M, A_, B, C, D
        self.sym, self.insym, self.outsym = sym, insym, outsym

    @property
    def dim(self):
        return self.A.shape[0]

    def dense_M(self):
        return _dense_mass(self.M, self.dim)

    def constraint_dim(self):
        """Number of algebraic states (linear_analysis_base.jl:97-98)."""
        if isinstance(self.M, comp.UniformScaling):
            return 0
        return int(self.dim - np.sum(self.dense_M()))

    def __call__(self, s):
        if self.B is None:
            raise _abi.ArgumentError("the system has no inputs and outputs (linearize with in/out channels)")
        val = self.C @ np.linalg.solve(self.dense_M() * s - self.A, self.B) + self.D
        if isinstance(self.insym, comp.VIndex) and isinstance(self.outsym, comp.VIndex):
            return val[0, 0]                     # SISO: a scalar (linear_analysis_base.jl:79-89)
        return val

    def __repr__(self):
        return (f"NetworkDescriptorSystem (dim {self.dim}, {self.constraint_dim()} algebraic states"
                + (f", {self.B.shape[1]} inputs, {self.C.shape[0]} outputs)" if self.B is not None else ")"))


def _dense_mass(M, n):
    if isinstance(M, comp.UniformScaling):
        return M.s * np.eye(n)
    if isinstance(M, comp.Diagonal):
        return M.toarray()
    return np.asarray(M, dtype=np.float64)


def _vertex_outsym(m):
    out = getattr(m, "outsym", None)
    return list(out) if out else [m.sym[j - 1] for j in m.mask_dst]


def _edge_out_names(m):
    from .network import _edge_outsym_flat
    flat = _edge_outsym_flat(m)
    return flat[:m.outdim_src], flat[m.outdim_src:]


def perturbation_channels(nw, channels):
    """_classify_perturbation_channels (linear_analysis.jl:629-680): (delta0, maps) for
    a list of VIndex / EIndex channels.  A parameter name gives a p_map entry (pflat
    index); a vertex output / input name a vo_map / vi_map entry (output buffer /
    aggregation buffer index); an edge output name (src side first) an eo_map entry, an
    edge input name (src₊... / dst₊..., from insym) an ei_map entry (gather buffer
    index).  Entries are 1-based (buffer index, channel) pairs."""
    v, e = nw.ranges()
    maps = dict(vo_map=[], vi_map=[], eo_map=[], ei_map=[], p_map=[])
    for k, ref in enumerate(channels, start=1):
        i = nw._component(ref)
        s = ref.sym
        if ref.kind == "V":
            m = nw.vertexm[i]
            if s in m.psym:
                maps["p_map"].append((int(v[i, 2, 0] + m.psym.index(s)), k))
            elif s in _vertex_outsym(m):
                maps["vo_map"].append((int(v[i, 1, 0] + _vertex_outsym(m).index(s)), k))
            elif m.insym and s in m.insym:
                maps["vi_map"].append((int(v[i, 3, 0] + m.insym.index(s)), k))
            else:
                raise _abi.ArgumentError(f"Perturbation channel {ref!r} is not an input or output of its component.")
        else:
            m = nw.edgem[i]
            osrc, odst = _edge_out_names(m)
            if s in m.psym:
                maps["p_map"].append((int(e[i, 3, 0] + m.psym.index(s)), k))
            elif s in osrc:
                maps["eo_map"].append((int(e[i, 1, 0] + osrc.index(s)), k))
            elif s in odst:
                maps["eo_map"].append((int(e[i, 2, 0] + odst.index(s)), k))
            elif m.insym and s in m.insym["src"]:
                maps["ei_map"].append((int(e[i, 4, 0] + m.insym["src"].index(s)), k))
            elif m.insym and s in m.insym["dst"]:
                maps["ei_map"].append((int(e[i, 5, 0] + m.insym["dst"].index(s)), k))
            else:
                raise _abi.ArgumentError(f"Perturbation channel {ref!r} is not an input or output of its component.")
    return np.zeros(len(channels)), maps


def _observable(nw, ref):
    """Where an output channel of linearize_network reads its value: ("u", i) a state,
    ("o", i) the output buffer, ("agg", i) the aggregation buffer, ("gbuf", g, i) the
    gather-buffer entry g, which holds output i (0-based indices, g 1-based)."""
    v, e = nw.ranges()
    i = nw._component(ref)
    s = ref.sym
    if ref.kind == "V":
        m = nw.vertexm[i]
        if s in m.sym:
            return ("u", int(v[i, 0, 0] - 1 + m.sym.index(s)))
        if s in _vertex_outsym(m):
            return ("o", int(v[i, 1, 0] - 1 + _vertex_outsym(m).index(s)))
        if m.insym and s in m.insym:
            return ("agg", int(v[i, 3, 0] - 1 + m.insym.index(s)))
    else:
        m = nw.edgem[i]
        osrc, odst = _edge_out_names(m)
        if s in m.sym:
            return ("u", int(e[i, 0, 0] - 1 + m.sym.index(s)))
        if s in osrc:
            return ("o", int(e[i, 1, 0] - 1 + osrc.index(s)))
        if s in odst:
            return ("o", int(e[i, 2, 0] - 1 + odst.index(s)))
        if m.insym and (s in m.insym["src"] or s in m.insym["dst"]):
            src = s in m.insym["src"]
            j = (m.insym["src"] if src else m.insym["dst"]).index(s)
            end = (nw.g.src if src else nw.g.dst)[i] - 1
            return ("gbuf", int(e[i, 4 if src else 5, 0] + j), int(v[end, 1, 0] - 1 + j))
    raise _abi.ArgumentError(f"Cannot observe {ref!r}: not a state, output or input of its component.")


def linearize_network(nw, u, p, t=0.0, in_=None, out=None):
    """linearize_network(s0; in, out), linear_analysis.jl:74-122.

    Without channels: NetworkDescriptorSystem(M = nw.mass_matrix, A = d(du)/du, sym).
    With perturbation channels in_ (VIndex / EIndex or a list of them: parameters,
    vertex outputs / inputs, edge outputs / inputs) and observations out (states,
    outputs, inputs): the full M x' = A x + B u, y = C x + D u.  B and D are the
    derivatives of f and of the observations in the channels at delta = 0, C those of
    the observations in the states: one dual-number pass of the perturbed loop per
    channel and per state column (parameters enter as dual numbers as well).  A single
    channel / observation (not a list) gives a SISO system whose transfer function is a
    scalar."""
    import torch
    dev = u.device if isinstance(u, torch.Tensor) else torch.device("cuda", torch.cuda.current_device())
    u = torch.as_tensor(u, dtype=torch.float64, device=dev).contiguous()
    pd = torch.as_tensor(p if len(p) else np.zeros(1), dtype=torch.float64, device=dev).contiguous()
    if in_ is None and out is None:
        Jd = nw.jacobian(u, pd, t).cpu().numpy()
        return NetworkDescriptorSystem(Jd, M=nw.mass_matrix, sym=nw.variable_symbols())
    if in_ is None:
        raise _abi.ArgumentError("Must specify `in` when `out` is given")
    if out is None:
        raise _abi.ArgumentError("Must specify `out` when `in` is given")
    single_in, single_out = isinstance(in_, comp.VIndex), isinstance(out, comp.VIndex)
    insym = [in_] if single_in else list(in_)
    outsym = [out] if single_out else list(out)
    delta0, maps = perturbation_channels(nw, insym)
    obs = [_observable(nw, r) for r in outsym]
    n, nin, nout = nw.dim, len(insym), len(outsym)
    A_ = nw.jacobian(u, pd, t).cpu().numpy()
    B = np.zeros((n, nin))
    C = np.zeros((nout, n))
    D = np.zeros((nout, nin))
    du = torch.empty(n, dtype=torch.float64, device=u.device)
    jv = torch.empty_like(du)
    ei = maps["ei_map"]

    def observe(bufs, v_host, dd_host):
        o, do, agg, dagg = (b.cpu().numpy() for b in bufs)
        y = np.empty(nout)
        for r, ob in enumerate(obs):
            if ob[0] == "u":
                y[r] = v_host[ob[1]] if v_host is not None else 0.0
            elif ob[0] == "o":
                y[r] = do[ob[1]]
            elif ob[0] == "agg":
                y[r] = dagg[ob[1]]
            else:                           # gbuf = gathered output + its edge-input channels
                y[r] = do[ob[2]] + sum(dd_host[k - 1] for g, k in ei if g == ob[1])
        return y

    p_chan = {k: b for b, k in maps["p_map"]}
    for c in range(nin):
        dd = np.zeros(nin)
        dp = None
        if c + 1 in p_chan:                 # parameter channel: p + delta at pflat index b
            dp = torch.zeros(max(nw.pdim, 1), dtype=torch.float64, device=u.device)
            dp[p_chan[c + 1] - 1] = 1.0
            dp = dp[:nw.pdim] if nw.pdim else dp
        else:
            dd[c] = 1.0
        bufs = nw.perturbed_jvp(du, jv, u, None, pd, t, delta0, dd, maps, buffers=True, dp=dp)
        B[:, c] = jv.cpu().numpy()
        D[:, c] = observe(bufs, None, dd)
    if any(ob[0] != "u" for ob in obs):
        for j in range(n):
            vj = np.zeros(n)
            vj[j] = 1.0
            bufs = nw.perturbed_jvp(du, jv, u, torch.from_numpy(vj).to(u.device), pd, t, delta0, None, maps,
                                    buffers=True)
            C[:, j] = observe(bufs, vj, np.zeros(nin))
    else:
        for r, ob in enumerate(obs):
            C[r, ob[1]] = 1.0
    return NetworkDescriptorSystem(A_, M=nw.mass_matrix, B=B, C=C, D=D, sym=nw.variable_symbols(),
                                   insym=in_ if single_in else insym, outsym=out if single_out else outsym)


def reduce_dae(sys):
    """The ODE system of a descriptor system with algebraic constraints
    (linear_analysis.jl:130-214).  Differential states are those with M_ii = 1,
    algebraic ones those with M_ii = 0; with x_a solved from 0 = g_x x_d + g_y x_a + B_a u:
        A_s = f_x - f_y g_y^-1 g_x,   B_s = B_d - f_y g_y^-1 B_a,
        C_s = C_d - C_a g_y^-1 g_x,   D_s = D - C_a g_y^-1 B_a.
    A singular g_y falls back to its pseudo-inverse with a warning, as the reference does.
    M = I returns the system unchanged."""
    if isinstance(sys.M, comp.UniformScaling):
        return NetworkDescriptorSystem(sys.A, M=comp.I, B=sys.B, C=sys.C, D=sys.D, sym=sys.sym,
                                       insym=sys.insym, outsym=sys.outsym)
    d = np.diag(sys.dense_M())
    d_idx = np.flatnonzero(d == 1)
    c_idx = np.flatnonzero(d == 0)
    Am = sys.A
    f_x = Am[np.ix_(d_idx, d_idx)]
    f_y = Am[np.ix_(d_idx, c_idx)]
    g_x = Am[np.ix_(c_idx, d_idx)]
    g_y = Am[np.ix_(c_idx, c_idx)]
    try:
        gy_gx = np.linalg.solve(g_y, g_x) if g_y.size else np.zeros((0, len(d_idx)))

        def gy_solve(x):
            return np.linalg.solve(g_y, x) if g_y.size else np.zeros((0,) + x.shape[1:])
    except np.linalg.LinAlgError as e:
        warnings.warn(f"Matrix g_y is singular or ill-conditioned, falling back to pseudoinverse ({e})")
        gy_inv = np.linalg.pinv(g_y)
        gy_gx = gy_inv @ g_x

        def gy_solve(x):
            return gy_inv @ x
    A_s = f_x - f_y @ gy_gx
    B_s, C_s, D_s = sys.B, sys.C, sys.D
    if sys.B is not None:
        B_d, B_a = sys.B[d_idx, :], sys.B[c_idx, :]
        C_d, C_a = sys.C[:, d_idx], sys.C[:, c_idx]
        B_s = B_d - f_y @ gy_solve(B_a)
        C_s = C_d - C_a @ gy_gx
        D_s = sys.D - C_a @ gy_solve(B_a)
    sym = None if sys.sym is None else [sys.sym[i] for i in d_idx]
    return NetworkDescriptorSystem(A_s, M=comp.I, B=B_s, C=C_s, D=D_s, sym=sym, insym=sys.insym,
                                   outsym=sys.outsym)


def _sort_eigs(ev):
    """Julia's eigvals order for a general matrix: lexicographic by (real, imag)."""
    return ev[np.lexsort((ev.imag, ev.real))]


def _as_system(x, u, p, t):
    if isinstance(x, NetworkDescriptorSystem):
        return x
    if u is None or p is None:
        raise _abi.ArgumentError("pass a NetworkDescriptorSystem, or a network with u and p")
    return linearize_network(x, u, p, t)


def jacobian_eigenvals(sys_or_nw, u=None, p=None, t=0.0, eigvalf=np.linalg.eigvals):
    """Eigenvalues of the reduced Jacobian, in (real, imag) order
    (linear_analysis.jl:230-237).  Takes a NetworkDescriptorSystem or (nw, u, p, t)."""
    red = reduce_dae(_as_system(sys_or_nw, u, p, t))
    return _sort_eigs(np.asarray(eigvalf(red.A)))


def is_linear_stable(nw, u, p, t=0.0, marginally_stable=False, atol=1e-14, tol=1e-10):
    """All eigenvalues of the reduced Jacobian have negative real part; with
    marginally_stable also those within atol of zero (linear_analysis.jl:252-261).
    The state must be a fixpoint."""
    if not isfixpoint(nw, u, p, t, tol=tol):
        raise _abi.ArgumentError("The state s0 is not a fixpoint of the network.")
    lam = jacobian_eigenvals(nw, u, p, t)
    if marginally_stable:
        return bool(np.all((lam.real < atol) | np.isclose(lam.real, 0.0, atol=atol, rtol=0.0)))
    return bool(np.all(lam.real < 0.0))


def participation_factors(sys_or_nw, u=None, p=None, t=0.0, normalize=True):
    """(eigenvalues, pfactors, sym) of the reduced system, pfactors[k, i] = |W[i,k] V[k,i]|
    with V the right eigenvectors and W = V^-1; normalize scales each mode's column to
    sum 1 (linear_analysis.jl:286-309).  Eigenvalues in the decomposition's order."""
    red = reduce_dae(_as_system(sys_or_nw, u, p, t))
    ev, V = np.linalg.eig(red.A)
    W = np.linalg.inv(V)
    pf = np.abs(W.T * V)
    if normalize:
        s = pf.sum(axis=0)
        nz = s > 0
        pf[:, nz] /= s[nz]
    return dict(eigenvalues=ev, pfactors=pf, sym=red.sym)
