"""Time integration of networks with algebraic states (diagonal mass matrices).

The reference hands `ODEProblem(nw, u0, tspan, p)` with `mass_matrix = nw.mass_matrix`
(construction.jl:409-453, post_utils.jl:43-82) to an implicit OrdinaryDiffEq solver
such as Rodas5P.  For index-1 networks -- every algebraic state (M_ii = 0) determined
by its own constraints, f_a(u) = 0 with a nonsingular block df_a/du_a -- the same
trajectories follow from the state-space form: the differential states advance with
a Runge-Kutta method and, at every stage, the algebraic states are solved from the
constraints.  Here that is classic RK4 on u_d' = f_d(u)/M_dd, each stage state made
consistent by simplified Newton on f_a with the block df_a/du_a of the colored sparse
Jacobian (one device JVP launch per color) factored once per step.  Every residual is a
device RHS launch; the Newton corrections are host sparse solves.
"""
import numpy as np

from . import _abi as A
from . import components as comp


class DAEError(RuntimeError):
    """The algebraic states could not be made consistent."""


def _split(nw):
    M = nw.mass_matrix
    if isinstance(M, comp.UniformScaling):
        d = np.full(nw.dim, M.s)
    elif isinstance(M, comp.Diagonal):
        d = M.diag.copy()
    else:
        raise A.UnsupportedError("the DAE integrator takes diagonal mass matrices")
    return np.flatnonzero(d != 0), np.flatnonzero(d == 0), d


class _Constraints:
    """Simplified Newton on f_a(u) = 0 over the algebraic states."""

    def __init__(self, nw, p, a_idx, tol, maxiter):
        import torch
        from .sparsity import color_columns, jac_prototype
        self.nw, self.p, self.tol, self.maxiter = nw, p, tol, maxiter
        self.a = a_idx
        self.a_d = torch.from_numpy(a_idx).to(p.device)
        self.lu = None
        self.iters = 0
        self.refactor = 0
        if len(a_idx):
            self.pattern = jac_prototype(nw)
            self.colors = color_columns(self.pattern)[0]

    def factor(self, y, t):
        from scipy.sparse.linalg import splu
        from .sparsity import sparse_jacobian
        J = sparse_jacobian(self.nw, y, self.p, t, self.pattern, self.colors).tocsc()
        self.lu = splu(J[self.a][:, self.a].tocsc())
        self.refactor += 1

    def solve(self, y, du, t):
        """Makes y consistent in place; leaves du = f(y)."""
        import torch
        nw, p = self.nw, self.p
        nw(du, y, p, t)
        if len(self.a) == 0:
            return
        for attempt in range(2):
            if self.lu is None:
                self.factor(y, t)
            prev = np.inf
            for _ in range(self.maxiter):
                r = du[self.a_d]
                res = float(r.abs().max())
                if res <= self.tol:
                    return
                if res > 0.5 * prev and prev < np.inf:
                    break                          # slow contraction: refresh the Jacobian
                prev = res
                step = torch.from_numpy(self.lu.solve(r.cpu().numpy())).to(y.device)
                y[self.a_d] -= step
                nw(du, y, p, t)
                self.iters += 1
            self.lu = None
        raise DAEError(f"constraints not met at t={t}: max |f_a| = {float(du[self.a_d].abs().max()):.3e}")


def rk4(nw, u0, p, t0, dt, nsteps, tol=1e-12, maxiter=20):
    """Fixed-step RK4 of M u' = f(u, p, t) for a diagonal mass matrix; u0 is made
    consistent first.  u0, p: host arrays or float64 device tensors.  Returns (u, info)
    with u the final state (device) and info the Newton iterations and Jacobian
    factorizations."""
    import torch
    d_idx, a_idx, mdiag = _split(nw)
    dev = torch.device("cuda", torch.cuda.current_device())
    y = torch.as_tensor(np.asarray(u0, dtype=np.float64) if not isinstance(u0, torch.Tensor) else u0,
                        dtype=torch.float64, device=dev).clone()
    pd = torch.as_tensor(np.asarray(p, dtype=np.float64) if not isinstance(p, torch.Tensor) else p,
                         dtype=torch.float64, device=dev).contiguous()
    if y.numel() != nw.dim:
        raise A.ArgumentError(f"u0 has length {y.numel()}, the network has {nw.dim} states")
    inv_m = torch.zeros(nw.dim, dtype=torch.float64, device=dev)
    inv_m[torch.from_numpy(d_idx).to(dev)] = torch.from_numpy(1.0 / mdiag[d_idx]).to(dev)
    dm = inv_m != 0                                    # differential states
    con = _Constraints(nw, pd, a_idx, tol, maxiter)
    du = torch.empty_like(y)
    ys = torch.empty_like(y)
    acc = torch.empty_like(y)
    con.solve(y, du, t0)
    half = 0.5 * dt
    for k in range(nsteps):
        tk = t0 + k * dt
        con.lu = None                                  # one factorization per step
        # stage 1 at y (consistent), derivative of the differential part only
        k1 = du * inv_m
        acc.copy_(k1)
        ys.copy_(y).add_(k1, alpha=half)
        con.solve(ys, du, tk + half)
        k2 = du * inv_m
        acc.add_(k2, alpha=2.0)
        ys[dm] = (y + half * k2)[dm]
        con.solve(ys, du, tk + half)
        k3 = du * inv_m
        acc.add_(k3, alpha=2.0)
        ys[dm] = (y + dt * k3)[dm]
        con.solve(ys, du, tk + dt)
        k4 = du * inv_m
        acc.add_(k4)
        y[dm] = (y + (dt / 6.0) * acc)[dm]
        y[~dm] = ys[~dm]                               # algebraic guess: the last stage's
        con.solve(y, du, tk + dt)
    return y, dict(newton_iters=con.iters, factorizations=con.refactor)
