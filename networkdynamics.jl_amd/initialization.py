"""find_fixpoint for any network the device plan runs (src/initialization.jl:44-144).

The reference wraps `nw` as a SteadyStateProblem at fixed t and hands it to
NonlinearSolve (SSRootfind).  Here Newton's method runs on the device network:
every residual is one `nw(du, x, p, t)` launch and every Jacobian column one
dual-number JVP launch (`Network.jacobian`).  The Newton step solves J s = -du on the
host: rank-revealing least squares on the dense Jacobian for small networks (the
global phase of a Kuramoto-type network leaves J singular, and least squares takes
the minimum-norm step), a sparse LU of the colored sparse Jacobian (one JVP launch
per color) above `dense_max` states, or GMRES whose products are device JVPs.  A
backtracking line search on ||du||_2 keeps flat starts robust.  Grids of swing/load
vertices in bulk belong to `FixpointSolver`, which runs thousands of them in one
launch.

Errors follow the reference: a NaN right-hand side at x0 raises NetworkInitError
naming the NaN states and the NaN inputs (initialization.jl:78-120); a solve that
does not reach the tolerance raises NetworkInitError with the residual
(initialization.jl:127-141).
"""
import math

import numpy as np

from . import _abi as A


class NetworkInitError(RuntimeError):
    """NetworkInitError (src/initialization.jl)."""


def _fmt(sym):
    kind, i, name = sym
    return f"{'VIndex' if kind == 'V' else 'EIndex'}({i}, :{name})"


def _nan_report(nw, x, p, du, t):
    lines = [f"Evaluation of network rhs at t={t} resulted in NaNs.)"]
    vs = nw.variable_symbols()
    nans = np.flatnonzero(np.isnan(du))
    lines.append("  NaN-Results:  " + ", ".join(_fmt(vs[i]) for i in nans))
    lines.append("Possible causes:")
    if np.isnan(x).any() or np.isnan(p).any():
        lines.append(" - Some input states/parameters contained NaN:")
        if np.isnan(x).any():
            lines.append("   Variables:  " + ", ".join(_fmt(vs[i]) for i in np.flatnonzero(np.isnan(x))))
        if np.isnan(p).any():
            ps = nw.parameter_symbols()
            lines.append("   Parameters: " + ", ".join(_fmt(ps[i]) for i in np.flatnonzero(np.isnan(p))))
    if math.isnan(t):
        lines.append(" - Network was evaluated at t=NaN, maybe your system has explicit time dependency? "
                     "Try specifying kw argument `t` to decide on time")
    return "\n".join(lines)


def _newton_step(nw, x, p, t, du, linsolve, gmres_rtol, sparse_state):
    """s with J(x) s = -du, as a host float64 vector.  dense: least squares on the dense
    device Jacobian; sparse: LU of J + mu*I on the colored sparse Jacobian, mu = 1e-12
    max|J| so that the null direction of a phase-invariant network gets a bounded
    component; gmres: matrix-free with device JVPs."""
    import torch
    n = nw.dim
    rhs = -du.cpu().numpy()
    if linsolve == "dense":
        from scipy.linalg import lstsq
        J = nw.jacobian(x, p, t).cpu().numpy()
        return lstsq(J, rhs, lapack_driver="gelsy")[0]
    if linsolve == "sparse":
        from scipy.sparse import identity
        from scipy.sparse.linalg import splu
        from .sparsity import color_columns, jac_prototype, sparse_jacobian
        if "pattern" not in sparse_state:
            sparse_state["pattern"] = jac_prototype(nw)
            sparse_state["colors"] = color_columns(sparse_state["pattern"])[0]
        J = sparse_jacobian(nw, x, p, t, sparse_state["pattern"], sparse_state["colors"])
        mu = 1e-12 * (np.abs(J.data).max() if J.nnz else 1.0)
        return splu((J + mu * identity(n, format="csr")).tocsc()).solve(rhs)
    from scipy.sparse.linalg import LinearOperator, gmres
    v_d = torch.empty(n, dtype=torch.float64, device=x.device)
    jv = torch.empty_like(v_d)
    scratch = torch.empty_like(v_d)

    def matvec(v):
        v_d.copy_(torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64).reshape(-1)))
        nw.jvp(scratch, jv, x, v_d, p, t)
        return jv.cpu().numpy()

    op = LinearOperator((n, n), matvec=matvec, dtype=np.float64)
    s, _ = gmres(op, rhs, rtol=gmres_rtol, atol=0.0, restart=min(n, 100), maxiter=20)
    return s


def find_fixpoint(nw, x0, p, t=float("nan"), tol=1e-10, maxiters=50, linsolve="auto", dense_max=2048,
                  gmres_rtol=1e-10):
    """find_fixpoint(nw, x0, p; t) -- src/initialization.jl:73-144.  x0, p: host arrays or
    float64 device tensors.  linsolve: "dense", "sparse" (colored Jacobian, sparse LU),
    "gmres" (matrix-free), or "auto" (dense up to dense_max states, else sparse).
    Returns (x, info): x the fixpoint as a float64 device tensor (max |du| <= tol, the
    isfixpoint test of linear_analysis.jl:8-15), info the Newton iterations, line-search
    halvings and final residual."""
    if linsolve == "auto":
        linsolve = "dense" if nw.dim <= dense_max else "sparse"
    if linsolve not in ("dense", "sparse", "gmres"):
        raise A.ArgumentError(f"linsolve must be dense, sparse, gmres or auto, got {linsolve!r}")
    sparse_state = {}
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    x = torch.as_tensor(np.asarray(x0, dtype=np.float64) if not isinstance(x0, torch.Tensor) else x0,
                        dtype=torch.float64, device=dev).clone()
    pd = torch.as_tensor(np.asarray(p, dtype=np.float64) if not isinstance(p, torch.Tensor) else p,
                         dtype=torch.float64, device=dev).contiguous()
    if x.numel() != nw.dim:
        raise A.ArgumentError(f"x0 has length {x.numel()}, the network has {nw.dim} states")
    du = torch.empty_like(x)
    nw(du, x, pd, t)
    duh = du.cpu().numpy()
    if np.isnan(duh).any():
        raise NetworkInitError(_nan_report(nw, x.cpu().numpy(), pd.cpu().numpy(), duh, t))
    f0 = float(torch.dot(du, du))
    halvings = 0
    trial = torch.empty_like(x)
    dut = torch.empty_like(x)
    for it in range(maxiters + 1):
        res = float(du.abs().max())
        if res <= tol:
            return x, dict(iters=it, halvings=halvings, residual=res)
        if it == maxiters:
            break
        s = torch.from_numpy(_newton_step(nw, x, pd, t, du, linsolve, gmres_rtol, sparse_state)).to(dev)
        lam, accepted = 1.0, False
        while lam >= 1.0 / 1024:
            torch.add(x, s, alpha=lam, out=trial)
            nw(dut, trial, pd, t)
            f1 = float(torch.dot(dut, dut))
            if math.isfinite(f1) and f1 <= (1.0 - 1e-4 * lam) * f0:
                accepted = True
                break
            lam *= 0.5
            halvings += 1
        if not accepted:
            break
        x, trial = trial, x
        du, dut = dut, du
        f0 = f1
    res = float(du.abs().max())
    raise NetworkInitError(
        f"Could not find fixpoint, solver returned {'Stalled' if it < maxiters else 'MaxIters'} with residual "
        f"{res:.3e} (Newton on the device network, {it} iterations)!")
