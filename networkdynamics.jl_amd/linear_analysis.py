"""Fixpoint check and linearisation on the device (src/linear_analysis.jl).

isfixpoint          every |du_i| < tol at (u, p, t)                  (linear_analysis.jl:8-15)
linearize_network   A = d(du)/du, one dual-number JVP per column       (linear_analysis.jl:74-85)
jacobian_eigenvals  eigenvalues of A (identity mass matrix)           (linear_analysis.jl:232-238)
"""
import numpy as np


def isfixpoint(nw, u, p, t=0.0, tol=1e-10):
    import torch
    du = torch.empty(nw.dim, dtype=torch.float64, device=u.device)
    nw(du, u, p, t)
    return bool((du.abs() < tol).all().item())


def linearize_network(nw, u, p, t=0.0):
    """Dense state matrix A (numpy, dim x dim) of du = A x around u; the networks here
    have an identity mass matrix."""
    return nw.jacobian(u, p, t).cpu().numpy()


def jacobian_eigenvals(nw, u, p, t=0.0):
    """Eigenvalues of the linearisation, sorted by real part (largest last)."""
    ev = np.linalg.eigvals(linearize_network(nw, u, p, t))
    return ev[np.argsort(ev.real, kind="stable")]
