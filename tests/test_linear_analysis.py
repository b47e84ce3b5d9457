"""Descriptor systems and their reduction (src/linear_analysis.jl:130-309) on host
matrices, pinned by facts that do not restate the Schur-complement formula:
substituting explicitly solved algebraic states, the invariance of the transfer
function, the finite generalized eigenvalues of (A, M), and orthogonal eigenvectors
for a symmetric A.  No GPU."""
import warnings

import numpy as np
import pytest
import scipy.linalg as sla


def _system_with_explicit_constraints(rng, nd, na, ni=2, no=3):
    """States ordered [x_d, x_a] shuffled; constraints 0 = K x_d + L u - x_a."""
    f_x = rng.normal(size=(nd, nd))
    f_y = rng.normal(size=(nd, na))
    K = rng.normal(size=(na, nd))
    L = rng.normal(size=(na, ni))
    B_d = rng.normal(size=(nd, ni))
    C_d, C_a = rng.normal(size=(no, nd)), rng.normal(size=(no, na))
    D = rng.normal(size=(no, ni))
    n = nd + na
    perm = rng.permutation(n)                # flat position of [x_d, x_a][j]
    A = np.zeros((n, n))
    B = np.zeros((n, ni))
    C = np.zeros((no, n))
    d_pos, a_pos = perm[:nd], perm[nd:]
    A[np.ix_(d_pos, d_pos)] = f_x
    A[np.ix_(d_pos, a_pos)] = f_y
    A[np.ix_(a_pos, d_pos)] = K
    A[np.ix_(a_pos, a_pos)] = -np.eye(na)
    B[d_pos] = B_d
    B[a_pos] = L
    C[:, d_pos] = C_d
    C[:, a_pos] = C_a
    mdiag = np.zeros(n)
    mdiag[d_pos] = 1.0
    # substituting x_a = K x_d + L u by hand:
    order = np.argsort(d_pos)                # reduce_dae keeps the differential states in flat order
    exp = dict(A=(f_x + f_y @ K)[np.ix_(order, order)], B=(B_d + f_y @ L)[order],
               C=(C_d + C_a @ K)[:, order], D=D + C_a @ L)
    return A, B, C, D, mdiag, exp


@pytest.mark.parametrize("seed", range(4))
def test_reduce_dae_equals_substitution(ND, seed):
    LA = ND.linear_analysis
    rng = np.random.default_rng(seed)
    A, B, C, D, mdiag, exp = _system_with_explicit_constraints(rng, 5 + seed, 3 + seed)
    sys = LA.NetworkDescriptorSystem(A, M=ND.Diagonal(mdiag), B=B, C=C, D=D, insym=["a", "b"],
                                     outsym=["x", "y", "z"])
    assert sys.constraint_dim() == 3 + seed
    red = LA.reduce_dae(sys)
    assert red.M == ND.I
    for k in ("A", "B", "C", "D"):
        np.testing.assert_allclose(getattr(red, k), exp[k], rtol=1e-12, atol=1e-12)
    # the transfer function does not change under the reduction
    for s in (0.3 + 1.1j, -2.0 + 0.5j, 4.0j):
        np.testing.assert_allclose(red(s), sys(s), rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("seed", range(3))
def test_reduced_eigenvalues_are_the_finite_generalized_ones(ND, seed):
    LA = ND.linear_analysis
    rng = np.random.default_rng(10 + seed)
    n, na = 9, 4
    A = rng.normal(size=(n, n))
    mdiag = np.ones(n)
    mdiag[rng.choice(n, na, replace=False)] = 0.0
    sys = LA.NetworkDescriptorSystem(A, M=ND.Diagonal(mdiag))
    ev = LA.jacobian_eigenvals(sys)
    gen = sla.eigvals(A, np.diag(mdiag))
    gen = gen[np.isfinite(gen)]
    assert len(ev) == n - na == len(gen)
    np.testing.assert_allclose(np.sort_complex(ev), np.sort_complex(gen), rtol=1e-8, atol=1e-9)
    # (real, imag) order
    key = np.stack([ev.real, ev.imag], 1)
    assert all(tuple(key[i]) <= tuple(key[i + 1]) for i in range(len(ev) - 1))


def test_identity_mass_passes_through(ND):
    LA = ND.linear_analysis
    A = np.diag([-1.0, -2.0, 3.0])
    red = LA.reduce_dae(LA.NetworkDescriptorSystem(A))
    assert red.A is not None and np.array_equal(red.A, A)
    np.testing.assert_array_equal(LA.jacobian_eigenvals(LA.NetworkDescriptorSystem(A)), [-2.0, -1.0, 3.0])


def test_participation_factors_of_a_symmetric_matrix(ND):
    # V orthogonal, W = V^T: p_ki = V[k,i]^2, each column already sums to 1
    LA = ND.linear_analysis
    rng = np.random.default_rng(3)
    X = rng.normal(size=(6, 6))
    S = X + X.T
    out = LA.participation_factors(LA.NetworkDescriptorSystem(S), normalize=False)
    w, V = np.linalg.eigh(S)
    np.testing.assert_allclose(out["pfactors"].sum(axis=0), 1.0, rtol=1e-10)
    for i, lam in enumerate(out["eigenvalues"]):
        j = np.argmin(np.abs(w - lam.real))
        np.testing.assert_allclose(np.sort(out["pfactors"][:, i]), np.sort(V[:, j] ** 2), atol=1e-10)
    # a diagonal system: every mode is one state
    out = LA.participation_factors(LA.NetworkDescriptorSystem(np.diag([-1.0, -3.0, -2.0])))
    np.testing.assert_allclose(out["pfactors"], np.eye(3))


def test_singular_constraint_block_falls_back_to_pinv(ND):
    LA = ND.linear_analysis
    A = np.array([[-1.0, 1.0, 0.0],
                  [0.0, 0.0, 0.0],
                  [0.0, 0.0, 0.0]])
    sys = LA.NetworkDescriptorSystem(A, M=ND.Diagonal([1.0, 0.0, 0.0]))
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        red = LA.reduce_dae(sys)
    assert any("pseudoinverse" in str(x.message) for x in w)
    np.testing.assert_allclose(red.A, [[-1.0]])


def test_descriptor_argument_errors(ND):
    LA = ND.linear_analysis
    with pytest.raises(ND.ArgumentError):
        LA.NetworkDescriptorSystem(np.zeros((3, 3)), sym=["a"])
    with pytest.raises(ND.ArgumentError):
        LA.NetworkDescriptorSystem(np.zeros((3, 3)), B=np.zeros((3, 1)))
    with pytest.raises(ND.ArgumentError):
        LA.NetworkDescriptorSystem(np.zeros((3, 3)), B=np.zeros((2, 1)), C=np.zeros((1, 3)),
                                   D=np.zeros((1, 1)), insym=["u"], outsym=["y"])
    with pytest.raises(ND.ArgumentError):
        LA.NetworkDescriptorSystem(np.zeros((3, 3)))(1.0j)
