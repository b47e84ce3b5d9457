"""CPU reference of the warm-start message-passing network (test infrastructure only).

Float64 numpy restatement of the network the HIP kernels train:
    H0 = relu(X W_in + b_in)
    H_l = relu(A (H_{l-1} W_l) + H_{l-1} U_l + b_l),  l = 1..4
    out = H_4 w_out + b_out,   loss = mean((out - target)^2)
with A the row-normalised weighted adjacency.  The backward pass here is checked
against central finite differences in tests/test_warmstart.py before it is used to
check the GPU gradients.  (No reference counterpart: NetworkDynamics.jl has no
learned initialisation.)
"""
import numpy as np

H = 64
L = 4


def unpack(params, nfeat):
    p = np.asarray(params, np.float64)
    k = 0
    out = {}
    out["W_in"] = p[k:k + nfeat * H].reshape(nfeat, H); k += nfeat * H
    out["b_in"] = p[k:k + H]; k += H
    for l in range(L):
        out[f"WU{l}"] = p[k:k + H * 2 * H].reshape(H, 2 * H); k += H * 2 * H
        out[f"b{l}"] = p[k:k + H]; k += H
    out["w_out"] = p[k:k + H]; k += H
    out["b_out"] = p[k:k + 1]; k += 1
    assert k == len(p)
    return out


def pack(d, nfeat):
    parts = [d["W_in"].ravel(), d["b_in"]]
    for l in range(L):
        parts += [d[f"WU{l}"].ravel(), d[f"b{l}"]]
    parts += [d["w_out"], d["b_out"]]
    return np.concatenate(parts)


def adjacency(n, rowptr, nbr, a):
    import scipy.sparse as sp
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    return sp.csr_matrix((np.asarray(a, np.float64), (rows, np.asarray(nbr))), shape=(n, n))


def forward(params, X, A):
    X = np.asarray(X, np.float64)
    P = unpack(params, X.shape[1])
    Hs = [np.maximum(X @ P["W_in"] + P["b_in"], 0.0)]
    for l in range(L):
        Y = Hs[-1] @ P[f"WU{l}"]
        Hs.append(np.maximum(A @ Y[:, :H] + Y[:, H:] + P[f"b{l}"], 0.0))
    out = Hs[-1] @ P["w_out"] + P["b_out"][0]
    return out, Hs


def loss_and_grad(params, X, A, target):
    X = np.asarray(X, np.float64)
    n = X.shape[0]
    P = unpack(params, X.shape[1])
    out, Hs = forward(params, X, A)
    e = out - np.asarray(target, np.float64)
    loss = np.mean(e * e)
    G = {}
    dout = 2.0 * e / n
    G["w_out"] = Hs[-1].T @ dout
    G["b_out"] = np.array([dout.sum()])
    dH = np.outer(dout, P["w_out"])
    for l in reversed(range(L)):
        g = dH * (Hs[l + 1] > 0)
        G[f"b{l}"] = g.sum(axis=0)
        dYa = A.T @ g
        dY = np.concatenate([dYa, g], axis=1)
        G[f"WU{l}"] = Hs[l].T @ dY
        dH = dY @ P[f"WU{l}"].T
    g0 = dH * (Hs[0] > 0)
    G["b_in"] = g0.sum(axis=0)
    G["W_in"] = X.T @ g0
    return loss, pack(G, X.shape[1])
