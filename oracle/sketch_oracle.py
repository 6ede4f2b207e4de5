"""FP64 CPU restatement of the learned per-mode sketch's loss and gradient.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product. No reference
counterpart: the reference's range finder draws a dense Gaussian test matrix
(fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-444); the build
replaces it by a Khatri-Rao test matrix of per-mode sketches S_m (w x n_m). For a node t
the randomised method sketches, with column modes cols(t),

    Omega_t[c, j] = prod_{m in cols(t)} S_m[j, i_m(c)]        (c column-major over cols(t))
    Y = A Omega_t,  A = X_(t),  W = Y^T Y + ridge * tr(Y^T Y) / w * I,  E = Y^T A A^T Y
    loss_t = 1 - tr(W^-1 E) / ||X||^2        (the energy fraction range(Y) misses)
    dloss_t/dOmega = -(2/||X||^2) (A^T A Z W^-1 - Z C2 - (ridge/w) tr(C2) Z),
                     Z = A^T Y, C2 = W^-1 E W^-1 (the last term: the ridge follows tr(Y^T Y))
    dloss_t/dS_m[j, i] = sum_{c: i_m(c) = i} dloss_t/dOmega[c, j] prod_{m' != m} S_m'[j, i_m'(c)]

and the loss of a tensor is the mean over the sketched nodes.

Two independent evaluations:
  * ``dense_loss_grad``: explicit unfoldings and Omega (any tree the method handles, small n);
  * ``bump_loss_grad``: the Gaussian-bump family is separable, X = sum_q a_q g_q1 o ... o g_qd,
    so for node {1,2} of a 4th-order tensor A = U D V^T with U = [g_q1 (x) g_q2], V = [g_q3 (x) g_q4];
    with P = V^T Omega, K1 = D U^T U D, K2 = K1 V^T V K1: Y^T Y = P^T K1 P, Y^T A A^T Y = P^T K2 P, and
    d tr(W^-1 E)/dP = 2 K2 P W^-1 - 2 K1 P C2 - (2 ridge/w) tr(C2) K1 P;
    dP[q, j]/dS_3[j, i] = g_q3[i] (g_q4 . S_4[j, :]).
The tests pin each against the other, against central finite differences, and against
the exact optimum (the tail of A's singular values beyond w bounds every loss below).
"""
import numpy as np

from . import gen as G
from .ht_oracle import unfold

RIDGE = 1e-6  # relative ridge on Y^T Y, the value the library uses


def sketched_nodes(tree):
    """Nodes the randomised method runs a range finder on (a child of the root whose
    interior sibling comes out of its right singular vectors is the only one of the two)."""
    lo, hi, left, right = tree.lo, tree.hi, tree.left, tree.right
    order = hi[0]
    parent = {c: t for t in range(len(lo)) if left[t] >= 0 for c in (left[t], right[t])}
    done, out = set(), []
    for t in range(1, len(lo)):
        if left[t] < 0 or t in done:
            continue
        assert lo[t] == 0 or hi[t] == order
        out.append(t)
        if parent[t] == 0:
            sib = right[0] if left[0] == t else left[0]
            if left[sib] >= 0:
                done.add(sib)
    return out


def col_modes(tree, t):
    lo, hi = tree.lo[t], tree.hi[t]
    order = tree.hi[0]
    return list(range(hi, order)) if lo == 0 else list(range(0, lo))


def kr_omega(S, modes):
    """Omega[c, j] = prod_{m in modes} S[m][j, i_m(c)], c column-major over `modes`."""
    w = S[modes[0]].shape[0]
    om = np.ones((1, w))
    for m in modes:  # each further mode is a slower index
        om = np.einsum("cj,ji->cij", om, S[m]).reshape(-1, w, order="F")
    return om


def kr_backward(S, modes, dom):
    """{m: d<dom, Omega>/dS_m} for m in modes (w x n_m arrays)."""
    w = dom.shape[1]
    D = dom.reshape([S[m].shape[1] for m in modes] + [w], order="F")
    letters = "abcdefg"[:len(modes)]
    out = {}
    for k, m in enumerate(modes):
        ops, subs = [D], [letters + "z"]
        for kk, mm in enumerate(modes):
            if kk != k:
                ops.append(S[mm])
                subs.append("z" + letters[kk])
        out[m] = np.einsum(",".join(subs) + "->z" + letters[k], *ops)
    return out


def node_loss_grad(A, xsq, om, ridge=RIDGE):
    """(loss, dloss/dOmega) of one node from the dense unfolding A."""
    w = om.shape[1]
    Y = A @ om
    Wm = Y.T @ Y
    Wm = Wm + ridge * np.trace(Wm) / w * np.eye(w)
    Z = A.T @ Y
    E = Z.T @ Z
    Wi = np.linalg.inv(Wm)
    loss = 1.0 - np.trace(Wi @ E) / xsq
    C2 = Wi @ E @ Wi
    # the ridge follows tr(Y^T Y): d tr(W^-1 E) gains -(2 ridge / w) tr(C2) tr(Y^T dY)
    dom = -(2.0 / xsq) * (A.T @ (A @ (Z @ Wi)) - Z @ C2 - ridge / w * np.trace(C2) * Z)
    return loss, dom


def dense_loss_grad(x, dims, tree, S, ridge=RIDGE):
    """Mean loss over the sketched nodes and its gradient (list per mode; zeros for modes
    no sketched node uses). x is flat column-major; S a list of w x n_m arrays."""
    nodes = sketched_nodes(tree)
    xsq = float(np.dot(x, x))
    grad = [np.zeros_like(s) for s in S]
    loss = 0.0
    for t in nodes:
        A = unfold(x, dims, tree.lo[t], tree.hi[t])
        modes = col_modes(tree, t)
        f, dom = node_loss_grad(A, xsq, kr_omega(S, modes), ridge)
        loss += f / len(nodes)
        for m, g in kr_backward(S, modes, dom).items():
            grad[m] += g / len(nodes)
    return loss, grad


def bump_factors(n, seed, sample, n_bumps=6, order=4):
    """(a, g) with g[q, m, :] = exp(-(x - c_qm)^2 / (2 w_q^2)) on x_i = i/(n-1)."""
    a, c, w = G.bump_params(seed, sample, n_bumps, order)
    x = np.arange(n) / (n - 1.0)
    g = np.exp(-((x[None, None, :] - c[:, :, None]) ** 2) / (2.0 * w[:, None, None] ** 2))
    return a, g


def bump_loss_grad(a, g, S, ridge=RIDGE):
    """Loss and gradient for the 4th-order balanced tree (sketched node {1,2}, column
    modes 3 and 4) from the bump factors alone."""
    gg = np.einsum("qmi,pmi->mqp", g, g)
    D = np.diag(a)
    Mu = gg[0] * gg[1]
    Mv = gg[2] * gg[3]
    g3 = g[:, 2, :] @ S[2].T  # (q, w)
    g4 = g[:, 3, :] @ S[3].T
    P = g3 * g4
    K1 = D @ Mu @ D
    K2 = K1 @ Mv @ K1
    w = P.shape[1]
    Wm = P.T @ K1 @ P
    Wm = Wm + ridge * np.trace(Wm) / w * np.eye(w)
    E = P.T @ K2 @ P
    Wi = np.linalg.inv(Wm)
    xsq = float(np.sum(np.outer(a, a) * Mu * Mv))
    loss = 1.0 - np.trace(Wi @ E) / xsq
    C2 = Wi @ E @ Wi
    dP = -(2.0 * K2 @ P @ Wi - 2.0 * K1 @ P @ C2 - 2.0 * ridge / w * np.trace(C2) * K1 @ P) / xsq
    grad = [np.zeros_like(s) for s in S]
    grad[2] = np.einsum("qj,qi,qj->ji", dP, g[:, 2, :], g4)
    grad[3] = np.einsum("qj,qj,qi->ji", dP, g3, g[:, 3, :])
    return loss, grad


def bump_exact_miss(a, g, w):
    """Smallest possible loss of a width-w sketch for node {1,2}: the energy fraction of
    A's singular values beyond the w-th (exact FP64 SVD of the 6 x 6 core of A)."""
    U = np.einsum("qi,qj->ijq", g[:, 0], g[:, 1]).reshape(-1, len(a))
    V = np.einsum("qi,qj->ijq", g[:, 2], g[:, 3]).reshape(-1, len(a))
    _, Ru = np.linalg.qr(U)
    _, Rv = np.linalg.qr(V)
    s = np.linalg.svd(Ru @ np.diag(a) @ Rv.T, compute_uv=False)
    e = s ** 2
    return float(e[w:].sum() / e.sum())


def dense_bump(a, g):
    """The bump tensor from its factors (flat, column-major, FP64)."""
    d = g.shape[1]
    x = 0.0
    for q in range(len(a)):
        t = a[q] * g[q, 0]
        for m in range(1, d):
            t = np.multiply.outer(t, g[q, m])
        x = x + t
    return np.asarray(x).reshape(-1, order="F")
