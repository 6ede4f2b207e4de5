"""FP64 CPU oracle for hierarchical Tucker (HT) decomposition.

TEST INFRASTRUCTURE ONLY. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the checker
(or, for ``tensorly_ht_fp32``, as the timed CPU baseline). The product path is the
HIP library under the package's ``csrc/``; it never calls into this file.

What this restates (paths relative to the reference repository; abbreviations:
3B = third_order_tensor_decompsition/CUDA/Baseline, 3O = .../CUDA/Optimized,
4B/4O = fourth_order_tensor_decomposition/single_GPU/CUDA/{Baseline,Optimized},
3P/4P = the Python directories):

* layout: column-major, mode 1 fastest; element (i1,..,id) at i1 + n1*i2 + ...
  (3O/htd_tensorcore.cu:98-104, 249; 4O/htd4_tensorcore.cu:193).
* ``unfold``: the mode-t matricisation; its column order is the one the reference's
  permutation kernels produce (3B/kernel.cu:31-69, 4O/kernel.cu:51-63).
* ``plain_ht``: the reference's plain semantics -- every non-root node's basis is the
  exact truncated SVD of its own matricisation of X (3B/htd_unopt.cu:65-88 via
  qr_svd/qr_svd_2, 3B/kernel.cu:220-333,381-491), transfer tensors by the ttm
  contraction (3O/ttm.cu:2-36, 4O/ttm.cu:2-28), root from X (3B/htd_unopt.cu:159-179).
* ``plain_ht`` with ``interior='rsvd_baseline'``: 4B/htd4.cu:77-117 -- leaves exact,
  the {1,2}/{3,4} nodes by the 20-iteration subspace rSVD of 4B/kernel.cu:367-404.
* ``rsvd_opt_reference``: 4O/kernel.cu:419-504 (ks = k+8, no power iteration, SVD of
  the first ks columns of Q^T A only) -- restated to document that quirk.
* ``ltr_ht``: leaves-to-root truncation (leaf bases first, then interior bases from
  the leaf-projected core). No reference counterpart; it is the build's fast method.
* ``reconstruct`` / ``rel_error``: 3B/htd_unopt.cu:202-266, 4O/htd4_tensorcore.cu:232-291.
* ``tensorly_ht_fp32``: the CPU-baseline semantics of 3P/HT_TensorLy.py:51-87 and
  4P/HT_4d_TensorLy.py:50-98 (Gram + eig leaves, truncated SVD interior, B by mode
  products) in FP32, with sorted ``eigh`` instead of torch.eig's unsorted order.

Conventions of the outputs (shared with the HIP library):
* ``U[t]`` is (rows_t x k_t), columns in descending singular-value order, each column
  signed so that its entry of largest magnitude is positive.
* ``B[t]`` is (k_l, k_r, k_t) with the left child's rank index fastest, i.e. stored
  column-major exactly like the reference's d_B buffers (3B/htd_unopt.cu:105-108),
  and U_t[i_l + n_l*i_r, c] = sum_ab U_l[i_l,a] U_r[i_r,b] B_t[a,b,c].
"""
import numpy as np

from . import philox


# ----------------------------------------------------------------------------- trees

class Tree:
    """Binary dimension tree whose nodes are contiguous mode ranges [lo, hi).

    Node 0 is the root (covers all modes). ``left[t]``/``right[t]`` are child node
    indices (-1 for leaves). The left child always holds the lower modes, so the row
    index of X_(t) is i_left + N_left * i_right.
    """

    def __init__(self, lo, hi, left, right):
        self.lo = list(lo)
        self.hi = list(hi)
        self.left = list(left)
        self.right = list(right)
        self.n_nodes = len(self.lo)
        self.order = self.hi[0]
        self.validate()

    def validate(self):
        if self.lo[0] != 0:
            raise ValueError("root must start at mode 0")
        seen = set()
        for t in range(self.n_nodes):
            lo, hi, l, r = self.lo[t], self.hi[t], self.left[t], self.right[t]
            if not (0 <= lo < hi):
                raise ValueError(f"node {t}: empty mode range")
            if (l < 0) != (r < 0):
                raise ValueError(f"node {t}: must have 0 or 2 children")
            if l < 0:
                if hi != lo + 1:
                    raise ValueError(f"leaf {t} must hold exactly one mode")
                if lo in seen:
                    raise ValueError(f"mode {lo} appears in two leaves")
                seen.add(lo)
            else:
                if not (self.lo[l] == lo and self.hi[l] == self.lo[r] and self.hi[r] == hi):
                    raise ValueError(f"node {t}: children do not split [{lo},{hi}) contiguously")
        if seen != set(range(self.order)):
            raise ValueError("leaves must cover every mode exactly once")

    def is_leaf(self, t):
        return self.left[t] < 0

    def leaves(self):
        return [t for t in range(self.n_nodes) if self.is_leaf(t)]

    def interior(self):
        return [t for t in range(self.n_nodes) if not self.is_leaf(t)]

    def depth(self, t):
        d = 0
        parent = self.parents()
        while t != 0:
            t = parent[t]
            d += 1
        return d

    def parents(self):
        p = [-1] * self.n_nodes
        for t in range(self.n_nodes):
            if self.left[t] >= 0:
                p[self.left[t]] = t
                p[self.right[t]] = t
        return p

    def leaf_of_mode(self, mode):
        for t in self.leaves():
            if self.lo[t] == mode:
                return t
        raise KeyError(mode)


def tree_from_spec(spec):
    """Build a Tree from a nested tuple of 0-based modes, e.g. ((0, 1), (2, 3)).

    Nodes are numbered breadth-first from the root, which matches the reference's
    heap numbering for its complete trees (3P/HT_TensorLy.py:23, 4P/HT_4d_TensorLy.py:23).
    """
    lo, hi, left, right = [], [], [], []
    queue = [spec]
    nodes = []
    while queue:
        s = queue.pop(0)
        nodes.append(s)
        if isinstance(s, tuple):
            queue.extend(s)
    index = {id(s): i for i, s in enumerate(nodes)}

    def span(s):
        if isinstance(s, tuple):
            a, _ = span(s[0])
            _, b = span(s[1])
            return a, b
        return s, s + 1

    for s in nodes:
        a, b = span(s)
        lo.append(a)
        hi.append(b)
        if isinstance(s, tuple):
            if len(s) != 2:
                raise ValueError("tree must be binary")
            left.append(index[id(s[0])])
            right.append(index[id(s[1])])
        else:
            left.append(-1)
            right.append(-1)
    return Tree(lo, hi, left, right)


def tree_reference_3d():
    """3B/3O hard-coded tree {1,2,3} -> {1,2},{3} -> {1},{2} (3P/HT_TensorLy.py:23)."""
    return tree_from_spec(((0, 1), 2))


def tree_first_vs_rest_3d():
    """Tree {1},{2,3} used by the 3rd-order benchmark workloads (8^3 ranks on 128^3, 32^3 on 1024^3)."""
    return tree_from_spec((0, (1, 2)))


def tree_balanced_4d():
    """4B/4O balanced tree {1,2},{3,4} (4P/HT_4d_TensorLy.py:23)."""
    return tree_from_spec(((0, 1), (2, 3)))


# ------------------------------------------------------------------------- helpers

def prod(xs):
    p = 1
    for x in xs:
        p *= int(x)
    return p


def node_rows(dims, tree, t):
    return prod(dims[tree.lo[t]:tree.hi[t]])


def unfold(x, dims, lo, hi):
    """Matricisation X_(t) for t = modes [lo, hi) of a column-major tensor.

    Returns the (M, P*S) matrix with M = prod(dims[lo:hi]) rows whose column index is
    p + P*s (p over the modes before lo, s over the modes from hi on). This is exactly
    the order the reference's permutation kernels produce, e.g. tensorToMode2 /
    tensorToMode3 (3B/kernel.cu:45-69) and mode2 (4O/kernel.cu:51-63).
    """
    P = prod(dims[:lo])
    M = prod(dims[lo:hi])
    S = prod(dims[hi:])
    y = np.asarray(x).reshape((P, M, S), order="F")
    return y.transpose(1, 0, 2).reshape((M, P * S), order="F")


def fold(mat, dims, lo, hi):
    """Inverse of ``unfold``: returns the flat column-major tensor."""
    P = prod(dims[:lo])
    M = prod(dims[lo:hi])
    S = prod(dims[hi:])
    y = np.asarray(mat).reshape((M, P, S), order="F").transpose(1, 0, 2)
    return y.reshape(-1, order="F")


def canon_signs(U):
    """Sign each column so that its largest-magnitude entry is positive (ties: first)."""
    U = np.array(U, copy=True)
    if U.size == 0:
        return U
    idx = np.argmax(np.abs(U), axis=0)
    s = np.sign(U[idx, np.arange(U.shape[1])])
    s[s == 0] = 1.0
    return U * s


def top_left_singular(A, k):
    """Exact truncated SVD (FP64 LAPACK): top-k left singular vectors and values. When k
    exceeds min(A.shape) the basis is completed with orthonormal null-space directions
    (zero singular values)."""
    A = np.asarray(A, dtype=np.float64)
    U, s, _ = np.linalg.svd(A, full_matrices=k > min(A.shape))
    s = np.concatenate([s, np.zeros(max(0, k - s.size))])
    return canon_signs(U[:, :k]), s[:k]


def gram_topk(G, k):
    """Top-k eigenpairs of a symmetric PSD matrix, descending (FP64 LAPACK syevd)."""
    w, V = np.linalg.eigh(np.asarray(G, dtype=np.float64))
    order = np.argsort(w)[::-1][:k]
    return canon_signs(V[:, order]), w[order]


def contract_transfer(Ut, Ul, Ur, nl, nr):
    """B_t = (U_l (x) U_r)^T U_t reshaped (k_l, k_r, k_t) -- the ttm of 3O/ttm.cu:2-36."""
    kt = Ut.shape[1]
    Ut3 = Ut.reshape((nl, nr, kt), order="F")
    return np.einsum("ia,jb,ijc->abc", Ul, Ur, Ut3)


def expand_node(Bt, Ul, Ur):
    """U_t = (U_r (x) U_l) B_t as an (n_l*n_r, k_t) matrix (row index i_l + n_l*i_r)."""
    nl, nr = Ul.shape[0], Ur.shape[0]
    kt = Bt.shape[2]
    return np.einsum("ia,jb,abc->ijc", Ul, Ur, Bt).reshape((nl * nr, kt), order="F")


class HT:
    """An HT representation: leaf bases, transfer tensors and (optionally) interior bases."""

    def __init__(self, dims, tree, ranks):
        self.dims = list(dims)
        self.tree = tree
        self.ranks = list(ranks)
        self.U = {}
        self.B = {}
        self.sv = {}

    def full_basis(self, t):
        tr = self.tree
        if tr.is_leaf(t):
            return self.U[t]
        Ul = self.full_basis(tr.left[t])
        Ur = self.full_basis(tr.right[t])
        return expand_node(self.B[t], Ul, Ur)

    def reconstruct(self):
        """X_HT (flat, column-major) -- 3B/htd_unopt.cu:202-247, 4O/htd4_tensorcore.cu:233-281."""
        tr = self.tree
        Ul = self.full_basis(tr.left[0])
        Ur = self.full_basis(tr.right[0])
        M = Ul @ self.B[0][:, :, 0] @ Ur.T
        return M.reshape(-1, order="F")


def rel_error(x, xr):
    """||X - X_r||_F / ||X||_F in FP64 -- 3B/htd_unopt.cu:258-266."""
    x = np.asarray(x, dtype=np.float64)
    xr = np.asarray(xr, dtype=np.float64)
    return float(np.linalg.norm(x - xr) / np.linalg.norm(x))


def validate_ranks(dims, tree, ranks):
    if len(ranks) != tree.n_nodes:
        raise ValueError("one rank per node")
    if ranks[0] != 1:
        raise ValueError("root rank must be 1 (k[0]=1, 3B/ht_init.cu:16-19)")
    for t in range(1, tree.n_nodes):
        k = ranks[t]
        if k < 1:
            raise ValueError(f"node {t}: rank must be >= 1")
        if k > node_rows(dims, tree, t):
            raise ValueError(f"node {t}: rank {k} exceeds its row count")
        if not tree.is_leaf(t):
            kl, kr = ranks[tree.left[t]], ranks[tree.right[t]]
            if k > kl * kr:
                raise ValueError(f"node {t}: rank {k} exceeds k_left*k_right = {kl * kr}")


# -------------------------------------------------------------- reference rSVDs

def qr_q(A):
    Q, R = np.linalg.qr(A)
    return Q


def rsvd_baseline(A, k, q, omega):
    """4B/kernel.cu:367-404: block subspace iteration, no oversampling, q iterations.

    C = A W; QR(C); q times { B = A^T C; QR(B); C = A B; QR(C) }; B = C^T A;
    SVD(B) -> U_B; return C U_B (first k columns).
    """
    C = qr_q(A @ omega)
    for _ in range(q):
        Bm = qr_q(A.T @ C)
        C = qr_q(A @ Bm)
    Bm = C.T @ A
    Ub, s, _ = np.linalg.svd(Bm, full_matrices=False)
    return canon_signs((C @ Ub)[:, :k]), s[:k]


def rsvd_opt_reference(A, k, omega_ks):
    """4O/kernel.cu:419-504: ks = k+8, C = A W, QR(C), B = C^T A and gesvd applied to the
    leading ks x ks block of B only (lda = ks, :457-481), U = C U_svd, first k columns.
    Exact only when rank(A) <= ks. Restated to document the quirk; not used as the oracle.
    """
    ks = omega_ks.shape[1]
    C = qr_q(A @ omega_ks)
    Bm = C.T @ A
    Ub, s, _ = np.linalg.svd(Bm[:, :ks])
    return canon_signs((C @ Ub)[:, :k])


def rsvd_power(A, k, p, q, omega):
    """Randomised range finder with oversampling p and q power iterations, then the
    exact SVD of the projected (k+p) x n problem (Halko-Martinsson-Tropp, Alg. 4.4/5.1)."""
    Q = qr_q(A @ omega)
    for _ in range(q):
        Z = qr_q(A.T @ Q)
        Q = qr_q(A @ Z)
    Bm = Q.T @ A
    Ub, s, _ = np.linalg.svd(Bm, full_matrices=False)
    return canon_signs((Q @ Ub)[:, :k]), s[:k]


# ---------------------------------------------------------------------- plain HT

def plain_ht(x, dims, tree, ranks, interior="exact", q=20, omega_seed=1234, leaf="svd"):
    """Root-to-leaves HT: every non-root node's basis from its own matricisation of X.

    interior = "exact": exact truncated SVD (3B semantics, qr_svd 3B/kernel.cu:220-333).
    interior = "rsvd_baseline": 4B semantics (4B/htd4.cu:112-117, 4B/kernel.cu:367-404)
    with Gaussian test matrices drawn from Philox stream (omega_seed, node index).
    leaf = "svd" (qr_svd_2, 3B/kernel.cu:381-491) or "gram" (Gram + eig,
    3O/htd_tensorcore.cu:83-140); both give the exact truncated basis.
    """
    validate_ranks(dims, tree, ranks)
    x = np.asarray(x, dtype=np.float64)
    ht = HT(dims, tree, ranks)
    full = {}
    for t in range(1, tree.n_nodes):
        A = unfold(x, dims, tree.lo[t], tree.hi[t])
        k = ranks[t]
        if tree.is_leaf(t):
            if leaf == "gram":
                U, w = gram_topk(A @ A.T, k)
                s = np.sqrt(np.maximum(w, 0.0))
            else:
                U, s = top_left_singular(A, k)
        elif interior == "rsvd_baseline":
            omega = philox.normals(omega_seed, 256 + t, A.shape[1] * k).reshape((A.shape[1], k), order="F")
            U, s = rsvd_baseline(A, k, q, omega)
        else:
            U, s = top_left_singular(A, k)
        full[t] = U
        ht.sv[t] = s
        if tree.is_leaf(t):
            ht.U[t] = U
    for t in tree.interior():
        l, r = tree.left[t], tree.right[t]
        nl, nr = node_rows(dims, tree, l), node_rows(dims, tree, r)
        if t == 0:
            Xm = x.reshape((nl, nr), order="F")
            ht.B[0] = (full[l].T @ Xm @ full[r])[:, :, None]
        else:
            ht.B[t] = contract_transfer(full[t], full[l], full[r], nl, nr)
            ht.U[t] = full[t]
    return ht


# ------------------------------------------------------------------------ LTR HT

def project_modes(y, ydims, slot, V):
    """y x_slot V^T for a flat column-major tensor with dims ``ydims``."""
    A = unfold(y, ydims, slot, slot + 1)
    nd = list(ydims)
    nd[slot] = V.shape[1]
    return fold(V.T @ A, nd, slot, slot + 1), nd


def ltr_ht(x, dims, tree, ranks, leaf_groups=None):
    """Leaves-to-root HT truncation in FP64.

    1. Leaf bases: top-k eigenvectors of the Gram X_(n) X_(n)^T. ``leaf_groups`` lists
       groups of modes; the leaves of one group are computed from the current tensor,
       which is then projected onto them before the next group (None = one group of
       all modes, i.e. every leaf from X itself).
    2. Core C = X x_1 U_1^T ... x_d U_d^T.
    3. Interior nodes bottom-up (deepest level first, left to right): the node's basis
       V_t is the top-k_t left singular subspace of the current tensor matricised over
       its two child slots; B_t = V_t as (k_l, k_r, k_t); the two slots merge into one.
    4. Root: B_root is the remaining 2-slot tensor.
    """
    validate_ranks(dims, tree, ranks)
    x = np.asarray(x, dtype=np.float64)
    ht = HT(dims, tree, ranks)
    d = len(dims)
    if leaf_groups is None:
        leaf_groups = [list(range(d))]
    y = x
    ydims = list(dims)
    for group in leaf_groups:
        bases = {}
        for mode in group:
            t = tree.leaf_of_mode(mode)
            A = unfold(y, ydims, mode, mode + 1)
            U, w = gram_topk(A @ A.T, ranks[t])
            bases[mode] = U
            ht.U[t] = U
            ht.sv[t] = np.sqrt(np.maximum(w, 0.0))
        for mode in group:
            y, ydims = project_modes(y, ydims, mode, bases[mode])
    # slots: list of node ids in mode order
    slots = [tree.leaf_of_mode(m) for m in range(d)]
    todo = sorted([t for t in tree.interior() if t != 0], key=lambda t: (-tree.depth(t), tree.lo[t]))
    for t in todo:
        l, r = tree.left[t], tree.right[t]
        i = slots.index(l)
        assert slots[i + 1] == r
        A = unfold(y, ydims, i, i + 2)
        V, s = top_left_singular(A, ranks[t])
        ht.B[t] = V.reshape((ydims[i], ydims[i + 1], ranks[t]), order="F")
        ht.sv[t] = s
        y = fold(V.T @ A, ydims[:i] + [ranks[t]] + ydims[i + 2:], i, i + 1)
        ydims = ydims[:i] + [ranks[t]] + ydims[i + 2:]
        slots = slots[:i] + [t] + slots[i + 2:]
    assert len(slots) == 2
    ht.B[0] = y.reshape((ydims[0], ydims[1], 1), order="F")
    return ht


# ---------------------------------------------------- TensorLy CPU-baseline port

def tensorly_ht_fp32(x, dims, tree, ranks, progress=None):
    """FP32 CPU restatement of the reference's TensorLy path (3P/HT_TensorLy.py:51-87,
    4P/HT_4d_TensorLy.py:50-98): leaves by Gram X_(n) X_(n)^T + symmetric eig (top-k,
    sorted -- torch.eig's unsorted order is a documented deviation), interior nodes by
    a truncated SVD of the matricisation, B_t by mode products with the child bases,
    then reconstruction and the relative error. Returns (HT, rel_err). progress, if
    given, is called with a short text after each node and stage.
    """
    x32 = np.asarray(x, dtype=np.float32)
    ht = HT(dims, tree, ranks)
    full = {}
    for t in range(tree.n_nodes - 1, 0, -1):
        A = unfold(x32, dims, tree.lo[t], tree.hi[t])
        k = ranks[t]
        if tree.is_leaf(t):
            w, V = np.linalg.eigh(A @ A.T)
            U = V[:, ::-1][:, :k]
        else:
            if A.shape[0] <= A.shape[1]:
                w, V = np.linalg.eigh(A @ A.T)
                U = V[:, ::-1][:, :k]
            else:
                w, V = np.linalg.eigh(A.T @ A)
                Vk = V[:, ::-1][:, :k]
                U = A @ Vk
                U /= np.linalg.norm(U, axis=0, keepdims=True)
        full[t] = U.astype(np.float32)
        if tree.is_leaf(t):
            ht.U[t] = full[t]
        if progress:
            progress(f"node {t} ({A.shape[0]} x {A.shape[1]}) done")
    for t in tree.interior():
        l, r = tree.left[t], tree.right[t]
        nl, nr = node_rows(dims, tree, l), node_rows(dims, tree, r)
        if t == 0:
            Xm = x32.reshape((nl, nr), order="F")
            ht.B[0] = (full[l].T @ Xm @ full[r])[:, :, None]
        else:
            ht.B[t] = contract_transfer(full[t], full[l], full[r], nl, nr).astype(np.float32)
    if progress:
        progress("transfer tensors done")
    xr = ht.reconstruct().astype(np.float32)
    err = float(np.linalg.norm(x32 - xr) / np.linalg.norm(x32))
    if progress:
        progress("reconstruction and error done")
    return ht, err


# --------------------------------------------------------------------- bounds

def ht_error_bound(x, dims, tree, ranks):
    """Grasedyck's root-to-leaves bound: ||X - X_HT||^2 <= sum_t sum_{i>k_t} s_i(X_(t))^2
    over every non-root node except one child of the root (the two children of the
    root share their singular values). Returns the bound on the relative error."""
    x = np.asarray(x, dtype=np.float64)
    total = 0.0
    for t in range(1, tree.n_nodes):
        if t == tree.right[0]:
            continue
        A = unfold(x, dims, tree.lo[t], tree.hi[t])
        s = np.linalg.svd(A, compute_uv=False)
        total += float(np.sum(s[ranks[t]:] ** 2))
    return float(np.sqrt(total) / np.linalg.norm(x))


# ------------------------------------------------------------- randomised RTL HT

def rsvd_ht(x, dims, tree, ranks, p=8, q=1, seed=1234, max_block=64):
    """Root-to-leaves randomised HT as the engine's HT_METHOD_RSVD runs it (the 4O
    structure, fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/htd4_tensorcore.cu:
    174-228, with oversampling p, q power iterations and the full projected SVD):
    leaves by the Gram eigenvectors of X_(n); each interior node t whose mode range starts
    at mode 0 or ends at the last mode by a range finder on X_(t) with the FP32-rounded
    Philox test matrix of stream 256+t (l = min(k + p, 64, rows, cols) columns, k the
    larger of the node's and its complementary sibling's rank); the complementary sibling
    of a root child takes the right singular vectors of the same projected SVD. A node
    strictly inside the mode order (order >= 5 trees) takes the Gram eigenvectors of its
    X_(t), as the leaves do."""
    validate_ranks(dims, tree, ranks)
    x = np.asarray(x, dtype=np.float64)
    ht = HT(dims, tree, ranks)
    full = {}
    for t in tree.leaves():
        A = unfold(x, dims, tree.lo[t], tree.hi[t])
        U, w = gram_topk(A @ A.T, ranks[t])
        full[t] = ht.U[t] = U
        ht.sv[t] = np.sqrt(np.maximum(w, 0))
    parent = tree.parents()
    done = set()
    for t in range(1, tree.n_nodes):
        if tree.is_leaf(t) or t in done:
            continue
        A = unfold(x, dims, tree.lo[t], tree.hi[t])
        rows, cols = A.shape
        if not (tree.lo[t] == 0 or tree.hi[t] == len(dims)):
            U, w = gram_topk(A @ A.T, ranks[t])
            full[t] = ht.U[t] = U
            ht.sv[t] = np.sqrt(np.maximum(w, 0))
            done.add(t)
            continue
        sib = -1
        if parent[t] == 0:
            sib = tree.right[0] if tree.left[0] == t else tree.left[0]
        kW = ranks[sib] if sib >= 0 and not tree.is_leaf(sib) else 0
        k = ranks[t]
        l = int(min(max(k, kW) + p, max_block, rows, cols))
        om = philox.normals(seed, 256 + t, cols * l).astype(np.float32).astype(np.float64)
        om = om.reshape((cols, l), order="F")
        Q = qr_q(A @ om)
        for _ in range(q):
            Qz = qr_q(A.T @ Q)
            Q = qr_q(A @ Qz)
        Ub, s, Vbt = np.linalg.svd(Q.T @ A, full_matrices=False)
        full[t] = ht.U[t] = canon_signs((Q @ Ub)[:, :k])
        ht.sv[t] = s[:k]
        if kW:
            full[sib] = ht.U[sib] = canon_signs(Vbt.T[:, :kW])
            ht.sv[sib] = s[:kW]
            done.add(sib)
        done.add(t)
    for t in tree.interior():
        l, r = tree.left[t], tree.right[t]
        nl, nr = node_rows(dims, tree, l), node_rows(dims, tree, r)
        if t == 0:
            ht.B[0] = (full[l].T @ x.reshape((nl, nr), order="F") @ full[r])[:, :, None]
        else:
            ht.B[t] = contract_transfer(full[t], full[l], full[r], nl, nr)
    return ht


# ------------------------------------------------------------- rank-adaptive truncation

def choose_rank(energies, trace, tau2, cap):
    """Smallest r <= cap with trace - sum(energies[:r]) <= tau2 (energies descending), else cap."""
    kept = 0.0
    for i in range(min(cap, len(energies))):
        kept += max(float(energies[i]), 0.0)
        if trace - kept <= tau2:
            return i + 1
    return cap


def _nest(tree, ranks):
    """k_t <= k_l k_r, deepest interior nodes first."""
    inner = sorted([t for t in tree.interior() if t != 0], key=lambda t: -tree.depth(t))
    for t in inner:
        ranks[t] = min(ranks[t], ranks[tree.left[t]] * ranks[tree.right[t]])
    return ranks


def adaptive_caps(tree, max_ranks, max_rank=0):
    caps = list(max_ranks)
    for t in range(1, tree.n_nodes):
        if max_rank > 0:
            caps[t] = min(caps[t], max_rank)
    return _nest(tree, caps)


def plain_ht_adaptive(x, dims, tree, max_ranks, rel_eps, max_rank=0):
    """The rel_eps/max_rank truncation the reference's htd declares and ignores
    (third_order_tensor_decompsition/CUDA/Baseline/function.h:6), on the root-to-leaves
    (plain) HT: each non-root node's rank is the smallest within its cap whose SVD tail of
    X_(t) is at most tau^2 = (rel_eps ||X||)^2 / (2d - 3); the two root children share one
    spectrum and get one rank; then k_t <= k_l k_r. Returns (HT, ranks)."""
    x = np.asarray(x, dtype=np.float64)
    caps = adaptive_caps(tree, max_ranks, max_rank)
    xsq = float(np.sum(x * x))
    tau2 = rel_eps ** 2 * xsq / max(1, 2 * len(dims) - 3)
    ranks = list(caps)
    for t in range(1, tree.n_nodes):
        s = np.linalg.svd(unfold(x, dims, tree.lo[t], tree.hi[t]), compute_uv=False)
        ranks[t] = max(1, choose_rank(s ** 2, xsq, tau2, caps[t]))
    cl, cr = tree.left[0], tree.right[0]
    common = max(ranks[cl], ranks[cr])
    ranks[cl], ranks[cr] = min(common, caps[cl]), min(common, caps[cr])
    ranks = _nest(tree, ranks)
    return plain_ht(x, dims, tree, ranks), ranks


def ltr_ht_adaptive(x, dims, tree, max_ranks, rel_eps, max_rank=0, leaf_groups=None):
    """Leaves-to-root HT (ltr_ht) with the same per-node rule, the energy of each truncation
    measured in the tensor it is applied to (X for the first leaf group, the projected tensor
    for later groups, the core for interior nodes), tau^2 = (rel_eps ||X||)^2 / (2d - 2): the
    removed energies of leaves-to-root truncation add up exactly. Returns (HT, ranks)."""
    x = np.asarray(x, dtype=np.float64)
    d = len(dims)
    caps = adaptive_caps(tree, max_ranks, max_rank)
    ranks = list(caps)
    tau2 = rel_eps ** 2 * float(np.sum(x * x)) / max(1, 2 * d - 2)
    ht = HT(dims, tree, ranks)
    if leaf_groups is None:
        leaf_groups = [list(range(d))]
    y, ydims = x, list(dims)
    for group in leaf_groups:
        bases = {}
        energy = float(np.sum(y * y))
        for mode in group:
            t = tree.leaf_of_mode(mode)
            A = unfold(y, ydims, mode, mode + 1)
            U, w = gram_topk(A @ A.T, caps[t])
            ranks[t] = max(1, choose_rank(w, energy, tau2, caps[t]))
            bases[mode] = U[:, :ranks[t]]
            ht.U[t] = bases[mode]
            ht.sv[t] = np.sqrt(np.maximum(w[:ranks[t]], 0.0))
        for mode in group:
            y, ydims = project_modes(y, ydims, mode, bases[mode])
    slots = [tree.leaf_of_mode(m) for m in range(d)]
    todo = sorted([t for t in tree.interior() if t != 0], key=lambda t: (-tree.depth(t), tree.lo[t]))
    for t in todo:
        l, r = tree.left[t], tree.right[t]
        i = slots.index(l)
        A = unfold(y, ydims, i, i + 2)
        cap = min(caps[t], A.shape[0], A.shape[1])
        V, s = top_left_singular(A, cap)
        ranks[t] = max(1, choose_rank(s ** 2, float(np.sum(y * y)), tau2, cap))
        V, s = V[:, :ranks[t]], s[:ranks[t]]
        ht.B[t] = V.reshape((ydims[i], ydims[i + 1], ranks[t]), order="F")
        ht.sv[t] = s
        y = fold(V.T @ A, ydims[:i] + [ranks[t]] + ydims[i + 2:], i, i + 1)
        ydims = ydims[:i] + [ranks[t]] + ydims[i + 2:]
        slots = slots[:i] + [t] + slots[i + 2:]
    ht.B[0] = y.reshape((ydims[0], ydims[1], 1), order="F")
    ht.ranks = ranks
    return ht, ranks
