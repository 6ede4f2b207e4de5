"""On-disk HT format (NPZ) and the hand-off to HT-layer code.

The reference keeps no decomposition: its drivers free U_t and B_t before returning
(third_order_tensor_decompsition/CUDA/Baseline/htd_unopt.cu:273-296), and its HT-layer
application builds the factors in memory and hands them to the layer as one tuple
(HT_tensor_layer/FC_MNIST/mnist_ht.py:194-196: leaf bases, the {1,2} transfer tensor and the
root transfer matrix). This module defines a file for the engine's output and that tuple.

Format "ht-npz/1" (numpy .npz, no pickled objects):
  format              the string "ht-npz/1"
  dims                int64 [d]          mode sizes (column-major tensor, mode 1 fastest)
  lo, hi, left, right int32 [2d-1]       the dimension tree (node 0 = root, breadth-first)
  ranks               int32 [2d-1]       ranks[0] = 1
  rel_err             float64 []         relative error of the decomposition (NaN if unknown)
  U<t>                float32 [rows_t, k_t]   basis of leaf t (and of interior t, when kept)
  B<t>                float32 [k_l, k_r, k_t] transfer tensor of interior t (B0: k_t = 1)
  sv<t>               float32 [k_t]      singular values of non-root node t
U<t> and B<t> are stored as numpy arrays of those shapes with the usual meaning of the
indices: U_t[i_l + n_l * i_r, c] = sum_ab U_l[i_l, a] U_r[i_r, b] B_t[a, b, c].
"""
import numpy as np

from .tree import Tree

FORMAT = "ht-npz/1"


def _host(a):
    if hasattr(a, "detach"):
        import torch
        if a.is_cuda:
            torch.cuda.synchronize(a.device)
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _rows(dims, tree, t):
    r = 1
    for m in range(tree.lo[t], tree.hi[t]):
        r *= int(dims[m])
    return r


def save(path, res):
    """Write an HTensor (engine.decompose's result, device or host) to `path` (.npz)."""
    tr = res.tree
    out = {
        "format": np.array(FORMAT),
        "dims": np.asarray(res.dims, dtype=np.int64),
        "lo": np.asarray(tr.lo, dtype=np.int32), "hi": np.asarray(tr.hi, dtype=np.int32),
        "left": np.asarray(tr.left, dtype=np.int32), "right": np.asarray(tr.right, dtype=np.int32),
        "ranks": np.asarray(res.ranks, dtype=np.int32),
        "rel_err": np.array(float(res.rel_err), dtype=np.float64),
    }
    for t in res.U:
        out[f"U{t}"] = np.ascontiguousarray(_host(res.U[t]), dtype=np.float32)
    for t in res.B:
        out[f"B{t}"] = np.ascontiguousarray(_host(res.B[t]), dtype=np.float32)
    for t in res.sv:
        out[f"sv{t}"] = np.ascontiguousarray(_host(res.sv[t]), dtype=np.float32)
    np.savez(path, **out)


class HTFile:
    """A decomposition read back from an ht-npz/1 file (numpy arrays on the host)."""

    def __init__(self, dims, tree, ranks, U, B, sv, rel_err):
        self.dims, self.tree, self.ranks = dims, tree, ranks
        self.U, self.B, self.sv, self.rel_err = U, B, sv, rel_err

    @property
    def want_interior_u(self):
        return any(not self.tree.is_leaf(t) for t in self.U)

    def packed(self):
        """The engine's packed FP32 block (ht_output_layout order) of this decomposition."""
        from .engine import output_layout
        u_off, b_off, s_off, total = output_layout(self.dims, self.tree, self.ranks, self.want_interior_u)
        p = np.zeros(total, dtype=np.float32)
        for t in range(self.tree.n_nodes):
            if u_off[t] >= 0:
                p[u_off[t]:u_off[t] + self.U[t].size] = self.U[t].reshape(-1, order="F")
            if b_off[t] >= 0:
                p[b_off[t]:b_off[t] + self.B[t].size] = self.B[t].reshape(-1, order="F")
            if s_off[t] >= 0 and t in self.sv:
                p[s_off[t]:s_off[t] + self.sv[t].size] = self.sv[t]
        return p

    def to_device(self, device=0):
        """An engine.HTensor over a device copy of the packed block (for ht.residual and
        ht.reconstruct)."""
        import torch
        from .engine import HTensor, _unpack, output_layout
        pk = torch.from_numpy(self.packed()).to(f"cuda:{device}")
        res = HTensor(dims=list(self.dims), tree=self.tree, ranks=list(self.ranks), packed=pk,
                      rel_err=self.rel_err, want_interior_u=self.want_interior_u)
        _unpack(res, output_layout(self.dims, self.tree, self.ranks, self.want_interior_u))
        return res

    def full_basis(self, t):
        """U_t of any node, expanded from the leaves through the transfer tensors (FP64)."""
        tr = self.tree
        if tr.is_leaf(t):
            return self.U[t].astype(np.float64)
        Ul, Ur = self.full_basis(tr.left[t]), self.full_basis(tr.right[t])
        E = np.einsum("ia,jb,abc->ijc", Ul, Ur, self.B[t].astype(np.float64))
        return E.reshape((Ul.shape[0] * Ur.shape[0], E.shape[2]), order="F")

    def reconstruct(self):
        """X_HT as a flat column-major FP64 array (host)."""
        tr = self.tree
        Ul, Ur = self.full_basis(tr.left[0]), self.full_basis(tr.right[0])
        return (Ul @ self.B[0][:, :, 0].astype(np.float64) @ Ur.T).reshape(-1, order="F")

    def layer_factors(self):
        """(U_{1}, U_{2}, U_{3}, B_{1,2}, B_root) of a 3rd-order decomposition on the tree
        {1,2},{3} -- the tuple the reference's HT layer is built from
        (HT_tensor_layer/FC_MNIST/mnist_ht.py:194-196, which reshapes each leaf basis
        (N_i M_i, k) to (N_i, M_i, k) for its input/output split): leaf bases (n_i, k_i),
        B_{1,2} (k_1, k_2, k_{12}) and B_root (k_{12}, k_3)."""
        tr = self.tree
        if len(self.dims) != 3 or tr.lo[tr.left[0]] != 0 or tr.hi[tr.left[0]] != 2:
            raise ValueError("layer_factors needs a 3rd-order decomposition on the tree {1,2},{3}")
        t12, t3 = tr.left[0], tr.right[0]
        t1, t2 = tr.left[t12], tr.right[t12]
        return (self.U[t1], self.U[t2], self.U[t3], self.B[t12], self.B[0][:, :, 0])


def load(path):
    """Read an ht-npz/1 file (no pickled objects are loaded)."""
    with np.load(path, allow_pickle=False) as z:
        if str(z["format"]) != FORMAT:
            raise ValueError(f"{path}: not an {FORMAT} file")
        dims = [int(v) for v in z["dims"]]
        tree = Tree([int(v) for v in z["lo"]], [int(v) for v in z["hi"]], [int(v) for v in z["left"]],
                    [int(v) for v in z["right"]]).validate()
        ranks = [int(v) for v in z["ranks"]]
        U, B, sv = {}, {}, {}
        for t in range(tree.n_nodes):
            if f"U{t}" in z:
                U[t] = z[f"U{t}"].astype(np.float32)
                if U[t].shape != (_rows(dims, tree, t), ranks[t]):
                    raise ValueError(f"{path}: U{t} has shape {U[t].shape}")
            if f"B{t}" in z:
                B[t] = z[f"B{t}"].astype(np.float32)
                if B[t].shape != (ranks[tree.left[t]], ranks[tree.right[t]], ranks[t]):
                    raise ValueError(f"{path}: B{t} has shape {B[t].shape}")
            if f"sv{t}" in z:
                sv[t] = z[f"sv{t}"].astype(np.float32)
        for t in range(tree.n_nodes):
            if tree.is_leaf(t) and t not in U:
                raise ValueError(f"{path}: leaf {t} has no basis")
            if not tree.is_leaf(t) and t not in B:
                raise ValueError(f"{path}: node {t} has no transfer tensor")
        return HTFile(dims, tree, ranks, U, B, sv, float(z["rel_err"]))
