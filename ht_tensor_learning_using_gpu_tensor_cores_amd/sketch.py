"""Learned per-mode sketch: the test matrix of the randomised method, trained on the device.

No counterpart in the reference, whose range finder draws a dense Gaussian test matrix
(fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/kernel.cu:419-444). Mode m
owns a sketch S_m (width x n_m); the test matrix of a node the randomised method sketches
is the Khatri-Rao product of the sketches of its matricisation's column modes,

    Omega_t[c, j] = prod_{m in cols(t)} S_m[j, i_m(c)],

so a 128^4 tensor's {1,2} node (16384 columns) is sketched by two 20 x 128 matrices. The
sketches are trained with Adam on the loss

    loss = mean over sketched nodes of 1 - ||P_Y X_(t)||_F^2 / ||X||_F^2,  Y = X_(t) Omega_t,

on Gaussian-bump tensors the library generates on the device; the forward and backward
passes are the library's own kernels (ht_sketch_loss_grad / ht_sketch_train_bumps). The
trained sketch is stored rounded to BF16.

    sk = sketch.gaussian([128] * 4, 20)                      # untrained starting point
    sk, curve = sketch.train_bumps(128, 4, tree.balanced_4d(), 20, count=2000, init=sk)
    res = decompose(x, dims, tree, ranks, method="rsvd", power_iters=0, sketch=sk)
"""
import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import default_handle
from .tree import Tree


def _torch():
    import torch
    return torch


def layout(dims, width):
    """(offsets per mode, total floats) of the flat sketch array."""
    lib = _lib.load()
    off = (ctypes.c_int64 * len(dims))()
    tot = ctypes.c_int64()
    _lib.check(lib.ht_sketch_layout(len(dims), _lib.i64_array(dims), int(width), off, ctypes.byref(tot)),
               "ht_sketch_layout")
    return list(off), tot.value


@dataclass
class Sketch:
    """Per-mode sketches on the device: `weights` is the flat FP32 array (mode m's
    width x n_m row-major block at offsets[m])."""
    dims: list
    width: int
    weights: object

    @property
    def offsets(self):
        return layout(self.dims, self.width)[0]

    def mode(self, m):
        o = self.offsets[m]
        return self.weights[o:o + self.width * self.dims[m]].view(self.width, self.dims[m])

    def bf16(self):
        """The same sketch with every weight rounded to the nearest BF16 value."""
        torch = _torch()
        return Sketch(list(self.dims), self.width, self.weights.to(torch.bfloat16).to(torch.float32))

    def clone(self):
        return Sketch(list(self.dims), self.width, self.weights.clone())

    def save(self, path):
        """npz with the dims, the width and the weights as BF16 bit patterns (uint16)."""
        torch = _torch()
        bits = self.weights.to(torch.bfloat16).view(torch.int16).cpu().numpy().view(np.uint16)
        np.savez(path, dims=np.asarray(self.dims, dtype=np.int64), width=np.int64(self.width), weights_bf16=bits)

    @staticmethod
    def load(path, device=0):
        torch = _torch()
        with np.load(path, allow_pickle=False) as f:
            dims = [int(d) for d in f["dims"]]
            width = int(f["width"])
            bits = f["weights_bf16"].astype(np.uint32) << 16
        w = torch.from_numpy(bits.view(np.float32).copy()).to(f"cuda:{device}")
        return Sketch(dims, width, w)


_BUNDLED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sketches")


def bundled(name, device=0):
    """A sketch trained by `profile_run --workload c5` and shipped with the package:
    "bumps128_w20" (width 20) or "bumps128_w4" (width 4), both for 128^4 bump tensors."""
    return Sketch.load(os.path.join(_BUNDLED, f"sketch_{name}.npz"), device=device)


def gaussian(dims, width, seed=7, device=0, handle=None):
    """i.i.d. N(0,1) sketches: the untrained starting point (ht_sketch_init_gaussian)."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle(device)
    dims = [int(d) for d in dims]
    _, tot = layout(dims, width)
    w = torch.empty(tot, dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.synchronize(w.device)
    _lib.check(lib.ht_sketch_init_gaussian(h.ptr, len(dims), _lib.i64_array(dims), int(width), int(seed),
                                           ctypes.c_void_p(w.data_ptr())), "ht_sketch_init_gaussian")
    return Sketch(dims, int(width), w)


def _tree_args(tree: Tree):
    return (tree.n_nodes, _lib.i32_array(tree.lo), _lib.i32_array(tree.hi), _lib.i32_array(tree.left),
            _lib.i32_array(tree.right))


def loss_grad(x, dims, tree: Tree, sk: Sketch, want_grad=True, handle=None):
    """Loss of one device tensor and (when want_grad) its gradient, a flat tensor in the
    sketch's layout (ht_sketch_loss_grad)."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle()
    dims = [int(d) for d in dims]
    g = torch.empty_like(sk.weights) if want_grad else None
    loss = ctypes.c_double()
    torch.cuda.synchronize(x.device)
    _lib.check(lib.ht_sketch_loss_grad(h.ptr, ctypes.c_void_p(x.data_ptr()), len(dims), _lib.i64_array(dims),
                                       *_tree_args(tree), int(sk.width), ctypes.c_void_p(sk.weights.data_ptr()),
                                       ctypes.c_void_p(g.data_ptr()) if g is not None else None,
                                       ctypes.byref(loss)), "ht_sketch_loss_grad")
    return loss.value, g


def train_bumps(n, order, tree: Tree, width, count, seed=4, first=0, epochs=2, batch=8, lr=0.05, n_bumps=6,
                init=None, init_seed=7, min_loss=1e-5, handle=None):
    """Train per-mode sketches on `count` bump tensors (samples first.. of `seed`); steps
    whose mean loss is below `min_loss` are skipped (their gradient is rounding noise).
    Returns (trained Sketch in FP32, per-epoch mean training loss)."""
    lib = _lib.load()
    h = handle or default_handle()
    dims = [int(n)] * int(order)
    sk = init.clone() if init is not None else gaussian(dims, width, seed=init_seed)
    if sk.width != width or list(sk.dims) != dims:
        raise ValueError("initial sketch has the wrong shape")
    o = _lib.SketchTrainOptions()
    lib.ht_sketch_default_train_options(ctypes.byref(o))
    o.width, o.epochs, o.batch, o.lr = int(width), int(epochs), int(batch), float(lr)
    o.min_loss = float(min_loss)
    curve = (ctypes.c_double * int(epochs))()
    _lib.check(lib.ht_sketch_train_bumps(h.ptr, int(order), int(n), int(n_bumps), int(seed), int(first), int(count),
                                         *_tree_args(tree), ctypes.byref(o), ctypes.c_void_p(sk.weights.data_ptr()),
                                         curve), "ht_sketch_train_bumps")
    return sk, np.array(curve[:])


def eval_bumps(sk: Sketch, n, order, tree: Tree, count, seed=5, first=0, n_bumps=6, handle=None):
    """Per-tensor loss of a sketch on held-out bump tensors (ht_sketch_eval_bumps)."""
    lib = _lib.load()
    h = handle or default_handle()
    out = (ctypes.c_double * int(count))()
    _lib.check(lib.ht_sketch_eval_bumps(h.ptr, int(order), int(n), int(n_bumps), int(seed), int(first), int(count),
                                        *_tree_args(tree), int(sk.width), ctypes.c_void_p(sk.weights.data_ptr()),
                                        out), "ht_sketch_eval_bumps")
    return np.array(out[:])
