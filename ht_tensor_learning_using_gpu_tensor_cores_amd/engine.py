"""Python front end of the HT engine: a thin layer over the C ABI.

Tensors are column-major (mode 1 fastest), as in the reference
(third_order_tensor_decompsition/CUDA/Optimized/htd_tensorcore.cu:98-104). Device
buffers are torch tensors on a ROCm device (torch is used for memory and streams only;
every computation runs in libht_amd.so). Host inputs are numpy arrays.
"""
import ctypes
import math
from collections.abc import Mapping
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import _lib
from .tree import Tree


def _torch():
    import torch
    return torch


class Handle:
    """Owns a HIP stream, a device workspace and events (ht_create / ht_destroy)."""

    def __init__(self, device=0):
        self.lib = _lib.load()
        self._h = ctypes.c_void_p()
        _lib.check(self.lib.ht_create(ctypes.byref(self._h), int(device)), "ht_create")
        self.device = int(device)
        self._ext = None
        self._ev = None

    def order_after_torch(self, device):
        """Queue a wait of the handle's stream for what torch's current stream of `device`
        holds so far (an event wait on the device; the host does not block)."""
        torch = _torch()
        dev = torch.device(device)
        if dev.index is not None and dev.index != self.device:
            torch.cuda.synchronize(dev)
            return
        cs = torch.cuda.current_stream(dev)
        if cs.query():  # nothing queued there: no ordering needed
            return
        if self._ext is None:
            self._ext = torch.cuda.ExternalStream(self.lib.ht_get_stream(self.ptr), device=dev)
            self._ev = torch.cuda.Event()
        self._ev.record(cs)
        self._ext.wait_event(self._ev)

    @property
    def ptr(self):
        if self._h is None or not self._h.value:
            raise RuntimeError("handle is closed")
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            self.lib.ht_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = {}


def default_handle(device=0):
    h = _default.get(device)
    if h is None:
        h = Handle(device)
        _default[device] = h
    return h


def make_options(method="ltr", precision="bf16", err="residual", oversample=8, power_iters=1, seed=1234,
                 eig_tol=1e-10, max_eig_iters=200, want_interior_u=False, sketch=None, sketch_tol=1e-3,
                 check_finite=1):
    """ht_options. `sketch` (a sketch.Sketch on the device) selects the learned Khatri-Rao
    test matrix of the randomised method, with the Gaussian one as its fallback above
    `sketch_tol`; precision "fp8" (randomised method only) keeps its FP8 range finder
    within the same `sketch_tol`, else redoes it in BF16."""
    lib = _lib.load()
    o = _lib.Options()
    lib.ht_default_options(ctypes.byref(o))
    if sketch is not None:
        o.sketch = _lib.SKETCH["learned"]
        o.sketch_width = int(sketch.width)
        o.sketch_weights = ctypes.c_void_p(sketch.weights.data_ptr())
    o.sketch_tol = float(sketch_tol)
    o.method = _lib.METHOD[method]
    o.precision = _lib.PRECISION[precision]
    o.err_mode = _lib.ERR_MODE[err]
    o.oversample = int(oversample)
    o.power_iters = int(power_iters)
    o.seed = int(seed)
    o.eig_tol = float(eig_tol)
    o.max_eig_iters = int(max_eig_iters)
    o.want_interior_u = int(bool(want_interior_u))
    o.check_finite = int(bool(check_finite))
    return o


_LAYOUTS = {}
_OPTIONS = {}  # ht_options per call signature (the C side copies them; read-only here)


def _layout_entry(dims, tree: Tree, ranks, want_interior_u):
    """Validated layout and the ctypes arguments of one (dims, tree, ranks), made once."""
    key = (tuple(dims), tuple(tree.lo), tuple(tree.hi), tuple(tree.left), tuple(tree.right), tuple(ranks),
           bool(want_interior_u))
    e = _LAYOUTS.get(key)
    if e is None:
        tree.validate()
        layout = output_layout(dims, tree, ranks, want_interior_u)
        args = (_lib.i64_array(dims), tree.n_nodes, _lib.i32_array(tree.lo), _lib.i32_array(tree.hi),
                _lib.i32_array(tree.left), _lib.i32_array(tree.right), _lib.i32_array(ranks))
        if len(_LAYOUTS) >= 256:
            _LAYOUTS.clear()
        e = _LAYOUTS[key] = (layout, args)
    return e


def output_layout(dims, tree: Tree, ranks, want_interior_u=False):
    lib = _lib.load()
    n = tree.n_nodes
    u = (ctypes.c_int64 * n)()
    b = (ctypes.c_int64 * n)()
    s = (ctypes.c_int64 * n)()
    tot = ctypes.c_int64()
    _lib.check(lib.ht_output_layout(len(dims), _lib.i64_array(dims), n, _lib.i32_array(tree.lo),
                                    _lib.i32_array(tree.hi), _lib.i32_array(tree.left), _lib.i32_array(tree.right),
                                    _lib.i32_array(ranks), int(bool(want_interior_u)), u, b, s,
                                    ctypes.byref(tot)), "ht_output_layout")
    return list(u), list(b), list(s), tot.value


@dataclass
class HTensor:
    """Result of a decomposition: U[t] (rows_t x k_t), B[t] (k_l, k_r, k_t, column-major),
    sv[t] (k_t singular values), the relative error and the timing statistics."""
    dims: list
    tree: Tree
    ranks: list
    packed: object
    U: Dict[int, object] = field(default_factory=dict)
    B: Dict[int, object] = field(default_factory=dict)
    sv: Dict[int, object] = field(default_factory=dict)
    rel_err: float = float("nan")
    stats: Optional[dict] = None
    want_interior_u: bool = False


class _Views(Mapping):
    """Read-only node -> view mapping over the packed block; a view is made on first use."""

    def __init__(self, keys, make):
        self._keys = keys
        self._make = make
        self._cache = {}

    def __getitem__(self, t):
        v = self._cache.get(t)
        if v is None:
            if t not in self._keys:
                raise KeyError(t)
            v = self._cache[t] = self._make(t)
        return v

    def __iter__(self):
        return iter(self._keys)

    def __len__(self):
        return len(self._keys)


def _unpack(res: HTensor, layout):
    u_off, b_off, s_off, _ = layout
    tr, rk, pk = res.tree, res.ranks, res.packed

    def rows(t):
        r = 1
        for m in range(tr.lo[t], tr.hi[t]):
            r *= res.dims[m]
        return r

    def make_u(t):
        n = rows(t) * rk[t]
        return pk[u_off[t]:u_off[t] + n].reshape(rk[t], rows(t)).T

    def make_b(t):
        kl, kr = rk[tr.left[t]], rk[tr.right[t]]
        n = kl * kr * rk[t]
        if hasattr(pk, "permute"):
            return pk[b_off[t]:b_off[t] + n].reshape(rk[t], kr, kl).permute(2, 1, 0)
        return pk[b_off[t]:b_off[t] + n].reshape((kl, kr, rk[t]), order="F")

    def make_s(t):
        return pk[s_off[t]:s_off[t] + rk[t]]

    nodes = range(tr.n_nodes)
    res.U = _Views([t for t in nodes if u_off[t] >= 0], make_u)
    res.B = _Views([t for t in nodes if b_off[t] >= 0], make_b)
    res.sv = _Views([t for t in nodes if s_off[t] >= 0], make_s)


def decompose(x, dims, tree: Tree, ranks, method="ltr", precision="bf16", err="residual", handle=None,
              want_interior_u=False, **kw):
    """HT decomposition of the column-major FP32 tensor x (torch CUDA tensor or numpy).

    Mirrors the reference drivers htd/htd4 (third_order_tensor_decompsition/CUDA/Baseline/
    function.h:6, fourth_order_tensor_decomposition/single_GPU/CUDA/Baseline/function.h:6)
    with an explicit tree and the outputs returned instead of freed.
    """
    lib = _lib.load()
    h = handle or default_handle()
    dims = [int(d) for d in dims]
    ranks = [int(r) for r in ranks]
    okey = (method, precision, err, bool(want_interior_u), tuple(sorted(kw.items())))
    opt = _OPTIONS.get(okey) if "sketch" not in kw else None
    if opt is None:
        opt = make_options(method, precision, err, want_interior_u=want_interior_u, **kw)
        if "sketch" not in kw:
            if len(_OPTIONS) >= 64:
                _OPTIONS.clear()
            _OPTIONS[okey] = opt
    layout, targs = _layout_entry(dims, tree, ranks, want_interior_u)
    total = layout[3]
    on_dev = hasattr(x, "is_cuda") and x.is_cuda
    if on_dev:
        torch = _torch()
        if x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError("x must be a contiguous float32 tensor (column-major data)")
        out = torch.empty(total, dtype=torch.float32, device=x.device)
        xptr, optr = x.data_ptr(), out.data_ptr()
        h.order_after_torch(x.device)
    else:
        xa = np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))
        out = np.empty(total, dtype=np.float32)
        xptr, optr = xa.ctypes.data, out.ctypes.data
    if math.prod(dims) != (x.numel() if on_dev else xa.size):
        raise ValueError("x has the wrong number of elements")
    rel = ctypes.c_double()
    st = _lib.Stats()
    status = lib.ht_decompose(h.ptr, ctypes.c_void_p(xptr), int(on_dev), len(dims), *targs,
                              ctypes.byref(opt), ctypes.c_void_p(optr), int(on_dev), ctypes.byref(rel),
                              ctypes.byref(st))
    _lib.check(status, "ht_decompose")
    res = HTensor(dims=dims, tree=tree, ranks=ranks, packed=out, rel_err=rel.value, want_interior_u=want_interior_u,
                  stats=dict(total_ms=st.total_ms, h2d_ms=st.h2d_ms, passes=st.bytes_x_passes,
                             eig_iters_max=st.eig_iters_max, eig_calls=st.eig_calls,
                             sketch_fallbacks=st.sketch_fallbacks, sketch_nodes=st.sketch_nodes,
                             sketch_resid=st.sketch_resid, workspace_bytes=st.workspace_bytes,
                             replayed=st.replayed))
    _unpack(res, layout)
    return res


def decompose_adaptive(x, dims, tree: Tree, max_ranks, rel_eps, max_rank=0, method="plain", precision=None,
                       err="residual", handle=None, **kw):
    """ht_decompose_adaptive: each node's rank is the smallest (within max_ranks[t] and
    max_rank) that keeps ||X - X_HT|| <= rel_eps ||X|| (per-node energy budget
    (rel_eps ||X||)^2 / m, m = 2d-3 for plain, 2d-2 for ltr). The rel_eps/max_rank of
    htd (third_order_tensor_decompsition/CUDA/Baseline/function.h:6). Returns an HTensor
    whose `ranks` are the ranks chosen."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle()
    dims = [int(d) for d in dims]
    max_ranks = [int(r) for r in max_ranks]
    if precision is None:
        precision = "fp32" if method == "plain" else "bf16"
    opt = make_options(method, precision, err, **kw)
    layout, targs = _layout_entry(dims, tree, max_ranks, False)
    if not (hasattr(x, "is_cuda") and x.is_cuda) or x.dtype != torch.float32 or not x.is_contiguous():
        raise ValueError("x must be a contiguous float32 device tensor (column-major data)")
    out = torch.empty(layout[3], dtype=torch.float32, device=x.device)
    chosen = (ctypes.c_int32 * tree.n_nodes)()
    rel = ctypes.c_double()
    st = _lib.Stats()
    h.order_after_torch(x.device)
    _lib.check(lib.ht_decompose_adaptive(h.ptr, ctypes.c_void_p(x.data_ptr()), 1, len(dims), *targs, float(rel_eps),
                                         int(max_rank), ctypes.byref(opt), ctypes.c_void_p(out.data_ptr()), 1,
                                         chosen, ctypes.byref(rel), ctypes.byref(st)), "ht_decompose_adaptive")
    ranks = [int(v) for v in chosen]
    lay = output_layout(dims, tree, ranks)
    res = HTensor(dims=dims, tree=tree, ranks=ranks, packed=out[:lay[3]], rel_err=rel.value,
                  stats=dict(total_ms=st.total_ms, passes=st.bytes_x_passes, eig_calls=st.eig_calls))
    _unpack(res, lay)
    return res


def decompose_many(xs, dims, tree: Tree, ranks, method="ltr", precision="bf16", err="residual", handle=None,
                   want_interior_u=False, inflight=2, outs=None, **kw):
    """ht_decompose_many: decompose several device tensors of one shape with `inflight`
    decompositions running at once (their passes over X ordered on the device). `outs`
    (optional) is a (len(xs), layout total) float32 device tensor to write into. Returns
    (list of HTensor, stats)."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle()
    dims = [int(d) for d in dims]
    ranks = [int(r) for r in ranks]
    opt = make_options(method, precision, err, want_interior_u=want_interior_u, **kw)
    layout, targs = _layout_entry(dims, tree, ranks, want_interior_u)
    total = layout[3]
    n = math.prod(dims)
    for x in xs:
        if not (hasattr(x, "is_cuda") and x.is_cuda) or x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError("xs must be contiguous float32 device tensors (column-major data)")
        if x.numel() != n:
            raise ValueError("every tensor must have prod(dims) elements")
    count = len(xs)
    if outs is None:
        outs = torch.empty((count, total), dtype=torch.float32, device=xs[0].device)
    if tuple(outs.shape) != (count, total):
        raise ValueError("outs must be (len(xs), layout total)")
    xp = (ctypes.c_void_p * count)(*[x.data_ptr() for x in xs])
    op = (ctypes.c_void_p * count)(*[outs[i].data_ptr() for i in range(count)])
    errs = np.full(count, np.nan, dtype=np.float64)
    st = _lib.Stats()
    h.order_after_torch(xs[0].device)
    _lib.check(lib.ht_decompose_many(h.ptr, count, ctypes.cast(xp, ctypes.c_void_p), len(dims), *targs,
                                     ctypes.byref(opt), ctypes.cast(op, ctypes.c_void_p),
                                     errs.ctypes.data_as(_lib.PD), int(inflight), ctypes.byref(st)),
               "ht_decompose_many")
    results = []
    for i in range(count):
        res = HTensor(dims=dims, tree=tree, ranks=ranks, packed=outs[i], rel_err=float(errs[i]),
                      want_interior_u=want_interior_u)
        _unpack(res, layout)
        results.append(res)
    return results, dict(total_ms=st.total_ms, passes=st.bytes_x_passes, workspace_bytes=st.workspace_bytes,
                         replayed=st.replayed)


def residual(x, res: HTensor, handle=None):
    """||X - X_HT||_F / ||X||_F in one device pass (x and res.packed on the device)."""
    lib = _lib.load()
    h = handle or default_handle()
    tr = res.tree
    rel = ctypes.c_double()
    # the library's stream reads x and the packed block: it waits for what torch's
    # current stream has queued on them
    h.order_after_torch(x.device)
    _lib.check(lib.ht_residual(h.ptr, ctypes.c_void_p(x.data_ptr()), len(res.dims), _lib.i64_array(res.dims),
                               tr.n_nodes, _lib.i32_array(tr.lo), _lib.i32_array(tr.hi), _lib.i32_array(tr.left),
                               _lib.i32_array(tr.right), _lib.i32_array(res.ranks), int(res.want_interior_u),
                               ctypes.c_void_p(res.packed.data_ptr()), ctypes.byref(rel)), "ht_residual")
    return rel.value


def reconstruct(res: HTensor, handle=None):
    """Explicit X_HT (flat, column-major) on the device."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle()
    tr = res.tree
    out = torch.empty(int(np.prod(res.dims)), dtype=torch.float32, device=res.packed.device)
    # the library's stream writes `out` (a block torch's allocator may have just recycled
    # from work still queued on its current stream) and reads the packed block
    h.order_after_torch(res.packed.device)
    _lib.check(lib.ht_reconstruct(h.ptr, len(res.dims), _lib.i64_array(res.dims), tr.n_nodes,
                                  _lib.i32_array(tr.lo), _lib.i32_array(tr.hi), _lib.i32_array(tr.left),
                                  _lib.i32_array(tr.right), _lib.i32_array(res.ranks), int(res.want_interior_u),
                                  ctypes.c_void_p(res.packed.data_ptr()), ctypes.c_void_p(out.data_ptr())),
               "ht_reconstruct")
    return out


def generate_tucker(dims, core_dims, seed, noise_rel, device=0, handle=None):
    """Seeded Tucker tensor + Gaussian noise, generated on the device (flat, column-major)."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle(device)
    x = torch.empty(int(np.prod(dims)), dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.synchronize(x.device)
    _lib.check(lib.ht_generate_tucker(h.ptr, len(dims), _lib.i64_array(dims), _lib.i32_array(core_dims),
                                      int(seed), float(noise_rel), ctypes.c_void_p(x.data_ptr())),
               "ht_generate_tucker")
    return x


def generate_cp_decay(dims, rank, seed, noise_rel, exponent=1.5, device=0, handle=None):
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle(device)
    x = torch.empty(int(np.prod(dims)), dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.synchronize(x.device)
    _lib.check(lib.ht_generate_cp_decay(h.ptr, len(dims), _lib.i64_array(dims), int(rank), float(exponent),
                                        int(seed), float(noise_rel), ctypes.c_void_p(x.data_ptr())),
               "ht_generate_cp_decay")
    return x


def generate_bumps(n, order, seed, sample, n_bumps=6, device=0, handle=None):
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle(device)
    x = torch.empty(int(n) ** int(order), dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.synchronize(x.device)
    _lib.check(lib.ht_generate_bumps(h.ptr, int(order), int(n), int(n_bumps), int(seed), int(sample),
                                     ctypes.c_void_p(x.data_ptr())), "ht_generate_bumps")
    return x


class Batch:
    """A batch of small 3rd-order tensors packed in one device buffer (tree {1},{2,3})."""

    def __init__(self, dims, x, x_offsets):
        self.dims = np.asarray(dims, dtype=np.int64).reshape(-1, 3)
        self.x = x
        self.x_offsets = np.asarray(x_offsets, dtype=np.int64)

    @property
    def count(self):
        return self.dims.shape[0]

    def member(self, b):
        n = int(np.prod(self.dims[b]))
        o = int(self.x_offsets[b])
        return self.x[o:o + n]


def batch_shapes(count, seed, lo=16, hi=64, rlo=4, rhi=8):
    """Mode sizes and Tucker ranks of the batched workload (count x 3 each)."""
    lib = _lib.load()
    dims = np.zeros(3 * count, dtype=np.int64)
    core = np.zeros(3 * count, dtype=np.int32)
    _lib.check(lib.ht_batch_shapes(int(count), int(seed), lo, hi, rlo, rhi,
                                   dims.ctypes.data_as(_lib.PI64), core.ctypes.data_as(_lib.PI32)), "ht_batch_shapes")
    return dims.reshape(-1, 3), core.reshape(-1, 3)


def generate_batch(dims, core, seed, noise_rel, device=0, handle=None):
    """Seeded Tucker members (Philox streams 4096 + 16 b ...) packed back to back on the device."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle(device)
    dims = np.ascontiguousarray(np.asarray(dims, dtype=np.int64).reshape(-1, 3))
    core = np.ascontiguousarray(np.asarray(core, dtype=np.int32).reshape(-1, 3))
    sizes = np.prod(dims, axis=1)
    # 16-byte aligned members
    sizes_al = (sizes + 3) // 4 * 4
    offs = np.concatenate([[0], np.cumsum(sizes_al)[:-1]]).astype(np.int64)
    x = torch.zeros(int(sizes_al.sum()), dtype=torch.float32, device=f"cuda:{device}")
    torch.cuda.synchronize(x.device)
    _lib.check(lib.ht_generate_batch(h.ptr, dims.shape[0], dims.ctypes.data_as(_lib.PI64),
                                     core.ctypes.data_as(_lib.PI32), int(seed), float(noise_rel),
                                     ctypes.c_void_p(x.data_ptr()), offs.ctypes.data_as(_lib.PI64)),
               "ht_generate_batch")
    return Batch(dims, x, offs)


def batch_ranks(core):
    """Heap ranks [1, k1, k23 = k1, k2, k3] of the tree {1},{2,3} from Tucker ranks."""
    core = np.asarray(core, dtype=np.int32).reshape(-1, 3)
    r = np.empty((core.shape[0], 5), dtype=np.int32)
    r[:, 0] = 1
    r[:, 1] = core[:, 0]
    r[:, 2] = core[:, 0]
    r[:, 3] = core[:, 1]
    r[:, 4] = core[:, 2]
    return r


def decompose_batched(batch: Batch, ranks, handle=None):
    """ht_decompose_batched: returns (list of HTensor views, rel_err array, stats)."""
    torch = _torch()
    lib = _lib.load()
    h = handle or default_handle()
    ranks = np.ascontiguousarray(np.asarray(ranks, dtype=np.int32).reshape(-1, 5))
    count = batch.count
    dims = np.ascontiguousarray(batch.dims)
    offs = np.zeros(count, dtype=np.int64)
    total = ctypes.c_int64()
    _lib.check(lib.ht_batch_output_layout(count, dims.ctypes.data_as(_lib.PI64), ranks.ctypes.data_as(_lib.PI32),
                                          offs.ctypes.data_as(_lib.PI64), ctypes.byref(total)), "ht_batch_output_layout")
    out = torch.empty(int(total.value), dtype=torch.float32, device=batch.x.device)
    err = np.zeros(count, dtype=np.float64)
    st = _lib.Stats()
    torch.cuda.synchronize(batch.x.device)
    _lib.check(lib.ht_decompose_batched(h.ptr, ctypes.c_void_p(batch.x.data_ptr()), count,
                                        dims.ctypes.data_as(_lib.PI64), ranks.ctypes.data_as(_lib.PI32),
                                        batch.x_offsets.ctypes.data_as(_lib.PI64), ctypes.c_void_p(out.data_ptr()),
                                        offs.ctypes.data_as(_lib.PI64), err.ctypes.data_as(_lib.PD),
                                        ctypes.byref(st)), "ht_decompose_batched")
    from . import tree as _tree
    tr = _tree.first_vs_rest_3d()
    results = []
    for b in range(count):
        d = [int(v) for v in dims[b]]
        rk = [int(v) for v in ranks[b]]
        layout = output_layout(d, tr, rk, False)
        o = int(offs[b])
        res = HTensor(dims=d, tree=tr, ranks=rk, packed=out[o:o + layout[3]], rel_err=float(err[b]))
        _unpack(res, layout)
        results.append(res)
    return results, err, dict(total_ms=st.total_ms)


def _host_f32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(-1))


def htd(x, a, b, c, k, rel_eps=0.0, max_rank=0.0):
    """float htd(dt *x, long a, long b, long c, int *k, dt rel_eps, dt max_rank) --
    third_order_tensor_decompsition/CUDA/Baseline/function.h:6 (plain FP32, tree {1,2},{3},
    k heap-indexed with k[0]=1). Prints "cost time"/"error rate" like the reference and
    returns the relative error (negative on failure). With rel_eps > 0 the ranks chosen
    are written back into k (a list or an integer array), as the C entry writes them."""
    lib = _lib.load()
    xa = _host_f32(x)
    karr = _lib.i32_array(k)
    e = float(lib.htd(ctypes.c_void_p(xa.ctypes.data), a, b, c, karr, float(rel_eps), float(max_rank)))
    if rel_eps > 0:
        for i in range(len(k)):
            k[i] = int(karr[i])
    return e


def htd_tc(x, a, b, c, k):
    """The mixed-precision variant of htd (third_order_tensor_decompsition/CUDA/Optimized/function.h:6)."""
    lib = _lib.load()
    xa = _host_f32(x)
    return float(lib.htd_tc(ctypes.c_void_p(xa.ctypes.data), a, b, c, _lib.i32_array(k)))


def htd4(x, a, b, c, d, k):
    """void htd4(dt *x, long a, long b, long c, long d, int *k) --
    fourth_order_tensor_decomposition/single_GPU/CUDA/Baseline/function.h:6."""
    lib = _lib.load()
    xa = _host_f32(x)
    lib.htd4(ctypes.c_void_p(xa.ctypes.data), a, b, c, d, _lib.i32_array(k))


def htd4_tc(x, a, b, c, d, k):
    """fourth_order_tensor_decomposition/single_GPU/CUDA/Optimized/function.h:6 (handles ignored)."""
    lib = _lib.load()
    xa = _host_f32(x)
    lib.htd4_tc(ctypes.c_void_p(xa.ctypes.data), a, b, c, d, _lib.i32_array(k), None, None)
