"""ht_decompose_noalloc, the entry point that allocates no device memory, against
ht_decompose on the same inputs: two calls in a row (the 128^4 rank-16 leaves-to-root
workload, then a smaller 3rd-order plain decomposition), each compared with the
allocating entry point's output, then a problem too large for the fixed workspace, which
must be refused with HT_ERR_WORKSPACE.
"""
import ctypes

import numpy as np
import pytest

import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
from ht_tensor_learning_using_gpu_tensor_cores_amd import _lib
from oracle import gen as G

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 160 << 20  # HT_NOALLOC_SCRATCH_BYTES


def noalloc(h, x, dims, tr, ranks, method, precision):
    """One ht_decompose_noalloc call through the C ABI; returns (status, packed, rel_err, stats)."""
    import torch
    lib = _lib.load()
    opt = ht.engine.make_options(method, precision, "residual")
    layout = ht.output_layout(dims, tr, ranks)
    out = torch.full((layout[3],), float("nan"), dtype=torch.float32, device=x.device)
    rel = ctypes.c_double()
    st = _lib.Stats()
    torch.cuda.synchronize()
    status = lib.ht_decompose_noalloc(h.ptr, ctypes.c_void_p(x.data_ptr()), len(dims), _lib.i64_array(dims),
                                      tr.n_nodes, _lib.i32_array(tr.lo), _lib.i32_array(tr.hi),
                                      _lib.i32_array(tr.left), _lib.i32_array(tr.right), _lib.i32_array(ranks),
                                      ctypes.byref(opt), ctypes.c_void_p(out.data_ptr()), ctypes.byref(rel),
                                      ctypes.byref(st))
    torch.cuda.synchronize()
    return status, out, rel.value, st


def compare(h_ref, h_na, x, dims, tr, ranks, method, precision):
    import torch
    ref = ht.decompose(x, dims, tr, ranks, method=method, precision=precision, handle=h_ref)
    need = ref.stats["workspace_bytes"]
    print(f"{method}/{precision} dims={dims}: workspace {need / 2**20:.1f} MiB of {SCRATCH_BYTES >> 20} MiB",
          flush=True)
    assert 0 < need <= SCRATCH_BYTES
    free0 = torch.cuda.mem_get_info()[0]
    status, out, rel, st = noalloc(h_na, x, dims, tr, ranks, method, precision)
    free1 = torch.cuda.mem_get_info()[0]
    assert status == 0, _lib.STATUS_NAMES.get(status, status)
    assert st.workspace_bytes == need
    a, b = ref.packed.cpu().numpy(), out.cpu().numpy()
    assert np.all(np.isfinite(b))
    # the same kernels on the same inputs: equal outputs (deterministic reductions)
    np.testing.assert_allclose(b, a, rtol=0, atol=1e-6 * max(1.0, np.abs(a).max()))
    assert abs(rel - ref.rel_err) <= 1e-7, (rel, ref.rel_err)
    print(f"  bit-exact: {np.array_equal(a, b)}; rel_err {rel:.9f}; device free memory change "
          f"{free0 - free1} B", flush=True)
    return free0 - free1


def test_noalloc_matches_decompose_twice_then_refuses_too_large():
    import torch
    with ht.Handle() as h_ref, ht.Handle() as h_na:
        # call 1: the 128^4 rank-16 leaves-to-root workload
        dims4, tr4, r4 = [128] * 4, ht.tree.balanced_4d(), [1] + [16] * 6
        x4 = ht.generate_tucker(dims4, [16] * 4, seed=1, noise_rel=1e-4)
        compare(h_ref, h_na, x4, dims4, tr4, r4, "ltr", "bf16")
        # the same call again replays the recorded solver rounds; no device memory taken
        d = compare(h_ref, h_na, x4, dims4, tr4, r4, "ltr", "bf16")
        assert d == 0
        del x4
        # call 2: a smaller input, another method (the workspace keeps the last call's
        # contents; nothing may depend on them)
        dims3, tr3, r3 = [96, 80, 64], ht.tree.first_vs_rest_3d(), [1, 8, 8, 8, 8]
        x3 = torch.from_numpy(G.tucker(dims3, [6, 7, 8], seed=2, noise_rel=1e-3)).cuda()
        compare(h_ref, h_na, x3, dims3, tr3, r3, "plain", "fp32")
        compare(h_ref, h_na, x3, dims3, tr3, r3, "rsvd", "bf16")
        # too large for the fixed workspace: refused, nothing written past it
        dimsL, trL, rL = [1024, 1024, 256], ht.tree.first_vs_rest_3d(), [1, 32, 32, 32, 32]
        xL = ht.generate_tucker(dimsL, [32, 32, 32], seed=3, noise_rel=1e-3)
        ref = ht.decompose(xL, dimsL, trL, rL, method="ltr", precision="bf16", handle=h_ref)
        print(f"large case workspace {ref.stats['workspace_bytes'] / 2**20:.1f} MiB", flush=True)
        assert ref.stats["workspace_bytes"] > SCRATCH_BYTES
        status, _, _, _ = noalloc(h_na, xL, dimsL, trL, rL, "ltr", "bf16")
        assert _lib.STATUS_NAMES.get(status) == "HT_ERR_WORKSPACE"
        # and the entry point still works after the refusal
        compare(h_ref, h_na, x3, dims3, tr3, r3, "ltr", "bf16")
