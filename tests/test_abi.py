"""The C ABI on the CPU: the library loads, exports every declared symbol, and its host
logic (tree validation, output layout, Philox) behaves without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

from ht_tensor_learning_using_gpu_tensor_cores_amd import _lib, tree as T
from oracle import philox

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("ht_amd.h", "ht_amd_debug.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:ht_status|void|float|const char\s*\*|void\s*\*)\s*\**\s*([a-z_0-9]+)\s*\(",
                             src, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert {"ht_decompose", "htd", "htd4", "ht_tsqr"} <= names
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_lib.SIGNATURES) == names


def test_version(lib):
    assert b"gfx950" in lib.ht_version()


def _validate(lib, tr):
    return lib.ht_validate_tree(tr.order, tr.n_nodes, _lib.i32_array(tr.lo), _lib.i32_array(tr.hi),
                                _lib.i32_array(tr.left), _lib.i32_array(tr.right))


@pytest.mark.parametrize("tr", [T.reference_3d(), T.first_vs_rest_3d(), T.balanced_4d(), T.balanced(5),
                                T.balanced(8), T.from_spec(((0, (1, 2)), 3))])
def test_valid_trees_accepted(lib, tr):
    assert _validate(lib, tr) == 0


def test_reference_heap_numbering():
    t3 = T.reference_3d()
    assert (t3.lo, t3.hi, t3.left, t3.right) == ([0, 0, 2, 0, 1], [3, 2, 3, 1, 2], [1, 3, -1, -1, -1], [2, 4, -1, -1, -1])
    t4 = T.balanced_4d()
    assert t4.left[:3] == [1, 3, 5] and t4.lo[3:] == [0, 1, 2, 3]


@pytest.mark.parametrize("bad", [
    dict(lo=[0, 0, 1], hi=[2, 1, 2], left=[1, -1, -1], right=[-1, -1, -1]),       # one child
    dict(lo=[0, 0, 0], hi=[2, 1, 1], left=[1, -1, -1], right=[2, -1, -1]),        # mode 1 missing
    dict(lo=[0, 1, 0], hi=[2, 2, 1], left=[1, -1, -1], right=[2, -1, -1]),        # children out of order
    dict(lo=[0, 0, 1, 2, 0], hi=[3, 2, 3, 3, 1], left=[1, 4, -1, -1, -1], right=[2, 3, -1, -1, -1]),
])
def test_invalid_trees_rejected(lib, bad):
    tr = T.Tree(**bad)
    assert lib.ht_validate_tree(tr.order, tr.n_nodes, _lib.i32_array(tr.lo), _lib.i32_array(tr.hi),
                                _lib.i32_array(tr.left), _lib.i32_array(tr.right)) == 2
    with pytest.raises(ValueError):
        tr.validate()


def _layout(lib, dims, tr, ranks, want=0):
    n = tr.n_nodes
    u, b, s = (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)(), (ctypes.c_int64 * n)()
    tot = ctypes.c_int64()
    st = lib.ht_output_layout(len(dims), _lib.i64_array(dims), n, _lib.i32_array(tr.lo), _lib.i32_array(tr.hi),
                              _lib.i32_array(tr.left), _lib.i32_array(tr.right), _lib.i32_array(ranks), want, u, b,
                              s, ctypes.byref(tot))
    return st, list(u), list(b), list(s), tot.value


def test_output_layout_4d(lib):
    st, u, b, s, tot = _layout(lib, [128] * 4, T.balanced_4d(), [1] + [16] * 6)
    assert st == 0
    assert u[:3] == [-1, -1, -1] and u[3:] == [0, 2048, 4096, 6144]
    assert b[0] == 8192 and b[1] == 8192 + 256 and b[2] == 8192 + 256 + 4096
    assert tot == 8192 + 256 + 2 * 4096 + 6 * 16
    st, u, *_ = _layout(lib, [128] * 4, T.balanced_4d(), [1] + [16] * 6, want=1)
    assert u[1] == 0 and u[2] == 16384 * 16


@pytest.mark.parametrize("ranks,code", [([2, 4, 4, 4, 4, 4, 4], 3),            # root rank != 1
                                        ([1, 4, 4, 0, 4, 4, 4], 3),            # rank 0
                                        ([1, 17, 4, 4, 4, 4, 4], 3),           # k > k_l * k_r
                                        ([1, 4, 4, 9, 4, 4, 4], 3),            # k > n
                                        ([1, 65, 65, 65, 65, 65, 65], 3)])     # k > HT_MAX_RANK
def test_rank_validation(lib, ranks, code):
    st, *_ = _layout(lib, [8, 8, 100, 100] if ranks[3] != 65 else [100] * 4, T.balanced_4d(), ranks)
    assert st == code


def test_philox_host_matches_oracle(lib):
    n = 1003
    out = (ctypes.c_double * n)()
    lib.ht_philox_normals_host(99, 256 + 3, n, out)
    np.testing.assert_allclose(np.array(out), philox.normals(99, 256 + 3, n), rtol=1e-14, atol=1e-14)


def test_no_device_is_an_error_not_a_crash(lib):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert lib.ht_create(ctypes.byref(h), 0) == 9
