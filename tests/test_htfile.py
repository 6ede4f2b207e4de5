"""The on-disk HT format (ht-npz/1): save/load round trips, the reconstruction from the
file, the HT-layer tuple (HT_tensor_layer/FC_MNIST/mnist_ht.py:194-196) and malformed
files. GPU: a decomposition saved from the device, loaded back and checked by the fused
residual pass."""
import types

import numpy as np
import pytest

import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
from ht_tensor_learning_using_gpu_tensor_cores_amd import htfile
from oracle import gen as G
from oracle import ht_oracle as O


def as_result(h, tree, rel_err):
    """An oracle HT in the shape of engine.decompose's result (host arrays)."""
    U = {t: h.U[t].astype(np.float32) for t in tree.leaves()}
    B = {t: h.B[t].astype(np.float32) for t in tree.interior()}
    sv = {t: np.asarray(h.sv[t], dtype=np.float32) for t in range(1, tree.n_nodes) if t in h.sv}
    return types.SimpleNamespace(dims=h.dims, tree=tree, ranks=h.ranks, U=U, B=B, sv=sv, rel_err=rel_err)


@pytest.mark.parametrize("tree,dims,core,ranks", [
    (ht.tree.reference_3d(), [12, 10, 8], [3, 3, 3], [1, 6, 3, 3, 3]),
    (ht.tree.balanced_4d(), [8, 7, 6, 5], [2, 3, 2, 2], [1, 5, 4, 2, 3, 2, 2]),
    (ht.tree.balanced(5), [6, 5, 4, 5, 6], [2, 2, 2, 2, 2], [1, 4, 4, 4, 2, 2, 2, 2, 2]),
])
def test_roundtrip(tmp_path, tree, dims, core, ranks):
    x = G.tucker(dims, core, seed=3, noise_rel=1e-2)
    otr = O.Tree(tree.lo, tree.hi, tree.left, tree.right)
    h = O.plain_ht(x, dims, otr, ranks)
    e = O.rel_error(x, h.reconstruct())
    res = as_result(h, tree, e)
    path = tmp_path / "a.npz"
    htfile.save(path, res)
    f = htfile.load(path)
    assert f.dims == dims and f.ranks == ranks and f.tree.lo == tree.lo and f.tree.left == tree.left
    assert f.rel_err == e
    for t in res.U:
        assert np.array_equal(f.U[t], res.U[t])
    for t in res.B:
        assert np.array_equal(f.B[t], res.B[t])
    for t in res.sv:
        assert np.array_equal(f.sv[t], res.sv[t])
    # the file reproduces the decomposition's reconstruction and error
    xr = f.reconstruct()
    np.testing.assert_allclose(xr, h.reconstruct(), atol=1e-5 * np.abs(x).max())
    assert abs(O.rel_error(x, xr) - e) < 1e-5
    # the packed block has the library's layout
    p = f.packed()
    u_off, b_off, s_off, total = ht.output_layout(dims, tree, ranks)
    assert p.size == total
    t0 = tree.leaves()[0]
    assert np.array_equal(p[u_off[t0]:u_off[t0] + res.U[t0].size], res.U[t0].reshape(-1, order="F"))


def test_layer_factors_match_the_reference_contraction(tmp_path):
    """The tuple (u1, u2, u3, b2, b1) of the reference's HT layer: contracting it as the
    layer does (mnist_ht.py:232-241, without the input/output split) gives X_HT."""
    dims, ranks = [8, 6, 10], [1, 5, 4, 3, 3]
    tree = ht.tree.reference_3d()
    x = G.tucker(dims, [3, 3, 4], seed=5, noise_rel=1e-3)
    otr = O.Tree(tree.lo, tree.hi, tree.left, tree.right)
    h = O.plain_ht(x, dims, otr, ranks)
    path = tmp_path / "b.npz"
    htfile.save(path, as_result(h, tree, O.rel_error(x, h.reconstruct())))
    u1, u2, u3, b2, b1 = htfile.load(path).layer_factors()
    assert u1.shape == (8, 3) and u2.shape == (6, 3) and u3.shape == (10, 4) and b2.shape == (3, 3, 5)
    assert b1.shape == (5, 4)
    x_ht = np.einsum("ia,jb,abc,cd,kd->ijk", u1, u2, b2, b1, u3)
    np.testing.assert_allclose(x_ht.reshape(-1, order="F"), h.reconstruct(), atol=1e-5 * np.abs(x).max())
    with pytest.raises(ValueError):
        htfile.load(path).__class__(dims, ht.tree.first_vs_rest_3d(), ranks, {}, {}, {}, 0.0).layer_factors()


def test_malformed_files_are_refused(tmp_path):
    dims, ranks = [6, 5, 4], [1, 3, 3, 2, 2]
    tree = ht.tree.reference_3d()
    x = G.tucker(dims, [2, 2, 2], seed=1, noise_rel=1e-2)
    h = O.plain_ht(x, dims, O.Tree(tree.lo, tree.hi, tree.left, tree.right), ranks)
    good = tmp_path / "g.npz"
    htfile.save(good, as_result(h, tree, 0.1))
    z = dict(np.load(good, allow_pickle=False))
    bad = dict(z)
    bad["U3"] = z["U3"][:, :1]
    np.savez(tmp_path / "shape.npz", **bad)
    with pytest.raises(ValueError):
        htfile.load(tmp_path / "shape.npz")
    bad = dict(z)
    del bad["B1"]
    np.savez(tmp_path / "missing.npz", **bad)
    with pytest.raises(ValueError):
        htfile.load(tmp_path / "missing.npz")
    bad = dict(z)
    bad["format"] = np.array("something-else")
    np.savez(tmp_path / "fmt.npz", **bad)
    with pytest.raises(ValueError):
        htfile.load(tmp_path / "fmt.npz")


@pytest.mark.gpu
def test_gpu_decomposition_saved_and_reloaded(tmp_path):
    import torch
    dims, tree, ranks = [64, 48, 40, 32], ht.tree.balanced_4d(), [1, 24, 20, 8, 6, 8, 6]
    x = G.tucker(dims, [8, 6, 8, 6], seed=2, noise_rel=1e-3)
    xd = torch.from_numpy(x).cuda()
    res = ht.decompose(xd, dims, tree, ranks, method="ltr")
    htfile.save(tmp_path / "c.npz", res)
    f = htfile.load(tmp_path / "c.npz")
    assert abs(O.rel_error(x, f.reconstruct()) - res.rel_err) <= 1e-5
    back = f.to_device()
    assert np.array_equal(back.packed.cpu().numpy(), res.packed.cpu().numpy())
    assert abs(ht.residual(xd, back) - res.rel_err) <= 1e-7
