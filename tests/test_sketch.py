"""Learned per-mode sketch (the Khatri-Rao test matrix of the randomised method).

CPU: the oracle's two loss/gradient evaluations (dense unfoldings, and the bump family's
factored form) agree with each other, with central finite differences, with a
brute-force Khatri-Rao product, with the QR projector, and with the exact-SVD bound.
GPU: the library's forward/backward kernels match the oracle; training on generated bump
tensors beats the Gaussian sketch on held-out tensors; inside the decomposition the
learned sketch lowers the error of the q = 0 range finder, and the tolerance-driven
Gaussian fallback reproduces the Gaussian run exactly.
"""
import ctypes
import itertools

import numpy as np
import pytest

from oracle import ht_oracle as O
from oracle import sketch_oracle as SO

T4 = O.tree_balanced_4d()
T3 = O.tree_first_vs_rest_3d()


# ----------------------------------------------------------------------------- oracle pins

def test_oracle_sketched_nodes_and_column_modes():
    assert SO.sketched_nodes(T4) == [1] and SO.col_modes(T4, 1) == [2, 3]
    assert SO.sketched_nodes(T3) == [2] and SO.col_modes(T3, 2) == [0]


def test_oracle_kr_omega_brute_force(rng):
    S = [rng.standard_normal((3, k)) for k in (2, 3, 4, 5)]
    modes = [1, 2, 3]
    om = SO.kr_omega(S, modes)
    bf = np.zeros((60, 3))
    for i1, i2, i3 in itertools.product(range(3), range(4), range(5)):
        bf[i1 + 3 * i2 + 12 * i3] = S[1][:, i1] * S[2][:, i2] * S[3][:, i3]
    np.testing.assert_array_equal(om, bf)
    np.testing.assert_array_equal(SO.kr_omega(S, [0]), S[0].T)


def test_oracle_kr_backward_is_the_adjoint(rng):
    S = [rng.standard_normal((3, k)) for k in (4, 5, 6)]
    modes = [0, 1, 2]
    dom = rng.standard_normal((120, 3))
    g = SO.kr_backward(S, modes, dom)
    h = 1e-6
    for m, j, i in [(0, 1, 2), (1, 2, 4), (2, 0, 5)]:
        Sp = [s.copy() for s in S]
        Sm = [s.copy() for s in S]
        Sp[m][j, i] += h
        Sm[m][j, i] -= h
        fd = (np.sum(dom * SO.kr_omega(Sp, modes)) - np.sum(dom * SO.kr_omega(Sm, modes))) / (2 * h)
        assert abs(fd - g[m][j, i]) < 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("sample", [0, 1])
def test_oracle_factored_equals_dense(rng, sample):
    n, w = 9, 3
    a, g = SO.bump_factors(n, 4, sample)
    S = [rng.standard_normal((w, n)) for _ in range(4)]
    x = SO.dense_bump(a, g)
    l1, g1 = SO.bump_loss_grad(a, g, S)
    l2, g2 = SO.dense_loss_grad(x, [n] * 4, T4, S)
    assert abs(l1 - l2) < 1e-12
    for m in range(4):
        np.testing.assert_allclose(g1[m], g2[m], rtol=1e-6, atol=1e-12)
    assert not np.any(g2[0]) and not np.any(g2[1])  # modes 1, 2 are no sketched node's columns


@pytest.mark.parametrize("tree,dims", [(T4, [6, 7, 8, 9]), (T3, [11, 9, 10])])
def test_oracle_gradient_matches_finite_differences(rng, tree, dims):
    w = 3
    x = rng.standard_normal(int(np.prod(dims)))
    S = [rng.standard_normal((w, d)) for d in dims]
    _, g = SO.dense_loss_grad(x, dims, tree, S)
    h = 1e-5
    modes = SO.col_modes(tree, SO.sketched_nodes(tree)[0])
    for m in modes:
        for j, i in [(0, 0), (w - 1, dims[m] - 1), (1, dims[m] // 2)]:
            Sp = [s.copy() for s in S]
            Sm = [s.copy() for s in S]
            Sp[m][j, i] += h
            Sm[m][j, i] -= h
            fd = (SO.dense_loss_grad(x, dims, tree, Sp)[0] - SO.dense_loss_grad(x, dims, tree, Sm)[0]) / (2 * h)
            assert abs(fd - g[m][j, i]) <= 1e-6 * max(abs(fd), 1e-3), (m, j, i, fd, g[m][j, i])


def test_oracle_loss_is_the_qr_projector_miss(rng):
    # ridge -> 0: the loss is the energy fraction outside range(Y), computed through QR
    dims = [8, 9, 10]
    x = rng.standard_normal(720)
    S = [rng.standard_normal((4, d)) for d in dims]
    loss, _ = SO.dense_loss_grad(x, dims, T3, S, ridge=0.0)
    A = O.unfold(x, dims, 1, 3)
    Q, _ = np.linalg.qr(A @ S[0].T)
    B = Q.T @ A
    assert abs(loss - (1.0 - np.sum(B * B) / np.sum(A * A))) < 1e-12


def test_oracle_loss_bounds(rng):
    n = 10
    for sample in range(4):
        a, g = SO.bump_factors(n, 5, sample)
        for w in (1, 2, 4):
            S = [rng.standard_normal((w, n)) for _ in range(4)]
            loss, _ = SO.bump_loss_grad(a, g, S)
            assert SO.bump_exact_miss(a, g, w) - 1e-9 <= loss <= 1.0
        # width >= the family's rank (6 bumps): every generic sketch captures the range
        # (a small ridge, 1e-9, above the FP64 rounding of W); the library's relative ridge (1e-6 of tr(Y^T Y)/w) can only
        # lower the captured energy, since W + eps I >= W
        S = [rng.standard_normal((8, n)) for _ in range(4)]
        tiny = SO.bump_loss_grad(a, g, S, ridge=1e-9)[0]
        assert abs(tiny) < 1e-6
        assert tiny - 1e-7 <= SO.bump_loss_grad(a, g, S)[0] < 1e-3
        assert SO.bump_exact_miss(a, g, 6) < 1e-12


def test_library_sketch_layout(lib):
    off = (ctypes.c_int64 * 4)()
    tot = ctypes.c_int64()
    dims = (ctypes.c_int64 * 4)(128, 64, 32, 16)
    assert lib.ht_sketch_layout(4, dims, 20, off, ctypes.byref(tot)) == 0
    assert list(off) == [0, 2560, 3840, 4480] and tot.value == 4800
    assert lib.ht_sketch_layout(4, dims, 65, off, ctypes.byref(tot)) != 0
    assert lib.ht_sketch_layout(4, dims, 0, off, ctypes.byref(tot)) != 0


# ----------------------------------------------------------------------------- GPU

def _sk_list(sk):
    w = sk.weights.detach().cpu().numpy().astype(np.float64)
    out, o = [], 0
    for d in sk.dims:
        out.append(w[o:o + sk.width * d].reshape(sk.width, d))
        o += sk.width * d
    return out


def _grad_list(g, dims, width):
    g = g.detach().cpu().numpy().astype(np.float64)
    out, o = [], 0
    for d in dims:
        out.append(g[o:o + width * d].reshape(width, d))
        o += width * d
    return out


def _assert_grad_close(g, ref):
    # The gradient is the difference of two terms about sigma_1^2 / sigma_w^2 (1e3..1e4 here)
    # larger than itself, both from FP32 products of the FP32 tensor: its relative accuracy
    # is ~1e-6 times that ratio. Tolerance: 1e-2 relative in norm, direction to 1e-4.
    err = np.linalg.norm(g - ref) / np.linalg.norm(ref)
    cos = float(np.dot(g, ref) / (np.linalg.norm(g) * np.linalg.norm(ref)))
    assert err <= 1e-2 and cos >= 1 - 1e-4, (err, cos)


@pytest.mark.gpu
@pytest.mark.parametrize("w", [4, 20])
def test_gpu_loss_grad_matches_oracle_4d(w):
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n = 24
    print(f"sketch loss/grad 4d n={n} w={w}", flush=True)
    x = ht.generate_bumps(n, 4, 4, 0)
    sk = ht.sketch.gaussian([n] * 4, w, seed=11)
    loss, g = ht.sketch.loss_grad(x, [n] * 4, ht.tree.balanced_4d(), sk)
    S = _sk_list(sk)
    a, gf = SO.bump_factors(n, 4, 0)
    ol, og = SO.bump_loss_grad(a, gf, S)
    # the tensor reaches the device rounded to FP32 and every large product is FP32
    assert abs(loss - ol) < 2e-6, (loss, ol)
    gg = _grad_list(g, [n] * 4, w)
    assert not np.any(gg[0]) and not np.any(gg[1])
    if w == 4:
        _assert_grad_close(np.concatenate([gg[2].ravel(), gg[3].ravel()]),
                           np.concatenate([og[2].ravel(), og[3].ravel()]))
    else:
        # width 20 > rank 6: every sketch captures the range; what is left is the ridge's
        # small upward bias (and the gradient is mostly rounding noise)
        assert ol < 1e-5 and loss < 1e-5


@pytest.mark.gpu
def test_gpu_loss_grad_matches_oracle_3d():
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = 40, 5
    print(f"sketch loss/grad 3d n={n} w={w}", flush=True)
    x = ht.generate_bumps(n, 3, 9, 2)
    sk = ht.sketch.gaussian([n] * 3, w, seed=12)
    loss, g = ht.sketch.loss_grad(x, [n] * 3, ht.tree.first_vs_rest_3d(), sk)
    xd = x.detach().cpu().numpy().astype(np.float64)  # the same FP32 tensor
    ol, og = SO.dense_loss_grad(xd, [n] * 3, T3, _sk_list(sk))
    assert abs(loss - ol) < 1e-6, (loss, ol)
    gg = _grad_list(g, [n] * 3, w)
    assert not np.any(gg[1]) and not np.any(gg[2])
    _assert_grad_close(gg[0].ravel(), og[0].ravel())


@pytest.fixture(scope="module")
def trained():
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = 32, 4
    print(f"training sketch n={n} w={w}", flush=True)
    init = ht.sketch.gaussian([n] * 4, w, seed=7)
    sk, curve = ht.sketch.train_bumps(n, 4, ht.tree.balanced_4d(), w, count=96, seed=4, epochs=3, batch=8, lr=0.05,
                                      init=init)
    return dict(n=n, w=w, init=init, sk=sk, curve=curve)


@pytest.mark.gpu
def test_gpu_training_beats_gaussian_on_held_out(trained):
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = trained["n"], trained["w"]
    t = ht.tree.balanced_4d()
    lg = ht.sketch.eval_bumps(trained["init"], n, 4, t, 32, seed=5)
    ll = ht.sketch.eval_bumps(trained["sk"], n, 4, t, 32, seed=5)
    lb = ht.sketch.eval_bumps(trained["sk"].bf16(), n, 4, t, 32, seed=5)
    exact = np.array([SO.bump_exact_miss(*SO.bump_factors(n, 5, s), w) for s in range(32)])
    print(f"held-out loss: gaussian {lg.mean():.4e} learned {ll.mean():.4e} (bf16 {lb.mean():.4e}) "
          f"exact {exact.mean():.4e}; training curve {trained['curve']}", flush=True)
    assert np.all(ll >= exact - 1e-6) and np.all(lg >= exact - 1e-6)
    assert ll.mean() < 0.6 * lg.mean()
    assert abs(lb.mean() - ll.mean()) < 0.05 * ll.mean()
    # the device evaluation of the untrained sketch agrees with the oracle's
    S0 = _sk_list(trained["init"])
    og = np.array([SO.bump_loss_grad(*SO.bump_factors(n, 5, s), S0)[0] for s in range(32)])
    np.testing.assert_allclose(lg, og, rtol=0, atol=2e-6)


@pytest.mark.gpu
def test_gpu_learned_sketch_in_decomposition(trained):
    import torch
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = trained["n"], trained["w"]
    t = ht.tree.balanced_4d()
    ranks = [1] + [w] * 6
    el, eg, ex = [], [], []
    for s in range(6):
        x = ht.generate_bumps(n, 4, 5, s)
        rl = ht.decompose(x, [n] * 4, t, ranks, method="rsvd", precision="fp32", power_iters=0,
                          sketch=trained["sk"], sketch_tol=1.0)
        rg = ht.decompose(x, [n] * 4, t, ranks, method="rsvd", precision="fp32", power_iters=0, oversample=0)
        rp = ht.decompose(x, [n] * 4, t, ranks, method="plain", precision="fp32")
        assert rl.stats["sketch_nodes"] == 1 and rl.stats["sketch_fallbacks"] == 0
        assert rg.stats["sketch_nodes"] == 0
        el.append(rl.rel_err)
        eg.append(rg.rel_err)
        ex.append(rp.rel_err)
    el, eg, ex = map(np.array, (el, eg, ex))
    print(f"HT rel err (ranks {w}, q=0): learned {el.mean():.4e} gaussian {eg.mean():.4e} exact {ex.mean():.4e}",
          flush=True)
    assert np.all(ex <= el + 1e-6) and np.all(ex <= eg + 1e-6)
    assert el.mean() < eg.mean()
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_gpu_fallback_reproduces_gaussian(trained):
    import torch
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = trained["n"], trained["w"]
    t = ht.tree.balanced_4d()
    ranks = [1] + [w] * 6
    x = ht.generate_bumps(n, 4, 5, 7)
    rf = ht.decompose(x, [n] * 4, t, ranks, method="rsvd", precision="bf16", power_iters=1, oversample=0,
                      sketch=trained["sk"], sketch_tol=0.0)
    rg = ht.decompose(x, [n] * 4, t, ranks, method="rsvd", precision="bf16", power_iters=1, oversample=0)
    assert rf.stats["sketch_nodes"] == 1 and rf.stats["sketch_fallbacks"] == 1
    assert rf.stats["sketch_resid"] > 0
    assert torch.equal(rf.packed, rg.packed)
    assert rf.rel_err == rg.rel_err


@pytest.mark.gpu
def test_gpu_learned_sketch_rejects_bad_options(trained):
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n, w = trained["n"], trained["w"]
    x = ht.generate_bumps(n, 4, 5, 0)
    t = ht.tree.balanced_4d()
    with pytest.raises(ht.HTError) as e:  # width below the node ranks
        ht.decompose(x, [n] * 4, t, [1] + [w + 1] * 6, method="rsvd", sketch=trained["sk"])
    assert e.value.status == 3
    with pytest.raises(ht.HTError) as e:  # learned sketch outside the randomised method
        ht.decompose(x, [n] * 4, t, [1] + [w] * 6, method="ltr", sketch=trained["sk"])
    assert e.value.status == 1


@pytest.mark.gpu
def test_gpu_sketch_save_load_roundtrip(trained, tmp_path):
    import torch
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    p = tmp_path / "sk.npz"
    trained["sk"].save(p)
    sk2 = ht.sketch.Sketch.load(p)
    assert sk2.dims == trained["sk"].dims and sk2.width == trained["sk"].width
    assert torch.equal(sk2.weights, trained["sk"].bf16().weights)


@pytest.mark.gpu
@pytest.mark.parametrize("method,precision,tol", [("rsvd", "bf16", 2e-3), ("ltr", "bf16", 2e-3), ("plain", "fp32", 1e-4)])
def test_gpu_ranks_above_numerical_rank(method, precision, tol):
    # Bump tensors have CP rank 6 with mode spectra spanning five decades; ranks 12 ask the
    # eigensolvers for vectors of a cluster at the Gram's rounding level. Every method
    # must finish (no HT_ERR_NOCONV) with an error at that rounding level.
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    n = 64
    t = ht.tree.balanced_4d()
    for s in range(3):
        print(f"ranks 12 on bump sample {s}, n={n}, {method}/{precision}", flush=True)
        x = ht.generate_bumps(n, 4, 5, s)
        r = ht.decompose(x, [n] * 4, t, [1] + [12] * 6, method=method, precision=precision, power_iters=0)
        assert r.rel_err < tol, (s, r.rel_err)


@pytest.mark.gpu
def test_gpu_bundled_sketch_beats_gaussian_at_128():
    # the width-4 sketch shipped with the package (trained on 2000 tensors of seed 4) on
    # held-out tensors of seed 5 it has not seen in training
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    sk = ht.sketch.bundled("bumps128_w4")
    assert sk.dims == [128] * 4 and sk.width == 4
    t = ht.tree.balanced_4d()
    print("bundled width-4 sketch on 16 held-out 128^4 tensors", flush=True)
    ll = ht.sketch.eval_bumps(sk, 128, 4, t, 16, seed=5, first=300)
    lg = ht.sketch.eval_bumps(ht.sketch.gaussian([128] * 4, 4, seed=7), 128, 4, t, 16, seed=5, first=300)
    exact = np.array([SO.bump_exact_miss(*SO.bump_factors(128, 5, 300 + s), 4) for s in range(16)])
    print(f"  learned {ll.mean():.4e} gaussian {lg.mean():.4e} exact {exact.mean():.4e}", flush=True)
    assert np.all(ll >= exact - 1e-6)
    assert ll.mean() < 0.7 * lg.mean()
