"""Component parity of the gfx950 kernels against numpy FP64 (through the C ABI)."""
import ctypes

import numpy as np
import pytest

from ht_tensor_learning_using_gpu_tensor_cores_amd import _lib
from oracle import ht_oracle as O
from oracle import philox

pytestmark = pytest.mark.gpu


def dev(a, dtype=None):
    import torch
    a = np.asarray(a)
    t = torch.from_numpy(np.ascontiguousarray(a if dtype is None else a.astype(dtype)))
    return t.cuda()


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def host(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bf16_round(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def test_mfma_lane_layouts(lib, handle, rng):
    """The operand/accumulator lane layouts the kernels assume, checked per shape."""
    import torch
    a = rng.standard_normal((32, 16)).astype(np.float32)
    b = rng.standard_normal((32, 16)).astype(np.float32)
    da, db = dev(a), dev(b)
    o1 = torch.zeros(32 * 32, device="cuda")
    o2 = torch.zeros(32 * 32, device="cuda")
    o3 = torch.zeros(16 * 16, device="cuda")
    o4 = torch.zeros(16 * 16, dtype=torch.float64, device="cuda")
    o5 = torch.zeros(16 * 16, device="cuda")
    _lib.check(lib.ht_debug_mfma_probe(handle.ptr, ptr(da), ptr(db), ptr(o1), ptr(o2), ptr(o3), ptr(o4), ptr(o5)),
               "probe")
    # 32x32x16 bf16: D = A B with A = a (32x16), B = b^T (16x32)
    ab, bb = bf16_round(a).astype(np.float64), bf16_round(b).astype(np.float64)
    np.testing.assert_allclose(host(o1).reshape(32, 32), ab @ bb.T, rtol=1e-5, atol=1e-4)
    # 32x32x2 f32: A = a[:, :2], B = b[:, :2]^T
    np.testing.assert_allclose(host(o2).reshape(32, 32), a[:, :2].astype(np.float64) @ b[:, :2].T, rtol=1e-5,
                               atol=1e-5)
    # 16x16x4 f32 / f64: A = a[:16, :4], B = b[:16, :4]^T
    ref = a[:16, :4].astype(np.float64) @ b[:16, :4].T
    np.testing.assert_allclose(host(o3).reshape(16, 16), ref, rtol=1e-5, atol=1e-5)
    # f64 16x16x4: the same product; its accumulator rows interleave as 4 e + l / 16
    # (acc16_row_f64), which the CholeskyQR2 kernel relies on.
    np.testing.assert_allclose(host(o4).reshape(16, 16), ref, rtol=1e-12, atol=1e-12)
    # 16x16x32 bf16: A = a viewed as 16x32, B = (b viewed 16x32)^T
    a2, b2 = bf16_round(a.reshape(16, 32)).astype(np.float64), bf16_round(b.reshape(16, 32)).astype(np.float64)
    np.testing.assert_allclose(host(o5).reshape(16, 16), a2 @ b2.T, rtol=1e-5, atol=1e-4)


# More than 256 rows with 16-bit products take 256 x 256 tile pairs: column-contiguous rows
# (P == 1: (300,20,18) mode 0, (20,16,600) modes 0-1, (1024,3,5) mode 0), runs of 64
# (P % 64 == 0: (64,520,3) mode 1, (16,20,288) mode 2), partial column chunks and rows
# that fill no whole tile; (4,300,16) mode 1 (P == 4) stays on 128 x 128 tiles.
@pytest.mark.parametrize("dims,lo,hi", [((24, 20, 18), 0, 1), ((24, 20, 18), 1, 2), ((24, 20, 18), 2, 3),
                                        ((64, 32, 40), 1, 2), ((128, 64, 8, 4), 0, 1), ((16, 128, 8, 4), 1, 2),
                                        ((8, 8, 130, 4), 2, 3), ((12, 10, 6), 0, 2), ((3, 5, 7, 2), 1, 3),
                                        ((300, 20, 18), 0, 1), ((20, 16, 600), 0, 2), ((1024, 3, 5), 0, 1),
                                        ((64, 520, 3), 1, 2), ((16, 20, 288), 2, 3), ((4, 300, 16), 1, 2)])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_mode_gram(lib, handle, rng, dims, lo, hi, precision):
    x = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    A = O.unfold(x.astype(np.float64), list(dims), lo, hi)
    M = A.shape[0]
    import torch
    G = torch.zeros(M * M, dtype=torch.float64, device="cuda")
    s = ctypes.c_double()
    xd = dev(x)
    _lib.check(lib.ht_mode_gram(handle.ptr, ptr(xd), len(dims), _lib.i64_array(dims), lo, hi,
                                _lib.PRECISION[precision], ptr(G), ctypes.byref(s)), "gram")
    Gh = host(G).reshape((M, M), order="F")
    if precision == "fp32":
        ref = A @ A.T
        np.testing.assert_allclose(Gh, ref, rtol=1e-5, atol=1e-4 * np.abs(ref).max())
    elif precision == "bf16":
        Ab = O.unfold(bf16_round(x).astype(np.float64), list(dims), lo, hi)
        ref = Ab @ Ab.T
        np.testing.assert_allclose(Gh, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())
        # forward error bound against the exact FP64 Gram: |G_bf16 - G| <= 2^-7 |A||A|^T
        bound = 2.0 ** -7 * (np.abs(A) @ np.abs(A).T) + 1e-6
        assert np.all(np.abs(Gh - A @ A.T) <= bound)
    else:
        Ah = O.unfold(x.astype(np.float16).astype(np.float64), list(dims), lo, hi)
        ref = Ah @ Ah.T
        np.testing.assert_allclose(Gh, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())
    np.testing.assert_allclose(s.value, float(np.sum(x.astype(np.float64) ** 2)), rtol=1e-6)
    assert np.allclose(Gh, Gh.T)


# The 16-bit copy path: the Gram must equal, bit for bit, the one that reads FP32 X and
# rounds it on the way into LDS (same rounded values, same tiles, splits and chunk order;
# an odd chunk count adds one chunk of zeros). Rows contiguous (P == 1, M % 4 == 0:
# (300,20,18) mode 0 with a partial last column chunk, (260,7,9) with overlapped tiles,
# (1024,3,5), (20,16,600) modes 0-1) and runs of 64 ((64,520,3) mode 1, (16,20,288) mode 2,
# (64,1024,2) mode 1); odd tensor sizes exercise the copy's tail.
@pytest.mark.parametrize("dims,lo,hi", [((300, 20, 18), 0, 1), ((260, 7, 9), 0, 1), ((1024, 3, 5), 0, 1),
                                        ((20, 16, 600), 0, 2), ((64, 520, 3), 1, 2), ((16, 20, 288), 2, 3),
                                        ((64, 1024, 2), 1, 2), ((292, 3, 3), 0, 1)])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_mode_gram16_matches_fp32_read(lib, handle, rng, dims, lo, hi, precision):
    import torch
    x = rng.standard_normal(int(np.prod(dims))).astype(np.float32)
    M = int(np.prod(dims[lo:hi]))
    G = torch.zeros(M * M, dtype=torch.float64, device="cuda")
    G16 = torch.zeros(M * M, dtype=torch.float64, device="cuda")
    s, s16 = ctypes.c_double(), ctypes.c_double()
    xd = dev(x)
    args = (len(dims), _lib.i64_array(dims), lo, hi, _lib.PRECISION[precision])
    print(f"mode_gram16 dims={dims} modes {lo}:{hi} {precision}", flush=True)
    _lib.check(lib.ht_mode_gram(handle.ptr, ptr(xd), *args, ptr(G), ctypes.byref(s)), "gram")
    _lib.check(lib.ht_mode_gram16(handle.ptr, ptr(xd), *args, ptr(G16), ctypes.byref(s16)), "gram16")
    assert torch.equal(G, G16)
    np.testing.assert_allclose(s16.value, float(np.sum(x.astype(np.float64) ** 2)), rtol=1e-6)
    # and against the oracle's Gram of the rounded values
    xr = bf16_round(x) if precision == "bf16" else x.astype(np.float16).astype(np.float32)
    A = O.unfold(xr.astype(np.float64), list(dims), lo, hi)
    ref = A @ A.T
    np.testing.assert_allclose(host(G16).reshape((M, M), order="F"), ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


def test_mode_gram16_rejects_other_views(lib, handle, rng):
    import torch
    x = dev(rng.standard_normal(4 * 300 * 16).astype(np.float32))
    G = torch.zeros(300 * 300, dtype=torch.float64, device="cuda")
    # P == 4 (neither contiguous rows nor runs of 64) and FP32 take no 16-bit copy
    for dims, lo, hi, prec in (((4, 300, 16), 1, 2, "bf16"), ((300, 4, 16), 0, 1, "fp32"),
                               ((200, 6, 16), 0, 1, "bf16")):
        rc = lib.ht_mode_gram16(handle.ptr, ptr(x), 3, _lib.i64_array(dims), lo, hi, _lib.PRECISION[prec],
                                ptr(G), None)
        assert _lib.STATUS_NAMES[rc] == "HT_ERR_UNSUPPORTED"


@pytest.mark.parametrize("n1,n2,S", [(128, 128, 6), (64, 32, 10), (20, 12, 7), (128, 16, 3)])
def test_slab_gram(lib, handle, rng, n1, n2, S):
    import torch
    x = rng.standard_normal(n1 * n2 * S).astype(np.float32)
    G1 = torch.zeros(n1 * n1, dtype=torch.float64, device="cuda")
    G2 = torch.zeros(n2 * n2, dtype=torch.float64, device="cuda")
    s = ctypes.c_double()
    xd = dev(x)
    _lib.check(lib.ht_slab_gram(handle.ptr, ptr(xd), n1, n2, S, _lib.PRECISION["bf16"], ptr(G1), ptr(G2),
                                ctypes.byref(s)), "slab_gram")
    xb = bf16_round(x).astype(np.float64)
    dims = [n1, n2, S]
    A1, A2 = O.unfold(xb, dims, 0, 1), O.unfold(xb, dims, 1, 2)
    np.testing.assert_allclose(host(G1).reshape(n1, n1), A1 @ A1.T, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(host(G2).reshape(n2, n2), A2 @ A2.T, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(s.value, float(np.sum(x.astype(np.float64) ** 2)), rtol=1e-6)


@pytest.mark.parametrize("n1,n2,S,r1,r2", [(128, 128, 5, 16, 16), (64, 48, 9, 8, 12), (32, 16, 4, 20, 30),
                                           (16, 128, 3, 32, 5)])
def test_slab_project(lib, handle, rng, n1, n2, S, r1, r2):
    import torch
    x = rng.standard_normal(n1 * n2 * S).astype(np.float32)
    U1 = rng.standard_normal((n1, r1)).astype(np.float32)
    U2 = rng.standard_normal((n2, r2)).astype(np.float32)
    z = torch.zeros(r1 * r2 * S, device="cuda")
    xd, u1d, u2d = dev(x), dev(U1.reshape(-1, order="F")), dev(U2.reshape(-1, order="F"))
    _lib.check(lib.ht_slab_project(handle.ptr, ptr(xd), n1, n2, S, ptr(u1d), r1, ptr(u2d), r2, ptr(z)),
               "slab_project")
    X = x.astype(np.float64).reshape((n1, n2, S), order="F")
    ref = np.einsum("ia,ijs,jb->abs", U1.astype(np.float64), X, U2.astype(np.float64))
    np.testing.assert_allclose(host(z).reshape((r1, r2, S), order="F"), ref, rtol=1e-4, atol=1e-3)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 31, 64])
def test_sym_eig(lib, handle, rng, n):
    import torch
    A = rng.standard_normal((n, n))
    A = A + A.T
    w = torch.zeros(n, dtype=torch.float64, device="cuda")
    V = torch.zeros(n * n, dtype=torch.float64, device="cuda")
    Ad = dev(A.reshape(-1, order="F"))
    _lib.check(lib.ht_sym_eig(handle.ptr, ptr(Ad), n, ptr(w), ptr(V)), "eig")
    wr, Vr = np.linalg.eigh(A)
    wr, Vr = wr[::-1], O.canon_signs(Vr[:, ::-1])
    np.testing.assert_allclose(host(w), wr, rtol=1e-12, atol=1e-12 * np.abs(wr).max())
    Vh = host(V).reshape((n, n), order="F")
    np.testing.assert_allclose(Vh.T @ Vh, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(Vh @ np.diag(host(w)) @ Vh.T, A, atol=1e-11 * np.abs(A).max())
    gaps = np.abs(np.diff(wr))
    if n > 1 and gaps.min() > 1e-3:
        np.testing.assert_allclose(Vh, Vr, atol=1e-9)


@pytest.mark.parametrize("n,k,decay", [(128, 16, "gap"), (256, 16, "flat"), (100, 8, "gap"), (40, 5, "flat")])
def test_sym_topk(lib, handle, rng, n, k, decay):
    import torch
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    if decay == "gap":
        lam = np.concatenate([np.linspace(10, 5, k), 1e-6 * rng.random(n - k)])
    else:
        lam = np.sort(rng.random(n) * 10)[::-1]
    G = (Q * lam) @ Q.T
    G = (G + G.T) / 2
    U = torch.zeros(n * k, dtype=torch.float64, device="cuda")
    w = torch.zeros(k, dtype=torch.float64, device="cuda")
    it = ctypes.c_int32()
    Gd = dev(G.reshape(-1, order="F"))
    _lib.check(lib.ht_sym_topk(handle.ptr, ptr(Gd), n, k, 1e-10, ptr(U), ptr(w), ctypes.byref(it)), "topk")
    np.testing.assert_allclose(host(w), lam[:k], rtol=1e-9)
    Uh = host(U).reshape((n, k), order="F")
    Ur = O.canon_signs(Q[:, :k])
    from scipy.linalg import subspace_angles
    assert np.max(subspace_angles(Uh, Ur)) < 1e-6


@pytest.mark.parametrize("m,b", [(16384, 64), (1000, 24), (300, 17), (256, 64), (40, 40), (30, 40), (70000, 40)])
def test_tsqr(lib, handle, rng, m, b):
    import torch
    W = rng.standard_normal((m, b)).astype(np.float32)
    if m >= b:
        W[:, -1] = W[:, 0] * 1e-3 + W[:, 1]  # mild ill-conditioning
    Wd = dev(W.reshape(-1, order="F"))
    Q = torch.zeros(m * b, device="cuda")
    R = torch.zeros(b * b, dtype=torch.float64, device="cuda")
    _lib.check(lib.ht_tsqr(handle.ptr, ptr(Wd), m, b, ptr(Q), ptr(R)), "tsqr")
    Qh = host(Q).reshape((m, b), order="F").astype(np.float64)
    Rh = host(R).reshape((b, b), order="F")
    kk = min(m, b)
    np.testing.assert_allclose(Qh @ Rh, W, atol=2e-5 * np.abs(W).max())
    if m >= b:
        np.testing.assert_allclose(Qh.T @ Qh, np.eye(b), atol=5e-6)
        Rn = np.linalg.qr(W.astype(np.float64))[1]
        np.testing.assert_allclose(np.abs(Rh), np.abs(Rn), atol=1e-4 * np.abs(Rn).max())
    assert np.allclose(np.tril(Rh[:kk, :kk], -1), 0)


# Shifted CholeskyQR3 (the randomised method's orthonormalisation of its sketches): a
# Gaussian block, one whose column norms span six decades (the first, equilibrated and
# shifted pass has to keep the small directions), and one with a repeated column (rank
# b - 1: the basis still spans Y, from CholeskyQR3 or its TSQR fallback). b = 40 is the
# 1024^3 workload's sketch width (1,048,576 and 1024 rows), b = 64 the largest.
@pytest.mark.parametrize("m,b,kind", [(300000, 40, "gauss"), (262144, 64, "graded"), (270001, 17, "graded"),
                                      (262144, 40, "repeat"), (1024, 40, "gauss"), (1000, 40, "graded"),
                                      (64, 64, "gauss")])
def test_tall_orth(lib, handle, rng, m, b, kind):
    import torch
    Y = rng.standard_normal((m, b)).astype(np.float32)
    if kind == "graded":
        Y *= np.logspace(0, -6, b).astype(np.float32)
    elif kind == "repeat":
        Y[:, 7] = Y[:, 3]
    Yd = dev(Y.reshape(-1, order="F"))
    Q = torch.zeros(m * b, device="cuda")
    fb = ctypes.c_int32(-1)
    _lib.check(lib.ht_tall_orth(handle.ptr, ptr(Yd), m, b, ptr(Q), ctypes.byref(fb)), "tall_orth")
    assert np.array_equal(host(Yd), Y.reshape(-1, order="F"))  # Y is left unchanged
    Qh = host(Q).reshape((m, b), order="F").astype(np.float64)
    Y64 = Y.astype(np.float64)
    print(f"tall_orth {kind} m={m} b={b}: fallback {fb.value}", flush=True)
    if kind != "repeat":
        assert fb.value == 0
        np.testing.assert_allclose(Qh.T @ Qh, np.eye(b), atol=2e-6)
        # Y = Q (Q^T Y): the basis spans Y column by column, the weakest to FP32 accuracy
        resid = Y64 - Qh @ (Qh.T @ Y64)
        assert np.all(np.linalg.norm(resid, axis=0) <= 1e-5 * np.linalg.norm(Y64, axis=0))
    else:
        # b - 1 independent columns: Q's columns are orthonormal up to one direction the
        # data leaves free, and they span Y
        sv = np.linalg.svd(Qh, compute_uv=False)
        assert np.all(sv[:b - 1] > 1 - 1e-5) and sv[0] < 1 + 1e-5
        resid = Y64 - Qh @ np.linalg.lstsq(Qh, Y64, rcond=None)[0]
        assert np.linalg.norm(resid) <= 1e-5 * np.linalg.norm(Y64)


# Shapes cover one and many reduction splits, b in one or two 32-column tiles, rows and
# reduction lengths that fill no whole tile or chunk, and the unaligned (scalar) loads:
# M % 4 != 0 for A^T B, K % 4 != 0 for A B.
@pytest.mark.parametrize("M,K,b,trans", [(4096, 3000, 64, 1), (4096, 3000, 24, 0), (1000, 777, 17, 1),
                                         (130, 5000, 8, 0), (16384, 512, 40, 1), (1022, 3001, 40, 1),
                                         (1024, 65537, 40, 0), (777, 4099, 33, 0), (300, 64, 64, 0),
                                         (62, 70001, 40, 1)])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_skinny_gemm(lib, handle, rng, M, K, b, trans, precision):
    import torch
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal(((M if trans else K), b)).astype(np.float32)
    Y = torch.zeros((K if trans else M) * b, device="cuda")
    Ad, Bd = dev(A.reshape(-1, order="F")), dev(B.reshape(-1, order="F"))
    _lib.check(lib.ht_skinny_gemm(handle.ptr, ptr(Ad), M, K, trans, ptr(Bd), b, _lib.PRECISION[precision], ptr(Y)),
               "gemm")
    rows = K if trans else M
    Yh = host(Y).reshape((rows, b), order="F")
    if precision == "bf16":
        A, B = bf16_round(A), bf16_round(B)
    elif precision == "fp16":
        A, B = A.astype(np.float16), B.astype(np.float16)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64.T @ B64 if trans else A64 @ B64
    np.testing.assert_allclose(Yh, ref, rtol=1e-4, atol=1e-4 * np.abs(ref).max())


# Reading a 16-bit copy of A instead of rounding FP32 A on its way to LDS gives the same
# LDS tiles, splits and chunk order: the products must agree bit for bit. Shapes with one
# and many splits, b in one or two tiles, partial chunks and tiles (M % 8 == 0 for A^T B,
# M % 4 == 0 for A B; K arbitrary), and the randomised method's 1024 x 2^20 shapes.
@pytest.mark.parametrize("M,K,b,trans", [(4096, 3000, 64, 1), (4096, 3000, 24, 0), (1000, 777, 17, 1),
                                         (132, 5000, 8, 0), (16384, 512, 40, 1), (1024, 65537, 40, 0),
                                         (780, 4099, 33, 0), (296, 64, 64, 0), (64, 70001, 40, 1),
                                         (1024, 1 << 20, 40, 1), (1024, 1 << 20, 40, 0)])
@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_skinny_gemm16_matches_fp32_read(lib, handle, rng, M, K, b, trans, precision):
    import torch
    print(f"skinny_gemm16 M={M} K={K} b={b} trans={trans} {precision}", flush=True)
    A = torch.randn(M * K, device="cuda", generator=torch.Generator(device="cuda").manual_seed(M + K))
    B = dev(rng.standard_normal(((M if trans else K), b)).astype(np.float32).reshape(-1, order="F"))
    rows = K if trans else M
    Y, Y16 = torch.zeros(rows * b, device="cuda"), torch.zeros(rows * b, device="cuda")
    p = _lib.PRECISION[precision]
    _lib.check(lib.ht_skinny_gemm(handle.ptr, ptr(A), M, K, trans, ptr(B), b, p, ptr(Y)), "gemm")
    _lib.check(lib.ht_skinny_gemm16(handle.ptr, ptr(A), M, K, trans, ptr(B), b, p, ptr(Y16)), "gemm16")
    assert torch.equal(Y, Y16)
    if M * K <= 1 << 24:  # and the product of the rounded values
        Ah = host(A).reshape((M, K), order="F")
        Bh = host(B).reshape((-1, b), order="F")
        r = bf16_round if precision == "bf16" else (lambda v: v.astype(np.float16).astype(np.float32))
        A64, B64 = r(Ah).astype(np.float64), r(Bh).astype(np.float64)
        ref = A64.T @ B64 if trans else A64 @ B64
        np.testing.assert_allclose(host(Y16).reshape((rows, b), order="F"), ref, rtol=1e-4,
                                   atol=1e-4 * np.abs(ref).max())


def test_skinny_gemm16_rejects_unsupported(lib, handle, rng):
    import torch
    A = dev(rng.standard_normal(1020 * 30).astype(np.float32))
    B = dev(rng.standard_normal(1020 * 8).astype(np.float32))
    Y = torch.zeros(1020 * 8, device="cuda")
    # A^T B with M % 8 != 0, and FP32
    for M, K, trans, prec in ((1020, 30, 1, "bf16"), (1020, 30, 0, "fp32")):
        rc = lib.ht_skinny_gemm16(handle.ptr, ptr(A), M, K, trans, ptr(B), 8, _lib.PRECISION[prec], ptr(Y))
        assert _lib.STATUS_NAMES[rc] == "HT_ERR_UNSUPPORTED"


# FP8 skinny products (OCP E4M3 with a power-of-two scale per operand tile and chunk).
# Small integers |v| <= 8 stay exact under every such scale (k 2^e, k <= 8, needs at most
# three mantissa bits), so the product of integer matrices must come out exact: this pins
# the operand layout, the per-tile scales and their removal. Gaussian data must stay
# within two FP8 roundings (2^-4 each) of the exact product, plus what underflows.
FP8_SHAPES = [(4096, 3000, 64, 1), (1022, 3001, 40, 1), (1024, 65537, 40, 0), (777, 4099, 33, 0),
              (62, 70001, 24, 1), (300, 64, 17, 0)]


@pytest.mark.parametrize("M,K,b,trans", FP8_SHAPES)
def test_skinny_gemm_fp8(lib, handle, rng, M, K, b, trans):
    import torch
    rows = K if trans else M

    def run(A, B):
        Y = torch.zeros(rows * b, device="cuda")
        Ad, Bd = dev(A.reshape(-1, order="F")), dev(B.reshape(-1, order="F"))
        _lib.check(lib.ht_skinny_gemm(handle.ptr, ptr(Ad), M, K, trans, ptr(Bd), b, _lib.PRECISION["fp8"], ptr(Y)),
                   "gemm")
        return host(Y).reshape((rows, b), order="F").astype(np.float64)

    A = rng.integers(-8, 9, size=(M, K)).astype(np.float32)
    B = rng.integers(-8, 9, size=((M if trans else K), b)).astype(np.float32)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64.T @ B64 if trans else A64 @ B64
    np.testing.assert_array_equal(run(A, B), ref)
    A = rng.standard_normal((M, K)).astype(np.float32)
    B = rng.standard_normal(((M if trans else K), b)).astype(np.float32)
    A64, B64 = A.astype(np.float64), B.astype(np.float64)
    ref = A64.T @ B64 if trans else A64 @ B64
    mag = np.abs(A64).T @ np.abs(B64) if trans else np.abs(A64) @ np.abs(B64)
    red = M if trans else K
    Yh = run(A, B)
    bound = 2.0 ** -3 * mag + 2.0 ** -10 / 112 * np.abs(A64).max() * np.abs(B64).max() * red
    assert np.all(np.abs(Yh - ref) <= bound)
    rel = np.linalg.norm(Yh - ref) / np.linalg.norm(ref)
    print(f"fp8 gemm M={M} K={K} b={b} trans={trans}: relative error {rel:.3e}", flush=True)
    assert rel < 0.05


def test_philox_device_matches_oracle(lib, handle):
    import torch
    n = 4099
    out = torch.zeros(n, dtype=torch.float64, device="cuda")
    _lib.check(lib.ht_philox_normals(handle.ptr, 7, 64, n, ptr(out)), "philox")
    np.testing.assert_allclose(host(out), philox.normals(7, 64, n), rtol=1e-13, atol=1e-13)
