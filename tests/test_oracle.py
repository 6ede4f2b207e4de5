"""Pins the FP64 oracle before anything is compared against it.

The reference ships no golden vectors, known-answer tests or fixtures for its HT path,
so each oracle function is checked against what the mathematics fixes (closed forms,
invariants, brute force on tiny inputs) and against literal restatements of the
reference's index formulas. Each check is chosen so that a plausible mistake (a dropped
term, a transposed operand, a swapped Kronecker order) fails it; test_planted_* plant
such mistakes in copies of the oracle and assert that the checks catch them.
"""
import numpy as np
import pytest
from scipy.linalg import subspace_angles

from oracle import gen as G
from oracle import ht_oracle as O
from oracle import philox


# ---------------------------------------------------------------- Philox (Random123 KATs)
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expect):
    out = philox.philox4x32_10(*[np.array([c], dtype=np.uint64) for c in ctr], key[0], key[1])
    assert tuple(int(o[0]) for o in out) == expect


def test_philox_normals_moments():
    z = philox.normals(5, 1, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    assert np.array_equal(philox.normals(5, 1, 10, start=7), philox.normals(5, 1, 17)[7:])


# ------------------------------------------------ unfoldings vs the reference's formulas
def ref_tensorToMode2(T1, m, n, k):
    """Literal restatement of tensorToMode2 (3B/kernel.cu:45-56)."""
    T2 = np.zeros_like(T1)
    for i in range(m * n * k):
        tube = i // (m * n)
        row = (i - tube * (m * n)) % m
        col = (i - tube * (m * n)) // m
        T2[tube * m * n + row * n + col] = T1[tube * m * n + col * m + row]
    return T2


def ref_tensorToMode3(T1, m, n, k):
    """Literal restatement of tensorToMode3 (3B/kernel.cu:58-69)."""
    T2 = np.zeros_like(T1)
    for i in range(m * n * k):
        tube = i // (m * n)
        row = (i - tube * (m * n)) % m
        col = (i - tube * (m * n)) // m
        T2[k * (col * m + row) + tube] = T1[tube * m * n + col * m + row]
    return T2


def ref_mode2(A, m, n, r):
    """Literal restatement of mode2 (4O/kernel.cu:51-63)."""
    B = np.zeros_like(A)
    for i in range(m * r * n):
        row, col, ge = i // n, i % n, i // (m * n)
        B[i] = A[(row - ge * m) + (col * m + ge * m * n)]
    return B


def test_unfold_matches_reference_permutation_kernels():
    rng = np.random.default_rng(0)
    a, b, c = 5, 4, 3
    x = rng.standard_normal(a * b * c)
    # mode-2 / mode-3 unfoldings of a 3rd-order tensor
    np.testing.assert_array_equal(ref_tensorToMode2(x, a, b, c).reshape((b, a * c), order="F"),
                                  O.unfold(x, [a, b, c], 1, 2))
    np.testing.assert_array_equal(ref_tensorToMode3(x, a, b, c).reshape((c, a * b), order="F"),
                                  O.unfold(x, [a, b, c], 2, 3))
    np.testing.assert_array_equal(ref_mode2(x, a, b, c).reshape((b, a * c), order="F"), O.unfold(x, [a, b, c], 1, 2))
    # 4th order, as 4B/htd4.cu:38-42 uses them: (ab, c, d) and (a, b, cd) views
    d4 = [3, 4, 5, 2]
    y = rng.standard_normal(int(np.prod(d4)))
    np.testing.assert_array_equal(ref_tensorToMode2(y, 3, 4, 10).reshape((4, 30), order="F"), O.unfold(y, d4, 1, 2))
    np.testing.assert_array_equal(ref_tensorToMode2(y, 12, 5, 2).reshape((5, 24), order="F"), O.unfold(y, d4, 2, 3))
    np.testing.assert_array_equal(ref_tensorToMode3(y, 12, 5, 2).reshape((2, 60), order="F"), O.unfold(y, d4, 3, 4))
    # mode 1 and the {1,2} matricisation are views of X (3O/htd_tensorcore.cu:98-104, :249)
    np.testing.assert_array_equal(O.unfold(y, d4, 0, 1), y.reshape((3, 40), order="F"))
    np.testing.assert_array_equal(O.unfold(y, d4, 0, 2), y.reshape((12, 10), order="F"))
    # fold inverts unfold
    for lo, hi in [(0, 1), (1, 3), (2, 4), (1, 2)]:
        np.testing.assert_array_equal(O.fold(O.unfold(y, d4, lo, hi), d4, lo, hi), y)


# ----------------------------------------------------------- ttm layout (B_t contraction)
def ref_ttm(U1, U2, U3, a, b, k1, k2, k3):
    """Literal restatement of ttm (3O/ttm.cu:2-36): B(k2,k3,k1) = sum U1(a,b,i) U2(a,k2) U3(b,k3),
    U1 stored a x b x k1 column-major, result k2 x k3 x k1 column-major."""
    B = np.zeros(k2 * k3 * k1)
    for i in range(k1):
        for p in range(k2):
            for q in range(k3):
                s = 0.0
                for ia in range(a):
                    for ib in range(b):
                        s += U1[ia + a * ib + a * b * i] * U2[ia + a * p] * U3[ib + b * q]
                B[p + k2 * q + k2 * k3 * i] = s
    return B


def test_contract_transfer_matches_reference_ttm():
    rng = np.random.default_rng(1)
    a, b, k1, k2, k3 = 4, 3, 2, 3, 2
    Ut = rng.standard_normal((a * b, k1))
    Ul = rng.standard_normal((a, k2))
    Ur = rng.standard_normal((b, k3))
    mine = O.contract_transfer(Ut, Ul, Ur, a, b).reshape(-1, order="F")
    ref = ref_ttm(Ut.reshape(-1, order="F"), Ul.reshape(-1, order="F"), Ur.reshape(-1, order="F"), a, b, k1, k2, k3)
    np.testing.assert_allclose(mine, ref, rtol=1e-13, atol=1e-13)
    # expand_node inverts it when U_t lies in span(U_r (x) U_l)
    Ql, _ = np.linalg.qr(Ul)
    Qr, _ = np.linalg.qr(Ur)
    Bt = rng.standard_normal((k2, k3, k1))
    Ut2 = O.expand_node(Bt, Ql, Qr)
    np.testing.assert_allclose(O.contract_transfer(Ut2, Ql, Qr, a, b), Bt, atol=1e-12)
    np.testing.assert_allclose(Ut2, np.kron(Qr, Ql) @ Bt.reshape((k2 * k3, k1), order="F"), atol=1e-12)


# ----------------------------------------------------------------------- closed forms
def test_first_vs_rest_tree_error_is_a_single_svd_tail():
    """Tree {1},{2,3} with full leaf ranks below {2,3}: the HT error equals the truncated
    SVD tail of X_(1) (the nodes {1} and {2,3} are complementary)."""
    dims = [9, 5, 4]
    x = G.tucker(dims, [3, 3, 3], seed=11, noise_rel=0.05)
    tr = O.tree_first_vs_rest_3d()
    for r in (1, 2, 3, 5):
        ranks = [1, r, r, 5, 4]
        h = O.plain_ht(x, dims, tr, ranks)
        s = np.linalg.svd(O.unfold(x.astype(np.float64), dims, 0, 1), compute_uv=False)
        closed = np.sqrt(np.sum(s[r:] ** 2)) / np.linalg.norm(x.astype(np.float64))
        assert abs(O.rel_error(x, h.reconstruct()) - closed) < 1e-12
        h2 = O.ltr_ht(x, dims, tr, ranks)
        assert abs(O.rel_error(x, h2.reconstruct()) - closed) < 1e-12


@pytest.mark.parametrize("tree,dims,core,ranks", [
    (O.tree_reference_3d, [8, 7, 6], [2, 3, 2], [1, 6, 2, 2, 3]),
    (O.tree_first_vs_rest_3d, [8, 7, 6], [3, 3, 3], [1, 3, 9, 3, 3]),
    (O.tree_balanced_4d, [5, 6, 4, 5], [2, 2, 3, 2], [1, 4, 6, 2, 2, 3, 2]),
])
def test_exact_low_rank_is_recovered(tree, dims, core, ranks):
    x = G.tucker(dims, core, seed=2, noise_rel=0.0, dtype=np.float64)
    tr = tree()
    for h in (O.plain_ht(x, dims, tr, ranks), O.ltr_ht(x, dims, tr, ranks),
              O.plain_ht(x, dims, tr, ranks, leaf="gram")):
        assert O.rel_error(x, h.reconstruct()) < 1e-12


def test_two_sided_ht_bounds():
    """max_t tail_t <= ||X - X_HT|| <= sqrt(sum_t tail_t^2) for the root-to-leaves HT-SVD
    (Grasedyck 2010), and the upper bound is within sqrt(2d-3) of the best."""
    dims = [6, 5, 7, 4]
    x = G.tucker(dims, [3, 3, 3, 3], seed=3, noise_rel=0.3, dtype=np.float64)
    tr = O.tree_balanced_4d()
    ranks = [1, 4, 4, 2, 2, 2, 2]
    h = O.plain_ht(x, dims, tr, ranks)
    e = O.rel_error(x, h.reconstruct())
    upper = O.ht_error_bound(x, dims, tr, ranks)
    nx = np.linalg.norm(x)
    tails = []
    for t in range(1, tr.n_nodes):
        s = np.linalg.svd(O.unfold(x, dims, tr.lo[t], tr.hi[t]), compute_uv=False)
        tails.append(np.sqrt(np.sum(s[ranks[t]:] ** 2)) / nx)
    assert max(tails) - 1e-12 <= e <= upper + 1e-12
    assert upper <= np.sqrt(2 * 4 - 3) * max(tails) * np.sqrt(2) + 1e-12


def test_ltr_equals_exact_when_leaf_ranks_match_the_tucker_rank():
    """The 4D configuration's structure (i.i.d. core, orthonormal factors, small noise):
    leaves-to-root reproduces the exact root-to-leaves HT error (the interior solves are
    exact SVDs of the leaf core's matricisations)."""
    dims = [12] * 4
    x = G.tucker(dims, [4] * 4, seed=1, noise_rel=1e-4)
    tr = O.tree_balanced_4d()
    ranks = [1, 4, 4, 4, 4, 4, 4]
    e_exact = O.rel_error(x, O.plain_ht(x, dims, tr, ranks).reconstruct())
    e_ltr = O.rel_error(x, O.ltr_ht(x, dims, tr, ranks).reconstruct())
    assert abs(e_exact - e_ltr) < 1e-8
    assert e_exact > 0.5  # the flat-spectrum regime: the error is dominated by the core


def test_gram_leaves_equal_svd_leaves():
    rng = np.random.default_rng(4)
    A = rng.standard_normal((10, 40)) * (0.7 ** np.arange(10))[:, None]
    U1, _ = O.top_left_singular(A, 4)
    U2, w = O.gram_topk(A @ A.T, 4)
    np.testing.assert_allclose(U1, U2, atol=1e-10)
    np.testing.assert_allclose(np.sqrt(w), np.linalg.svd(A, compute_uv=False)[:4], rtol=1e-12)


def test_brute_force_tiny_full_rank():
    """Full ranks everywhere: the HT format is exact (2x3x2 tensor, every node full)."""
    dims = [2, 3, 2]
    x = np.random.default_rng(5).standard_normal(12)
    tr = O.tree_first_vs_rest_3d()
    h = O.plain_ht(x, dims, tr, [1, 2, 6, 3, 2])
    assert O.rel_error(x, h.reconstruct()) < 1e-14


# ---------------------------------------------------------------------- rSVD variants
def test_rsvd_baseline_exact_for_low_rank():
    rng = np.random.default_rng(6)
    A = rng.standard_normal((60, 5)) @ rng.standard_normal((5, 80))
    om = philox.normals(1234, 256, 80 * 5).reshape((80, 5), order="F")
    U, s = O.rsvd_baseline(A, 5, 3, om)
    Ue, se = O.top_left_singular(A, 5)
    np.testing.assert_allclose(s, se, rtol=1e-10)
    np.testing.assert_allclose(np.abs(U.T @ Ue), np.eye(5), atol=1e-8)


def test_rsvd_opt_reference_quirk():
    """4O's rsvd takes the SVD of the first ks columns of Q^T A only: when rank(A) == k the
    returned subspace is exact but its vectors are not A's singular vectors; in general
    it is not the truncated SVD."""
    rng = np.random.default_rng(7)
    A = rng.standard_normal((50, 4)) @ rng.standard_normal((4, 70))
    om = rng.standard_normal((70, 12))
    U = O.rsvd_opt_reference(A, 4, om)
    Ue, _ = O.top_left_singular(A, 4)
    assert np.max(subspace_angles(U, Ue)) < 1e-8          # the subspace is exact ...
    assert np.max(np.abs(np.abs(U.T @ Ue) - np.eye(4))) > 1e-3  # ... its basis is not A's singular vectors
    B = rng.standard_normal((50, 70)) * np.linspace(1, 0.1, 70)
    U = O.rsvd_opt_reference(B, 4, rng.standard_normal((70, 12)))
    Ue, _ = O.top_left_singular(B, 4)
    assert np.max(np.abs(np.abs(U.T @ Ue) - np.eye(4))) > 1e-3


def test_rsvd_power_improves_on_flat_spectrum():
    dims = [10] * 4
    x = G.tucker(dims, [4] * 4, seed=1, noise_rel=1e-3, dtype=np.float64)
    A = O.unfold(x, dims, 0, 2)
    om = philox.normals(1, 256, 100 * 8).reshape((100, 8), order="F")
    e = []
    for q in (0, 1, 4):
        U, _ = O.rsvd_power(A, 4, 4, q, om)
        e.append(np.linalg.norm(A - U @ (U.T @ A)))
    Ue, _ = O.top_left_singular(A, 4)
    best = np.linalg.norm(A - Ue @ (Ue.T @ A))
    assert e[0] >= e[1] >= e[2] >= best - 1e-9


# ------------------------------------------------------------------ TensorLy restatement
def test_tensorly_fp32_port_matches_oracle():
    dims = [16, 12, 10]
    x = G.tucker(dims, [4, 4, 4], seed=0, noise_rel=1e-3)
    tr = O.tree_reference_3d()
    ranks = [1, 8, 4, 4, 4]
    _, e32 = O.tensorly_ht_fp32(x, dims, tr, ranks)
    e64 = O.rel_error(x, O.plain_ht(x, dims, tr, ranks).reconstruct())
    assert abs(e32 - e64) < 2e-5


# -------------------------------------------------------------------------- generators
def test_generator_properties():
    dims, core = [20, 16, 12], [4, 3, 5]
    x, c, U = G.tucker(dims, core, seed=3, noise_rel=1e-3, dtype=np.float64, return_parts=True)
    for n, Un in enumerate(U):
        np.testing.assert_allclose(Un.T @ Un, np.eye(core[n]), atol=1e-12)
    x0, _ = G.tucker_expand(c, core, U)
    np.testing.assert_allclose(np.linalg.norm(x - x0) / np.linalg.norm(x0), 1e-3, rtol=1e-9)
    np.testing.assert_allclose(np.linalg.norm(x0), np.linalg.norm(c), rtol=1e-12)
    # diagonal core: every mode spectrum is k^-1.5
    y = G.cp_decay([15, 14, 13], 6, seed=2, noise_rel=0.0, dtype=np.float64)
    for m in range(3):
        s = np.linalg.svd(O.unfold(y, [15, 14, 13], m, m + 1), compute_uv=False)[:6]
        np.testing.assert_allclose(s, np.arange(1, 7) ** -1.5, rtol=1e-10)
    # Gaussian bumps: CP rank <= 6, so every matricisation has rank <= 6
    z = G.bumps(16, seed=4, sample=0).astype(np.float64)
    for lo, hi in [(0, 1), (0, 2), (2, 4)]:
        s = np.linalg.svd(O.unfold(z, [16] * 4, lo, hi), compute_uv=False)
        assert s[6] < 1e-5 * s[0]


# ------------------------------------------------------------------ planted mistakes
def test_planted_swapped_kron_order_is_caught():
    """A transfer tensor built with the children swapped fails the ttm restatement and
    the exact-recovery check."""
    rng = np.random.default_rng(8)
    a, b, k1, k2, k3 = 4, 3, 2, 3, 2
    Ut = rng.standard_normal((a * b, k1))
    Ul = rng.standard_normal((a, k2))
    Ur = rng.standard_normal((b, k3))

    def wrong_contract(Ut, Ul, Ur, nl, nr):
        kt = Ut.shape[1]
        return np.einsum("ia,jb,jic->abc", Ul, Ur, Ut.reshape((nr, nl, kt), order="F"))

    ref = ref_ttm(Ut.reshape(-1, order="F"), Ul.reshape(-1, order="F"), Ur.reshape(-1, order="F"), a, b, k1, k2, k3)
    assert not np.allclose(wrong_contract(Ut, Ul, Ur, a, b).reshape(-1, order="F"), ref)


def test_planted_transposed_unfolding_is_caught():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(60)

    def wrong_unfold(x, dims, lo, hi):
        return np.asarray(x).reshape(dims, order="C").reshape(dims[lo], -1)

    assert not np.array_equal(ref_tensorToMode2(x, 5, 4, 3).reshape((4, 15), order="F"),
                              wrong_unfold(x, [5, 4, 3], 1, 2))


def test_planted_dropped_noise_term_is_caught():
    """An error formula that drops the cross term (||X||^2 - ||X_HT||^2 with a
    non-orthogonal reconstruction) disagrees with the residual definition."""
    dims = [6, 5, 4]
    x = G.tucker(dims, [2, 2, 2], seed=5, noise_rel=0.2, dtype=np.float64)
    h = O.plain_ht(x, dims, O.tree_first_vs_rest_3d(), [1, 2, 2, 2, 2])
    xr = h.reconstruct()
    B = h.B[0]
    h.B[0] = B * 1.3  # a wrong root transfer tensor
    xw = h.reconstruct()
    identity = np.sqrt(max(np.sum(x ** 2) - np.sum(xw ** 2), 0)) / np.linalg.norm(x)
    assert abs(identity - O.rel_error(x, xw)) > 1e-3
    assert abs(O.rel_error(x, xr) - np.sqrt(np.sum(x ** 2) - np.sum(xr ** 2)) / np.linalg.norm(x)) < 1e-10


# ------------------------------------------------------- randomised root-to-leaves HT
def _rsvd_variant(old, new):
    """A copy of oracle.rsvd_ht with one source fragment replaced (a planted mistake)."""
    import inspect
    src = inspect.getsource(O.rsvd_ht)
    assert old in src
    ns = dict(vars(O))
    exec(compile(src.replace(old, new), "<planted rsvd_ht>", "exec"), ns)
    return ns["rsvd_ht"]


def _rsvd_matches_plain_on_captured_range(rsvd):
    """When the sketched matricisation X_(12) has rank <= l = k + p the range finder spans
    it exactly, so the randomised HT equals the exact HT (and the complementary sibling
    {3,4} gets the exact basis from the right singular vectors)."""
    tr = O.tree_balanced_4d()
    dims = [6, 5, 5, 6]  # X_(12) square, so a sibling given the wrong side still has the right shape
    x = G.tucker(dims, [2, 2, 2, 2], seed=3, noise_rel=0.0, dtype=np.float64)
    ranks = [1, 3, 3, 2, 2, 2, 2]  # rank(X_(12)) = 4 > 3: the interior nodes truncate
    e = O.plain_ht(x, dims, tr, ranks)
    ee = O.rel_error(x, e.reconstruct())
    ok = True
    for q in (0, 1):
        r = rsvd(x, dims, tr, ranks, p=8, q=q)
        ok &= abs(O.rel_error(x, r.reconstruct()) - ee) <= 1e-12
        for t in range(1, tr.n_nodes):
            ok &= np.max(subspace_angles(r.full_basis(t), e.full_basis(t))) <= 1e-10
    return ok, ee


def test_rsvd_ht_equals_plain_when_the_sketch_captures_the_range():
    ok, ee = _rsvd_matches_plain_on_captured_range(O.rsvd_ht)
    assert ok
    assert ee > 1e-2  # the interior truncation is not trivial


def test_rsvd_ht_converges_to_plain_with_power_iterations():
    """q power iterations drive the range finder to the dominant subspace: on a noisy
    input the error gap to the exact HT shrinks with q and vanishes at q = 30."""
    tr = O.tree_balanced_4d()
    dims = [6, 6, 6, 6]
    x = G.tucker(dims, [3, 3, 3, 3], seed=4, noise_rel=0.3, dtype=np.float64)
    ranks = [1, 4, 4, 3, 3, 3, 3]
    e = O.plain_ht(x, dims, tr, ranks)
    ee = O.rel_error(x, e.reconstruct())
    gaps = []
    for q in (0, 2, 30):
        r = O.rsvd_ht(x, dims, tr, ranks, p=2, q=q)
        gaps.append(abs(O.rel_error(x, r.reconstruct()) - ee))
        ang = max(np.max(subspace_angles(r.full_basis(t), e.full_basis(t))) for t in (1, 2))
    assert gaps[0] > 1e-2 and gaps[1] < gaps[0] and gaps[2] < 1e-12
    assert ang < 1e-10


def test_rsvd_ht_uses_the_documented_test_matrix():
    """At q = 0 node t's basis is the top-k left singular subspace of A projected onto
    range(A Omega_t), Omega_t = the FP32-rounded Philox normals of stream 256 + t (the
    engine's test matrix), formed here as P_Y A instead of Q^T A."""
    tr = O.tree_balanced_4d()
    dims = [5, 6, 4, 5]
    x = G.tucker(dims, [3, 3, 3, 3], seed=6, noise_rel=0.5, dtype=np.float64)
    ranks = [1, 4, 4, 3, 3, 3, 3]
    r = O.rsvd_ht(x, dims, tr, ranks, p=3, q=0, seed=77)
    A = O.unfold(x, dims, 0, 2)
    l = 4 + 3
    om = philox.normals(77, 256 + 1, A.shape[1] * l).astype(np.float32).astype(np.float64)
    Y = A @ om.reshape((A.shape[1], l), order="F")
    Qy, _ = np.linalg.qr(Y)
    U, _ = O.top_left_singular(Qy @ (Qy.T @ A), 4)
    # (U[t] is the node's own basis; full_basis(t) would be its projection onto the
    # children's bases)
    assert np.max(subspace_angles(r.U[1], U)) < 1e-10
    # the sibling {3,4}: right singular vectors of the same projected problem
    _, _, Vt = np.linalg.svd(Qy @ (Qy.T @ A), full_matrices=False)
    assert np.max(subspace_angles(r.U[2], Vt[:4].T)) < 1e-10


def test_planted_rsvd_sibling_left_vectors_is_caught():
    """A sibling that takes the node's left instead of the right singular vectors fails
    the captured-range closed form."""
    bad = _rsvd_variant("canon_signs(Vbt.T[:, :kW])", "canon_signs((Q @ Ub)[:, :kW])")
    ok, _ = _rsvd_matches_plain_on_captured_range(bad)
    assert not ok


def test_planted_rsvd_wrong_stream_is_caught(monkeypatch):
    """A test matrix from another Philox stream fails the documented-test-matrix check."""
    bad = _rsvd_variant("philox.normals(seed, 256 + t,", "philox.normals(seed, 257 + t,")
    monkeypatch.setattr(O, "rsvd_ht", bad)
    with pytest.raises(AssertionError):
        test_rsvd_ht_uses_the_documented_test_matrix()


def test_planted_rsvd_dropped_power_iteration_is_caught():
    """An rSVD that ignores q fails the convergence check."""
    bad = _rsvd_variant("for _ in range(q):", "for _ in range(0):")
    tr = O.tree_balanced_4d()
    dims = [6, 6, 6, 6]
    x = G.tucker(dims, [3, 3, 3, 3], seed=4, noise_rel=0.3, dtype=np.float64)
    ranks = [1, 4, 4, 3, 3, 3, 3]
    ee = O.rel_error(x, O.plain_ht(x, dims, tr, ranks).reconstruct())
    assert abs(O.rel_error(x, bad(x, dims, tr, ranks, p=2, q=30).reconstruct()) - ee) > 1e-3



def _middle_node_case():
    """Order 5, tree ((1,(2,3)),(4,5)): node {2,3} lies strictly inside the mode order."""
    dims = [6, 5, 4, 5, 6]
    tr = O.tree_from_spec(((0, (1, 2)), (3, 4)))
    mid = [t for t in tr.interior() if t != 0 and tr.lo[t] != 0 and tr.hi[t] != len(dims)]
    assert len(mid) == 1
    ranks = [1] + [3] * (tr.n_nodes - 1)
    x = G.tucker(dims, [3] * 5, seed=2, noise_rel=0.5, dtype=np.float64)
    return x, dims, tr, ranks, mid[0]


def _middle_angle(rsvd):
    x, dims, tr, ranks, t = _middle_node_case()
    r = rsvd(x, dims, tr, ranks, p=0, q=0)
    e = O.plain_ht(x, dims, tr, ranks)
    return float(np.max(subspace_angles(r.U[t], e.U[t])))


def test_rsvd_ht_middle_node_takes_the_exact_basis():
    """A node strictly inside the mode order takes its Gram eigenvectors: the exact top-k
    left singular subspace of X_(t), even with no oversampling or power iteration on a
    half-noise input."""
    assert _middle_angle(O.rsvd_ht) < 1e-8


def test_planted_rsvd_middle_node_sketched_is_caught():
    bad = _rsvd_variant("if not (tree.lo[t] == 0 or tree.hi[t] == len(dims)):", "if False:")
    assert _middle_angle(bad) > 1e-3

# ---------------------------------------------------------------- rank-adaptive truncation
ADAPT_DIMS = [8, 7, 6, 8]
ADAPT_FULL = [1, 48, 48, 8, 7, 6, 8]   # the full rank of every matricisation: no cap binds


def _adaptive_input(seed):
    return G.tucker(ADAPT_DIMS, [4, 4, 4, 4], seed=seed, noise_rel=0.3, dtype=np.float64)


@pytest.mark.parametrize("eps", [0.05, 0.1, 0.2, 0.3])
@pytest.mark.parametrize("seed", [0, 1])
def test_adaptive_error_within_rel_eps(seed, eps):
    """With caps that do not bind, ||X - X_HT|| <= rel_eps ||X|| for both truncation orders,
    and the tail-sum bound of the chosen ranks (ht_error_bound) also sits below rel_eps."""
    x = _adaptive_input(seed)
    tr = O.tree_balanced_4d()
    for fn in (O.plain_ht_adaptive, O.ltr_ht_adaptive):
        h, ranks = fn(x, ADAPT_DIMS, tr, ADAPT_FULL, eps)
        assert O.rel_error(x, h.reconstruct()) <= eps
    _, ranks = O.plain_ht_adaptive(x, ADAPT_DIMS, tr, ADAPT_FULL, eps)
    assert O.ht_error_bound(x, ADAPT_DIMS, tr, ranks) <= eps


def test_adaptive_ranks_are_minimal_for_plain():
    """Lowering any one node's chosen rank (keeping the shared root-children rank and the
    nesting) pushes that node's own SVD tail past the per-node budget."""
    x = _adaptive_input(0)
    tr = O.tree_balanced_4d()
    eps = 0.2
    _, ranks = O.plain_ht_adaptive(x, ADAPT_DIMS, tr, ADAPT_FULL, eps)
    xsq = float(np.sum(x * x))
    tau2 = eps ** 2 * xsq / (2 * len(ADAPT_DIMS) - 3)
    for t in range(1, tr.n_nodes):
        s = np.linalg.svd(O.unfold(x, ADAPT_DIMS, tr.lo[t], tr.hi[t]), compute_uv=False)
        tail = lambda r: float(np.sum(s[r:] ** 2))  # noqa: E731
        assert tail(ranks[t]) <= tau2 * (1 + 1e-12)
        if t not in (tr.left[0], tr.right[0]) and ranks[t] > 1:
            assert tail(ranks[t] - 1) > tau2


def test_adaptive_closed_form_tucker_ranks():
    """An exact Tucker tensor (plus 1e-6 noise) with core (3,4,2,5): a small rel_eps recovers
    the multilinear ranks at the leaves and min(3*4, 2*5) = 10 at the two interior nodes."""
    dims = [9, 8, 7, 10]
    x = G.tucker(dims, [3, 4, 2, 5], seed=3, noise_rel=1e-6, dtype=np.float64)
    tr = O.tree_balanced_4d()
    caps = [1, 20, 20, 9, 8, 7, 10]
    for fn in (O.plain_ht_adaptive, O.ltr_ht_adaptive):
        h, ranks = fn(x, dims, tr, caps, 1e-3)
        assert ranks == [1, 10, 10, 3, 4, 2, 5], (fn.__name__, ranks)
        assert O.rel_error(x, h.reconstruct()) <= 1e-5
    # 3-D, tree {1,2},{3}: leaves (4,3), node {1,2} and node {3} both rank 5
    x3 = G.tucker([12, 10, 9], [4, 3, 5], seed=4, noise_rel=1e-6, dtype=np.float64)
    tr3 = O.tree_reference_3d()
    h, ranks = O.plain_ht_adaptive(x3, [12, 10, 9], tr3, [1, 9, 9, 12, 10], 1e-3)
    assert ranks == [1, 5, 5, 4, 3]


def test_adaptive_max_rank_caps_and_nests():
    x = _adaptive_input(1)
    tr = O.tree_balanced_4d()
    for fn in (O.plain_ht_adaptive, O.ltr_ht_adaptive):
        _, ranks = fn(x, ADAPT_DIMS, tr, ADAPT_FULL, 1e-6, max_rank=3)
        assert ranks[1:] == [3] * 6, ranks
    caps = O.adaptive_caps(tr, [1, 48, 48, 2, 2, 6, 8], 0)
    assert caps[1] == 4 and caps[2] == 48


def _planted_adaptive(fn, old, new):
    import inspect
    src = inspect.getsource(fn)
    assert old in src
    ns = dict(vars(O))
    exec(src.replace(old, new), ns)
    return ns[fn.__name__]


def test_planted_adaptive_budget_count_is_caught():
    """Spending the whole budget at every node (dividing by 1 instead of the number of
    truncations) breaks the rel_eps guarantee that the bound test checks."""
    bad_plain = _planted_adaptive(O.plain_ht_adaptive, "max(1, 2 * len(dims) - 3)", "1")
    bad_ltr = _planted_adaptive(O.ltr_ht_adaptive, "max(1, 2 * d - 2)", "1")
    x = _adaptive_input(0)
    tr = O.tree_balanced_4d()
    for bad in (bad_plain, bad_ltr):
        h, _ = bad(x, ADAPT_DIMS, tr, ADAPT_FULL, 0.1)
        assert O.rel_error(x, h.reconstruct()) > 0.1

