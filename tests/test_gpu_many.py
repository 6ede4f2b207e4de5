"""ht_decompose_many: several tensors of one shape with 1-3 decompositions in flight (worker
streams, passes over X ordered on the device) must give what ht_decompose gives one tensor
at a time: exactly, except that with several in flight the slab passes of the leaves-to-root
method (two leading modes of at most 128, multiples of 16) run on 7/8 of the CUs and so split
their partial sums differently (equal to rounding)."""
import numpy as np
import pytest

import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
from oracle import gen as G
from oracle import ht_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("method,precision", [("ltr", "bf16"), ("plain", "fp32")])
def test_many_equals_one_at_a_time(method, precision):
    import torch
    dims, tr, ranks = [48, 40, 32, 24], ht.tree.balanced_4d(), [1, 12, 12, 6, 6, 6, 6]
    xs = [torch.from_numpy(G.tucker(dims, [6, 6, 6, 6], seed=s, noise_rel=1e-2)).cuda() for s in range(5)]
    with ht.Handle() as h1:
        ref = [ht.decompose(x, dims, tr, ranks, method=method, precision=precision, handle=h1) for x in xs]
        ref_p = [r.packed.cpu().numpy() for r in ref]
    e0 = O.rel_error(xs[0].cpu().numpy(), O.plain_ht(xs[0].cpu().numpy(), dims, O.Tree(tr.lo, tr.hi, tr.left,
                                                                                        tr.right), ranks).reconstruct())
    assert abs(ref[0].rel_err - e0) <= 1e-4
    with ht.Handle() as h:
        for inflight in (1, 2, 3):
            for rep in range(2):  # the second call replays the workers' recorded plans
                print(f"{method}: 5 tensors, {inflight} in flight, call {rep}", flush=True)
                res, st = ht.decompose_many(xs, dims, tr, ranks, method=method, precision=precision, handle=h,
                                            inflight=inflight)
                assert st["total_ms"] > 0
                for i, r in enumerate(res):
                    assert r.rel_err == ref[i].rel_err, (inflight, i, r.rel_err, ref[i].rel_err)
                    assert np.array_equal(r.packed.cpu().numpy(), ref_p[i]), (inflight, i)


def test_many_rejects_bad_arguments():
    import torch
    dims, tr, ranks = [16, 16, 16], ht.tree.first_vs_rest_3d(), [1, 4, 4, 4, 4]
    xs = [torch.from_numpy(G.tucker(dims, [4, 4, 4], seed=1, noise_rel=1e-3)).cuda()]
    for inflight in (0, 9):
        with pytest.raises(ht.HTError) as e:
            ht.decompose_many(xs, dims, tr, ranks, inflight=inflight)
        assert e.value.status == 1
    with pytest.raises(ht.HTError) as e:
        ht.decompose_many(xs, dims, tr, [1, 4, 4, 4, 99])
    assert e.value.status == 3


def test_many_slab_shape_matches_to_rounding():
    import torch
    dims, tr, ranks = [64, 64, 32, 16], ht.tree.balanced_4d(), [1, 12, 12, 6, 6, 6, 6]
    xs = [torch.from_numpy(G.tucker(dims, [6, 6, 6, 6], seed=s, noise_rel=1e-2)).cuda() for s in range(4)]
    with ht.Handle() as h1:
        ref = [ht.decompose(x, dims, tr, ranks, method="ltr", precision="bf16", handle=h1) for x in xs]
    with ht.Handle() as h:
        for inflight in (1, 3):
            print(f"slab shape: 4 tensors, {inflight} in flight", flush=True)
            res, _ = ht.decompose_many(xs, dims, tr, ranks, method="ltr", precision="bf16", handle=h,
                                       inflight=inflight)
            for i, r in enumerate(res):
                if inflight == 1:
                    assert r.rel_err == ref[i].rel_err
                    assert np.array_equal(r.packed.cpu().numpy(), ref[i].packed.cpu().numpy())
                else:
                    assert abs(r.rel_err - ref[i].rel_err) <= 1e-7 * ref[i].rel_err, (i, r.rel_err, ref[i].rel_err)
                    np.testing.assert_allclose(r.packed.cpu().numpy(), ref[i].packed.cpu().numpy(), rtol=1e-4,
                                               atol=1e-5)
