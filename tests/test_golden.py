"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle and the library's host Philox reproduce the fixtures.
GPU: the HIP library matches the fixtures on the fixtures' own inputs.
"""
import ctypes
import os
import sys

import numpy as np
import pytest
from scipy.linalg import subspace_angles

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)
import make_golden as MG  # noqa: E402

from oracle import philox as P  # noqa: E402


def load(name):
    with np.load(os.path.join(GOLD, f"{name}.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_philox_golden_bit_exact():
    g = load("philox")
    for s in (1, 16, 64, 256):
        assert np.array_equal(P.normals(1234, s, 16), g[f"normals_1234_{s}"])
        assert np.array_equal(P.uniforms(1234, s, 16), g[f"uniforms_1234_{s}"])


def test_library_host_philox_matches_golden(lib):
    g = load("philox")
    for s in (1, 16, 64, 256):
        out = (ctypes.c_double * 16)()
        lib.ht_philox_normals_host(1234, s, 16, out)
        np.testing.assert_allclose(np.array(out), g[f"normals_1234_{s}"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_oracle_reproduces_golden(name):
    g = load(name)
    r = MG.run_case(name)
    assert set(r) == set(g)
    for k in g:
        np.testing.assert_allclose(r[k], g[k], rtol=1e-10, atol=1e-12, err_msg=f"{name}:{k}")


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(MG.CASES))
def test_gpu_matches_golden(name):
    import torch
    import ht_tensor_learning_using_gpu_tensor_cores_amd as ht
    g = load(name)
    dims, core, seed, noise, tname, ranks, method = MG.CASES[name]
    tr = {"first_vs_rest_3d": ht.tree.first_vs_rest_3d, "reference_3d": ht.tree.reference_3d,
          "balanced_4d": ht.tree.balanced_4d}[tname]()
    x = torch.from_numpy(g["x"]).cuda()
    print(f"golden {name}: dims {dims} ranks {ranks} method {method}", flush=True)
    res = ht.decompose(x, dims, tr, ranks, method=method, precision="fp32")
    torch.cuda.synchronize()
    tol = 1e-5 if method == "plain" else 1e-4
    assert abs(res.rel_err - float(g["rel_err"])) <= tol, (res.rel_err, float(g["rel_err"]))
    for t in range(len(ranks)):
        if f"U{t}" in g:
            Ug = res.U[t].detach().cpu().numpy().astype(np.float64)
            ks = min(ranks[t], core[tr.lo[t]])
            assert np.max(subspace_angles(Ug[:, :ks], g[f"U{t}"][:, :ks])) < 1e-3
            # canonical signs: the signal columns agree entrywise
            np.testing.assert_allclose(Ug[:, :ks], g[f"U{t}"][:, :ks], atol=2e-3)
