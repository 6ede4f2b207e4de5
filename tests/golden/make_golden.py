"""Writes the golden fixtures of this directory from the FP64 oracle (test data only).

The reference ships no test vectors for the HT path, so these fixtures pin the oracle's
own outputs on small seeded inputs: the checks in tests/test_golden.py regenerate them
(CPU, the oracle must reproduce them) and compare the HIP library against them (GPU).
Each .npz holds plain arrays only (numpy.load with allow_pickle=False).

  python tests/golden/make_golden.py        (from the repository root)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import gen as G  # noqa: E402
from oracle import ht_oracle as O  # noqa: E402
from oracle import philox as P  # noqa: E402

CASES = {
    # name: (dims, core, seed, noise, tree, ranks, method)
    "ht3_plain": ([12, 10, 8], [3, 3, 3], 0, 1e-3, "first_vs_rest_3d", [1, 3, 3, 3, 3], "plain"),
    "ht3_plain_ref_tree": ([9, 8, 7], [3, 2, 3], 5, 1e-3, "reference_3d", [1, 4, 3, 3, 2], "plain"),
    "ht4_ltr": ([8, 8, 8, 8], [3, 3, 2, 2], 1, 1e-4, "balanced_4d", [1, 4, 4, 3, 3, 2, 2], "ltr"),
    "ht4_plain": ([8, 6, 8, 6], [3, 3, 3, 3], 7, 1e-3, "balanced_4d", [1, 6, 6, 3, 3, 3, 3], "plain"),
}


def tree_of(name):
    return {"first_vs_rest_3d": O.tree_first_vs_rest_3d, "reference_3d": O.tree_reference_3d,
            "balanced_4d": O.tree_balanced_4d}[name]()


def run_case(name):
    dims, core, seed, noise, tname, ranks, method = CASES[name]
    x = G.tucker(dims, core, seed=seed, noise_rel=noise)
    tr = tree_of(tname)
    if method == "plain":
        ht = O.plain_ht(x, dims, tr, ranks)
    else:
        ht = O.ltr_ht(x, dims, tr, ranks, leaf_groups=[[0, 1], [2, 3]])
    err = O.rel_error(x, ht.reconstruct())
    arrays = {"x": np.asarray(x, dtype=np.float32), "dims": np.array(dims), "ranks": np.array(ranks),
              "rel_err": np.array(err)}
    for t in range(tr.n_nodes):
        if tr.is_leaf(t):
            arrays[f"U{t}"] = ht.U[t]
        else:
            arrays[f"B{t}"] = ht.B[t]
    return arrays


def philox_case():
    out = {}
    for s in (1, 16, 64, 256):
        out[f"normals_1234_{s}"] = P.normals(1234, s, 16)
        out[f"uniforms_1234_{s}"] = P.uniforms(1234, s, 16)
    return out


def main():
    np.savez(os.path.join(HERE, "philox.npz"), **philox_case())
    for name in CASES:
        np.savez(os.path.join(HERE, f"{name}.npz"), **run_case(name))
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
