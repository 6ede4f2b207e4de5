"""FP64 oracle of one member of the batched workload, for worker processes.

Test infrastructure (imported by tests/test_gpu_configs.py only): each worker regenerates
member b of the seeded batch with the oracle's generator and returns its plain HT
(oracle.ht_oracle.plain_ht) on the tree {1},{2,3}.
"""
import numpy as np

from oracle import gen as G
from oracle import ht_oracle as O


def member(args):
    seed, b, d, c, noise = args
    x = G.batch_member(seed, b, d, c, noise)
    rk = [1, c[0], c[0], c[1], c[2]]
    ref = O.plain_ht(x, d, O.tree_first_vs_rest_3d(), rk)
    err = O.rel_error(x, ref.reconstruct())
    return b, err, {t: ref.U[t] for t in (1, 3, 4)}, ref.B[0], ref.B[2], float(np.abs(x).max())
