"""Seeded synthetic inputs for the HT workloads (numpy, FP64 then rounded to FP32).

TEST INFRASTRUCTURE ONLY (see oracle/ht_oracle.py). The HIP library has its own
generator that draws the same Philox streams on the device; tests compare the two.

The reference builds its inputs with ``genHtensor`` from ``srand(time(NULL))``
(3B/kernel.cu:94-202, 4O/kernel.cu:106-248), which is not reproducible. These
generators define the workloads instead:

* ``tucker``: core G (i.i.d. N(0,1), column-major) x_n U_n with U_n the Q factor of a
  Householder QR of an i.i.d. N(0,1) matrix (signs fixed so diag(R) > 0), plus
  i.i.d. Gaussian noise E scaled so ||E||_F = noise_rel * ||X0||_F.
* ``cp_decay``: diagonal core with sigma_k = k**-1.5 (so every mode's singular values
  decay as k**-1.5) times orthonormal factors, plus noise.
* ``bumps``: f(x) = sum_j a_j exp(-||x - c_j||^2 / (2 w_j^2)) on a uniform [0,1]^d grid.

Philox stream ids (key word 1 = 0x5EED0000 ^ stream):
  core 1, factor of mode n 16+n, noise 64, bump parameters 96, rSVD test matrix of
  node t 256+t, batched tensor b: 4096 + 16*b + role.
"""
import numpy as np

from . import philox
from .ht_oracle import prod

STREAM_CORE = 1
STREAM_FACTOR0 = 16
STREAM_NOISE = 64
STREAM_BUMPS = 96
STREAM_OMEGA0 = 256
STREAM_BATCH0 = 4096


def orthonormal_factor(seed, stream, n, r):
    """Q of the Householder QR of an n x r N(0,1) matrix, with diag(R) > 0."""
    A = philox.normals(seed, stream, n * r).reshape((n, r), order="F")
    Q, R = np.linalg.qr(A)
    s = np.sign(np.diag(R))
    s[s == 0] = 1.0
    return Q * s


def mode_product(x, dims, mode, U):
    """x x_mode U (U is new_n x n): flat column-major in, flat column-major out."""
    P = prod(dims[:mode])
    S = prod(dims[mode + 1:])
    y = np.asarray(x).reshape((P, dims[mode], S), order="F")
    z = np.einsum("pms,nm->pns", y, U)
    nd = list(dims)
    nd[mode] = U.shape[0]
    return z.reshape(-1, order="F"), nd


def tucker_expand(core, core_dims, factors):
    y, yd = core, list(core_dims)
    for n, U in enumerate(factors):
        y, yd = mode_product(y, yd, n, U)
    return y, yd


def add_noise(x0, seed, noise_rel, stream=STREAM_NOISE):
    if noise_rel == 0.0:
        return x0
    z = philox.normals(seed, stream, x0.size)
    return x0 + noise_rel * np.linalg.norm(x0) * z / np.linalg.norm(z)


def tucker(dims, core_dims, seed, noise_rel, dtype=np.float32, return_parts=False):
    core = philox.normals(seed, STREAM_CORE, prod(core_dims))
    factors = [orthonormal_factor(seed, STREAM_FACTOR0 + n, dims[n], core_dims[n])
               for n in range(len(dims))]
    x0, _ = tucker_expand(core, core_dims, factors)
    x = add_noise(x0, seed, noise_rel).astype(dtype)
    if return_parts:
        return x, core, factors
    return x


def cp_decay(dims, rank, seed, noise_rel, exponent=1.5, dtype=np.float32):
    core_dims = [rank] * len(dims)
    core = np.zeros(prod(core_dims))
    stride = sum(prod(core_dims[:n]) for n in range(len(dims)))
    for k in range(rank):
        core[k * stride] = (k + 1.0) ** (-exponent)
    factors = [orthonormal_factor(seed, STREAM_FACTOR0 + n, dims[n], rank) for n in range(len(dims))]
    x0, _ = tucker_expand(core, core_dims, factors)
    return add_noise(x0, seed, noise_rel).astype(dtype)


def bump_params(seed, sample, n_bumps=6, order=4):
    """(a, c, w) for one sample of the Gaussian-bump family."""
    per = n_bumps * (2 + order)
    u = philox.uniforms(seed, STREAM_BUMPS, per, start=sample * per).reshape((n_bumps, 2 + order))
    a = 0.5 + 1.5 * u[:, 0]
    w = 0.1 + 0.3 * u[:, 1]
    c = u[:, 2:]
    return a, c, w


def bumps(n, seed, sample, n_bumps=6, order=4, dtype=np.float32):
    """f on the grid x_i = i/(n-1), column-major, as a flat array."""
    a, c, w = bump_params(seed, sample, n_bumps, order)
    g = np.arange(n) / (n - 1.0)
    x = np.zeros([n] * order, order="F")
    for j in range(n_bumps):
        f = [np.exp(-((g - c[j, m]) ** 2) / (2.0 * w[j] ** 2)) for m in range(order)]
        t = a[j] * f[0]
        for m in range(1, order):
            t = np.multiply.outer(t, f[m])
        x += t
    return x.reshape(-1, order="F").astype(dtype)


def batch_shapes(count, seed, lo=16, hi=64, rlo=4, rhi=8, order=3):
    """Per-tensor mode sizes in [lo, hi] and Tucker ranks in [rlo, rhi] (uniform)."""
    u = philox.uniforms(seed, STREAM_BATCH0 - 1, count * 2 * order).reshape((count, 2 * order))
    dims = lo + np.floor(u[:, :order] * (hi - lo + 1)).astype(int)
    ranks = rlo + np.floor(u[:, order:] * (rhi - rlo + 1)).astype(int)
    return np.minimum(dims, hi), np.minimum(ranks, rhi)


def batch_member(seed, b, dims, core_dims, noise_rel, dtype=np.float32):
    base = STREAM_BATCH0 + 16 * b
    core = philox.normals(seed, base, prod(core_dims))
    factors = [orthonormal_factor(seed, base + 1 + n, dims[n], core_dims[n]) for n in range(len(dims))]
    x0, _ = tucker_expand(core, core_dims, factors)
    return add_noise(x0, seed, noise_rel, stream=base + 15).astype(dtype)
