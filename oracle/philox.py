"""Philox4x32-10 counter-based RNG and Box-Muller normals (numpy, vectorised).

TEST INFRASTRUCTURE ONLY. This module belongs to the CPU oracle: only tests/,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may use it,
and only as the checker. The product (the HIP library) carries its own device-side
Philox and is never routed through this file.

The reference seeds its inputs with ``srand(time(NULL))`` (3B/kernel.cu:96,
4O/kernel.cu:108) and its rSVD test matrix with cuRAND seed 1234
(3O/kernel.cu:604-607); neither stream is reproducible or reproduced here. The
synthetic inputs are instead defined by this published generator (Salmon, Moraes,
Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11): the same
integers come out of the device implementation, so CPU and GPU see the same draws.

Stream layout (shared with the device code):
  key     = (seed & 0xffffffff, 0x5EED0000 ^ stream)
  counter = (j & 0xffffffff, j >> 32, 0, 0) for block j = element_index // 4
  normal[i] = Box-Muller of the block's words: words (0,1) give z0 (cos), z1 (sin);
              words (2,3) give z2 (cos), z3 (sin); u = (w + 0.5) * 2**-32.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint64(0x9E3779B9)
W1 = np.uint64(0xBB67AE85)
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds; all arguments uint32-valued arrays."""
    c0 = np.asarray(c0, dtype=np.uint64)
    c1 = np.asarray(c1, dtype=np.uint64)
    c2 = np.asarray(c2, dtype=np.uint64)
    c3 = np.asarray(c3, dtype=np.uint64)
    k0 = np.asarray(k0, dtype=np.uint64) & MASK
    k1 = np.asarray(k1, dtype=np.uint64) & MASK
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & MASK, lo1, (hi0 ^ c3 ^ k1) & MASK, lo0
        k0 = (k0 + W0) & MASK
        k1 = (k1 + W1) & MASK
    return (c0.astype(np.uint32), c1.astype(np.uint32),
            c2.astype(np.uint32), c3.astype(np.uint32))


def stream_key(seed, stream):
    return (np.uint64(seed & 0xFFFFFFFF),
            np.uint64((0x5EED0000 ^ stream) & 0xFFFFFFFF))


def normals(seed, stream, count, start=0):
    """``count`` standard normals (float64) of the given stream, from element ``start``."""
    if count <= 0:
        return np.zeros(0)
    first_blk = start // 4
    last_blk = (start + count - 1) // 4
    j = np.arange(first_blk, last_blk + 1, dtype=np.uint64)
    k0, k1 = stream_key(seed, stream)
    w = philox4x32_10(j & MASK, j >> np.uint64(32), np.zeros_like(j), np.zeros_like(j), k0, k1)
    u = [(wi.astype(np.float64) + 0.5) * (2.0 ** -32) for wi in w]
    r01 = np.sqrt(-2.0 * np.log(u[0]))
    r23 = np.sqrt(-2.0 * np.log(u[2]))
    t1 = 2.0 * np.pi * u[1]
    t3 = 2.0 * np.pi * u[3]
    z = np.empty((j.size, 4))
    z[:, 0] = r01 * np.cos(t1)
    z[:, 1] = r01 * np.sin(t1)
    z[:, 2] = r23 * np.cos(t3)
    z[:, 3] = r23 * np.sin(t3)
    z = z.reshape(-1)
    off = start - 4 * first_blk
    return z[off:off + count]


def uniforms(seed, stream, count, start=0):
    """``count`` uniforms in (0,1) (float64): word w of block j gives (w + 0.5) * 2**-32."""
    if count <= 0:
        return np.zeros(0)
    first_blk = start // 4
    last_blk = (start + count - 1) // 4
    j = np.arange(first_blk, last_blk + 1, dtype=np.uint64)
    k0, k1 = stream_key(seed, stream)
    w = philox4x32_10(j & MASK, j >> np.uint64(32), np.zeros_like(j), np.zeros_like(j), k0, k1)
    u = np.stack([(wi.astype(np.float64) + 0.5) * (2.0 ** -32) for wi in w], axis=1).reshape(-1)
    off = start - 4 * first_blk
    return u[off:off + count]
