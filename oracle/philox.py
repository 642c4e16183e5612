"""CPU twin of the counter-based field stream (Philox4x64-10 + Box-Muller).

TEST INFRASTRUCTURE ONLY.  This module is the test oracle, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it.

The reference draws its auxiliary fields and the SR comb offset from JAX's threefry stream
(trot/prop/afqmc.py:91-93, trot/prop/blocks.py:270-271); that stream needs JAX, which is absent,
so the GPU path draws from Philox4x64-10 instead (Salmon, Moraes, Dror and Shaw, "Parallel random
numbers: as easy as 1, 2, 3", SC'11).  This file restates the published generator in vectorised
NumPy.  It is pinned by numpy.random.Philox, an independent implementation of the same generator
(numpy advances its 256-bit counter by one before each block, so numpy at counter c - 1 gives our
block at counter c), and by the zero-key / zero-counter known answer of the Random123 suite.

Stream layout (shared with the HIP kernels):

* block j of draw c for purpose p = Philox4x64-10(counter words (j, c, p, 0), key words (seed, 0));
* normals (p = 0): Box-Muller on the 53-bit uniforms of (x0, x1) and (x2, x3) of block j give
  values 4j .. 4j+3:  u1 = ((x >> 11) + 1) 2^-53,  u2 = (x >> 11) 2^-53,
  z = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2);
* uniform (p = 1): (x0 >> 11) 2^-53 of block 0.
"""

from __future__ import annotations

import numpy as np

M0 = np.uint64(0xD2E7470EE14C6C93)
M1 = np.uint64(0xCA5A826395121157)
W0 = np.uint64(0x9E3779B97F4A7C15)
W1 = np.uint64(0xBB67AE8584CAA73B)
_M32 = np.uint64(0xFFFFFFFF)
_S32 = np.uint64(32)


def _mulhilo(a: np.ndarray, m: np.uint64):
    """(high, low) 64-bit words of the 128-bit product a * m, from 32-bit halves."""
    a0, a1 = a & _M32, a >> _S32
    m0, m1 = m & _M32, m >> _S32
    p00, p01, p10, p11 = a0 * m0, a0 * m1, a1 * m0, a1 * m1
    mid = (p01 & _M32) + (p10 & _M32) + (p00 >> _S32)
    hi = p11 + (p01 >> _S32) + (p10 >> _S32) + (mid >> _S32)
    return hi, a * m


def philox4x64_10(ctr: np.ndarray, key) -> np.ndarray:
    """Philox4x64-10 of counters ctr (n, 4) uint64 under key (k0, k1); returns (n, 4) uint64."""
    with np.errstate(over="ignore"):
        c = [np.array(ctr[:, i], dtype=np.uint64) for i in range(4)]
        k0, k1 = np.uint64(key[0]), np.uint64(key[1])
        for r in range(10):
            if r:
                k0, k1 = k0 + W0, k1 + W1
            hi0, lo0 = _mulhilo(c[0], M0)
            hi1, lo1 = _mulhilo(c[2], M1)
            c = [hi1 ^ c[1] ^ k0, lo1, hi0 ^ c[3] ^ k1, lo0]
    return np.stack(c, axis=1)


def raw_blocks(seed: int, counter: int, purpose: int, n_blocks: int) -> np.ndarray:
    u = np.uint64
    ctr = np.zeros((n_blocks, 4), dtype=np.uint64)
    ctr[:, 0] = np.arange(n_blocks, dtype=np.uint64)
    ctr[:, 1] = u(counter & (2**64 - 1))
    ctr[:, 2] = u(purpose & (2**64 - 1))
    return philox4x64_10(ctr, (u(seed & (2**64 - 1)), u(0)))


def _box_muller(a: np.ndarray, b: np.ndarray):
    u1 = ((a >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0**-53
    u2 = (b >> np.uint64(11)).astype(np.float64) * 2.0**-53
    r = np.sqrt(-2.0 * np.log(u1))
    th = 6.283185307179586 * u2
    return r * np.cos(th), r * np.sin(th)


def normals(seed: int, counter: int, n: int) -> np.ndarray:
    """n N(0,1) FP64 values of draw ``counter`` (purpose 0)."""
    nb = (n + 3) // 4
    x = raw_blocks(seed, counter, 0, nb)
    z = np.empty((nb, 4))
    z[:, 0], z[:, 1] = _box_muller(x[:, 0], x[:, 1])
    z[:, 2], z[:, 3] = _box_muller(x[:, 2], x[:, 3])
    return z.reshape(-1)[:n]


def uniform(seed: int, counter: int) -> float:
    """The comb offset of draw ``counter`` (purpose 1)."""
    x = raw_blocks(seed, counter, 1, 1)
    return float((x[0, 0] >> np.uint64(11)).astype(np.float64) * 2.0**-53)
