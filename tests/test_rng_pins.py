"""Pins of the counter-based field stream's CPU twin (oracle/philox.py), without a GPU.

* the raw Philox4x64-10 words against numpy.random.Philox, an independent implementation of the
  same published generator, over word-boundary counters and keys;
* the Random123 known answer for the zero counter and zero key;
* Box-Muller by what it must satisfy: each pair's squared radius is -2 ln u1 and its angle 2 pi u2
  for the block's own 53-bit uniforms, and 200 000 values pass a Kolmogorov-Smirnov test against
  N(0,1);
* the comb offset is the first word's top 53 bits.
"""

from __future__ import annotations

import numpy as np
import pytest
import scipy.stats

from oracle import philox as P

U64 = 2**64 - 1


def _numpy_block(seed, counter, purpose, j):
    c = j + (counter << 64) + (purpose << 128)
    return np.random.Philox(key=seed, counter=(c - 1) % 2**256).random_raw(4).astype(np.uint64)


@pytest.mark.parametrize("seed,counter,purpose", [(0, 0, 0), (7, 3, 0), (U64, 2**63 + 5, 1),
                                                  (123456789, U64, 0), (2**32 - 1, 2**32, 2)])
def test_raw_words_match_numpy_philox(seed, counter, purpose):
    ours = P.raw_blocks(seed, counter, purpose, 6)
    for j in range(6):
        assert np.array_equal(ours[j], _numpy_block(seed, counter, purpose, j)), j


def test_random123_zero_known_answer():
    x = P.philox4x64_10(np.zeros((1, 4), dtype=np.uint64), (0, 0))[0]
    assert [int(v) for v in x] == [0x16554D9ECA36314C, 0xDB20FE9D672D0FDC, 0xD7E772CEE186176B, 0x7E68B68AEC7BA23B]


def test_box_muller_pairs_follow_their_uniforms():
    seed, counter, nb = 99, 4, 512
    z = P.normals(seed, counter, 4 * nb).reshape(nb, 4)
    x = P.raw_blocks(seed, counter, 0, nb)
    for a, b, col in ((0, 1, 0), (2, 3, 2)):
        u1 = ((x[:, a] >> np.uint64(11)).astype(np.float64) + 1.0) / 2.0**53
        u2 = (x[:, b] >> np.uint64(11)).astype(np.float64) / 2.0**53
        r2 = z[:, col] ** 2 + z[:, col + 1] ** 2
        assert np.allclose(r2, -2.0 * np.log(u1), rtol=1e-13, atol=1e-13)
        ang = np.mod(np.arctan2(z[:, col + 1], z[:, col]), 2 * np.pi)
        d = np.abs(ang - 2 * np.pi * u2)
        d = np.minimum(d, 2 * np.pi - d)
        assert d.max() < 1e-10


def test_normals_are_standard_normal():
    z = P.normals(2024, 11, 200_000)
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    assert abs(z.mean()) < 0.01 and abs(z.var() - 1.0) < 0.01
    # draws differ by counter, not only by position
    assert not np.allclose(P.normals(2024, 12, 16), z[:16])
    # a ragged length is the prefix of the full block stream
    assert np.array_equal(P.normals(2024, 11, 7), z[:7])


def test_uniform_is_top_bits_of_first_word():
    for seed, ctr in ((0, 0), (5, 17), (U64, 3)):
        w = _numpy_block(seed, ctr, 1, 0)[0]
        assert P.uniform(seed, ctr) == int(w) // 2**11 / 2.0**53
        assert 0.0 <= P.uniform(seed, ctr) < 1.0
