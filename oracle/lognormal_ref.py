"""CPU oracle for the GPU's lognormal parameter samples (test infrastructure
only; mcs_lognormal_samples draws them on the GPU).

Row r, variable i: one Philox4x32-10 block per pair of variables (2j, 2j+1)
with key (seed mod 2^32, seed >> 32) and counter (j, r mod 2^32, r >> 32,
0x4C4E524D); u1 = (top 53 bits of (y0, y1) + 1/2) / 2^53, u2 = top 53 bits
of (y2, y3) / 2^53; Box-Muller z_{2j} = sqrt(-2 ln u1) cos(2 pi u2),
z_{2j+1} = sqrt(-2 ln u1) sin(2 pi u2); p = min(1, exp(mu + sigma z)).

Philox4x32-10 here is numpy over many counters at once: the round of
src/mc/prng/philox128.h:8-35 (products by 0xD2511F53 and 0xCD9E8D57, the
Weyl key increments 0x9E3779B9 and 0xBB67AE85), with any key. It is pinned
in tests/test_lognormal.py by the Random123 known-answer vectors, the
reference header's own vectors (tests/golden/philox_ref.json) and the C
oracle's Philox on random counters and keys; the Box-Muller transform by
its closed form at chosen uniforms and by the normal distribution of many
draws. The lognormal parametrisation (ln p = mu + sigma z) is the one of
src/expression/random_deviate.cc:91-108; the reference draws its deviates
from std::mt19937, a stream this engine does not reproduce.
"""
from __future__ import annotations

import numpy as np

M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
W0, W1 = 0x9E3779B9, 0xBB67AE85
TAG = 0x4C4E524D
MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """ctr: 4 arrays (or scalars) of uint32 counter words; key: (k0, k1).
    -> 4 uint32 arrays, the Philox4x32-10 output words."""
    x = [np.asarray(c, dtype=np.uint64) & MASK for c in ctr]
    x = np.broadcast_arrays(*x)
    x0, x1, x2, x3 = (a.copy() for a in x)
    k0, k1 = int(key[0]) & 0xFFFFFFFF, int(key[1]) & 0xFFFFFFFF
    for _ in range(10):
        p0 = x0 * M0  # < 2^64: exact in uint64
        p1 = x2 * M1
        h0, l0 = p0 >> np.uint64(32), p0 & MASK
        h1, l1 = p1 >> np.uint64(32), p1 & MASK
        x0, x1, x2, x3 = h1 ^ x1 ^ np.uint64(k0), l1, h0 ^ x3 ^ np.uint64(k1), l0
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return tuple(a.astype(np.uint32) for a in (x0, x1, x2, x3))


def box_muller(y0, y1, y2, y3):
    """Two standard normals per Philox block: -> (z_even, z_odd) float64."""
    b1 = ((y0.astype(np.uint64) << np.uint64(32)) | y1.astype(np.uint64)) >> np.uint64(11)
    b2 = ((y2.astype(np.uint64) << np.uint64(32)) | y3.astype(np.uint64)) >> np.uint64(11)
    two53 = 9007199254740992.0
    u1 = (b1.astype(np.float64) + 0.5) / two53
    u2 = b2.astype(np.float64) / two53
    rad = np.sqrt(-2.0 * np.log(u1))
    ang = 2.0 * np.pi * u2
    return rad * np.cos(ang), rad * np.sin(ang)


def normals(n_vars: int, seed: int, first_row: int, n_rows: int) -> np.ndarray:
    """[n_rows, n_vars] float64 standard normals of rows first_row ..."""
    pairs = (n_vars + 1) // 2
    r = np.uint64(first_row) + np.arange(n_rows, dtype=np.uint64)[:, None]
    j = np.arange(pairs, dtype=np.uint64)[None, :]
    y = philox4x32_10((j, r & MASK, r >> np.uint64(32), np.uint64(TAG)),
                      (seed & 0xFFFFFFFF, (seed >> 32) & 0xFFFFFFFF))
    za, zb = box_muller(*y)
    z = np.empty((n_rows, 2 * pairs))
    z[:, 0::2], z[:, 1::2] = za, zb
    return z[:, :n_vars]


def lognormal_samples(mu, sigma, seed: int, first_row: int, n_rows: int):
    """-> (z [n_rows, n_vars] float64, p [n_rows, n_vars] float64)."""
    mu = np.asarray(mu, dtype=np.float64)
    sigma = np.asarray(sigma, dtype=np.float64)
    z = normals(mu.shape[0], seed, first_row, n_rows)
    return z, np.minimum(1.0, np.exp(mu[None, :] + sigma[None, :] * z))
