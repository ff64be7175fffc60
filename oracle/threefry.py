"""CPU restatement (numpy) of the random stream the reference draws its
auxiliary fields and comb uniforms from.  Test infrastructure only: imported by
tests/, never by the package.

The reference calls, with x64 on and ``jax_threefry_partitionable=False``
(config.py:24-27):

    key = jax.random.PRNGKey(seed)                      prop/afqmc.py:32
    key, subkey = jax.random.split(key)                 prop/afqmc.py:86, prop/blocks.py:81
    fields = jax.random.normal(subkey, (nw, n_fields))  prop/afqmc.py:88
    zeta = jax.random.uniform(subkey)                   prop/blocks.py:82

The algorithm lives in the third-party ``jax`` package (unpinned in the
reference's pyproject.toml; the config comment targets the jax >= 0.5 behaviour
with the partitionable flag off), which is not installed here.  Its published
algorithm, restated:

* Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror, Shaw, "Parallel random
  numbers: as easy as 1, 2, 3", SC 2011): rotations 13 15 26 6 17 29 16 24, key
  schedule (k0, k1, k0 ^ k1 ^ 0x1BD11BDA), key injection after every 4 rounds.
* PRNGKey(seed) = (seed >> 32, seed & 0xffffffff).
* threefry_2x32(key, counts): the counts (padded to even length) are cut in two
  halves x0, x1; pair i is (x0[i], x1[i]); the outputs are concatenated.
* split(key, 2) = threefry_2x32(key, [0, 1, 2, 3]) reshaped (2, 2): the new key
  is row 0, the subkey row 1.
* 64-bit random bits for n values: threefry_2x32(key, iota(2n)) = (A, B) halves,
  value i = (A[i] << 32) | B[i].
* uniform in [lo, hi): f = bitcast((bits >> 12) | bits(1.0)) - 1, then
  max(lo, f * (hi - lo) + lo).
* normal: sqrt(2) * erfinv(uniform(lo = nextafter(-1, 0), hi = 1)).

Pinning (tests/test_rng.py): the Threefry-2x32-20 block function against the
published known-answer vectors; split against the keys jax's documentation
prints for split(PRNGKey(0)) (bit exact); the count layout, odd-count padding,
mantissa recipe, [nextafter(-1, 0), 1) interval and sqrt(2) erfinv transform
through the float32 path (x64 off) against the normals jax's documentation
prints for normal(PRNGKey(0), (10,)) and normal(PRNGKey(42)).  What those
leave **parity unpinned** is only the x64 specialisation: the two 32-bit words
per value (value i from outputs i and n + i) and the 52-bit mantissa shift;
checking them needs jax itself.  erfinv here is scipy.special.erfinv (not XLA's
polynomial), so normals agree with jax's to a few ulp, not bit for bit.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erfinv

_ROT = (13, 15, 26, 6, 17, 29, 16, 24)
_M32 = np.uint64(0xFFFFFFFF)


def _rotl(x, r):
    return ((x << np.uint64(r)) | (x >> np.uint64(32 - r))) & _M32


def threefry2x32(k0, k1, c0, c1):
    """Threefry-2x32-20 on arrays of 32-bit counter pairs (c0, c1) under key
    (k0, k1); returns the two output words as uint32 arrays."""
    k0, k1 = np.uint64(k0), np.uint64(k1)
    ks = (k0, k1, k0 ^ k1 ^ np.uint64(0x1BD11BDA))
    x0 = (np.asarray(c0, np.uint64) + ks[0]) & _M32
    x1 = (np.asarray(c1, np.uint64) + ks[1]) & _M32
    for r in range(20):
        x0 = (x0 + x1) & _M32
        x1 = _rotl(x1, _ROT[r % 8]) ^ x0
        if r % 4 == 3:
            s = (r + 1) // 4
            x0 = (x0 + ks[s % 3]) & _M32
            x1 = (x1 + ks[(s + 1) % 3] + np.uint64(s)) & _M32
    return x0.astype(np.uint32), x1.astype(np.uint32)


def prng_key(seed: int) -> np.ndarray:
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return np.array([seed >> 32, seed & 0xFFFFFFFF], dtype=np.uint32)


def threefry_counts(key, counts) -> np.ndarray:
    """jax's threefry_2x32(key, counts) for a flat uint32 count array."""
    c = np.asarray(counts, np.uint32).ravel()
    odd = c.size % 2
    if odd:
        c = np.concatenate([c, np.zeros(1, np.uint32)])
    h = c.size // 2
    a, b = threefry2x32(key[0], key[1], c[:h], c[h:])
    out = np.concatenate([a, b])
    return out[:-1] if odd else out


def split(key):
    """(new key, subkey) of jax.random.split(key) in the original scheme."""
    out = threefry_counts(key, np.arange(4, dtype=np.uint32)).reshape(2, 2)
    return out[0].copy(), out[1].copy()


def random_bits64(key, n: int) -> np.ndarray:
    out = threefry_counts(key, np.arange(2 * n, dtype=np.uint32))
    hi, lo = out[:n].astype(np.uint64), out[n:].astype(np.uint64)
    return (hi << np.uint64(32)) | lo


def random_bits32(key, n: int) -> np.ndarray:
    """32-bit random bits (x64 off): threefry_2x32(key, iota(n)), value i = output i."""
    return threefry_counts(key, np.arange(n, dtype=np.uint32))


def uniform(key, shape=(), lo=0.0, hi=1.0, dtype=np.float64) -> np.ndarray:
    """jax.random.uniform; float64 (the reference's x64 mode) or float32."""
    n = int(np.prod(shape)) if shape else 1
    if np.dtype(dtype) == np.float32:
        bits = random_bits32(key, n)
        f = ((bits >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
        lo = np.float32(lo)
        hi = np.float32(hi)
    else:
        bits = random_bits64(key, n)
        f = ((bits >> np.uint64(12)) | np.uint64(0x3FF0000000000000)).view(np.float64) - 1.0
        lo = np.float64(lo)
        hi = np.float64(hi)
    u = np.maximum(lo, f * (hi - lo) + lo)
    return u.reshape(shape) if shape else u[0]


def normal(key, shape, dtype=np.float64) -> np.ndarray:
    """jax.random.normal; float64 or float32 (erfinv evaluated in float64)."""
    dt = np.dtype(dtype)
    u = uniform(key, shape, np.nextafter(dt.type(-1.0), dt.type(0.0)), 1.0, dt)
    return (np.sqrt(2.0) * erfinv(np.asarray(u, np.float64))).astype(dt)
