"""CPU restatement of the JAX random streams the reference's sampler draws from — TEST INFRASTRUCTURE ONLY
(see ``oracle/__init__.py``).

The reference pins jax/jaxlib 0.4.23 (uv.lock:492-515), whose default PRNG is threefry2x32 with the
original (non-partitionable) split and bit generation.  Restated here with NumPy uint32 arithmetic:

* ``threefry2x32``  Threefry-2x32 with 20 rounds (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
  easy as 1, 2, 3", SC 2011; the Random123 reference), as jax.random's threefry2x32 primitive;
* ``prng_key``      jax.random.PRNGKey(seed) = [seed >> 32, seed & 0xffffffff];
* ``split``         jax.random.split: the counts 0 .. 2 num - 1 hashed as pairs (i, num + i), outputs
  concatenated and reshaped to [num, 2];
* ``random_bits``   32-bit words for `size` values: counts 0 .. size - 1 (padded with one 0 when odd), pairs
  (i, ceil(size / 2) + i), outputs concatenated, the padding word dropped;
* ``uniform``       mantissa fill (bits >> 9 | 0x3f800000) - 1, scaled to [minval, maxval), max with minval;
* ``normal``        sqrt(2) erfinv(uniform(-1 + 2^-24, 1)) with XLA's single-precision erfinv: M. Giles,
  "Approximating the erfinv function" (GPU Computing Gems, 2011), the two-branch degree-8 polynomial in
  w = -log1p(-x^2), evaluated in float32;
* ``mcmc_draws``    the reference sampler's per-walker stream (mcmc.py:175-180, 345-378): at each step the
  walker key k gives s = split(k)[1] and the next key split(k)[0]; the proposal is normal(s, (n, 3)) and
  the accept threshold uniform(s, ()) - the same key s (the reference's shared-key quirk);
* ``init_walkers``  MCMCState.initialize_around_nuclei, Gaussian branch (mcmc.py:38-91): rng_r, rng_spin,
  rng = split(rng, 3); r0 = R[mapping] + normal(rng_r, [B, n, 3]); walker keys split(rng, B).
"""
from __future__ import annotations

import numpy as np

_ROT = ((13, 15, 26, 6), (17, 29, 16, 24))
U32 = np.uint32


def _rotl(x, r):
    return ((x << U32(r)) | (x >> U32(32 - r))).astype(U32)


def threefry2x32(key, x0, x1):
    """Threefry-2x32-20 of the counter pairs (x0[i], x1[i]) under key = (k0, k1); uint32 arrays."""
    k0, k1 = U32(key[0]), U32(key[1])
    ks = (k0, k1, U32(k0 ^ k1 ^ U32(0x1BD11BDA)))
    with np.errstate(over="ignore"):
        x0 = (np.asarray(x0, U32) + ks[0]).astype(U32)
        x1 = (np.asarray(x1, U32) + ks[1]).astype(U32)
        for blk in range(5):
            for r in _ROT[blk % 2]:
                x0 = (x0 + x1).astype(U32)
                x1 = _rotl(x1, r) ^ x0
            x0 = (x0 + ks[(blk + 1) % 3]).astype(U32)
            x1 = (x1 + ks[(blk + 2) % 3] + U32(blk + 1)).astype(U32)
    return x0, x1


def prng_key(seed: int) -> np.ndarray:
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], U32)


def split(key, num: int = 2) -> np.ndarray:
    i = np.arange(num, dtype=U32)
    y0, y1 = threefry2x32(key, i, i + U32(num))
    return np.concatenate([y0, y1]).reshape(num, 2)


def random_bits(key, size: int) -> np.ndarray:
    half = (size + 1) // 2
    cnt = np.zeros(2 * half, U32)
    cnt[:size] = np.arange(size, dtype=U32)
    y0, y1 = threefry2x32(key, cnt[:half], cnt[half:])
    return np.concatenate([y0, y1])[:size]


def _unit(bits):
    return ((bits >> U32(9)) | U32(0x3F800000)).view(np.float32) - np.float32(1.0)


def uniform(key, shape=(), minval=0.0, maxval=1.0) -> np.ndarray:
    size = int(np.prod(shape))
    f = _unit(random_bits(key, size))
    lo, hi = np.float32(minval), np.float32(maxval)
    return np.maximum(lo, f * (hi - lo) + lo).reshape(shape).astype(np.float32)


_ERFINV_LT5 = np.array([2.81022636e-08, 3.43273939e-07, -3.5233877e-06, -4.39150654e-06, 0.00021858087,
                        -0.00125372503, -0.00417768164, 0.246640727, 1.50140941], np.float32)
_ERFINV_GE5 = np.array([-0.000200214257, 0.000100950558, 0.00134934322, -0.00367342844, 0.00573950773,
                        -0.0076224613, 0.00943887047, 1.00167406, 2.83297682], np.float32)


def erfinv32(x) -> np.ndarray:
    """Single-precision erfinv (Giles 2011), in float32 arithmetic."""
    x = np.asarray(x, np.float32)
    w = -np.log1p(-x * x).astype(np.float32)
    lt = w < np.float32(5.0)
    w = np.where(lt, w - np.float32(2.5), np.sqrt(w).astype(np.float32) - np.float32(3.0)).astype(np.float32)
    p = np.where(lt, _ERFINV_LT5[0], _ERFINV_GE5[0]).astype(np.float32)
    for i in range(1, 9):
        p = (np.where(lt, _ERFINV_LT5[i], _ERFINV_GE5[i]) + p * w).astype(np.float32)
    out = (p * x).astype(np.float32)
    return np.where(np.abs(x) == np.float32(1.0), x * np.float32(np.finfo(np.float32).max), out)


def normal(key, shape) -> np.ndarray:
    lo = np.nextafter(np.float32(-1.0), np.float32(0.0))
    u = uniform(key, shape, lo, 1.0)
    return (np.float32(np.sqrt(2.0)) * erfinv32(u)).astype(np.float32)


def mcmc_draws(keys: np.ndarray, n_el: int, n_steps: int):
    """Proposal normals [S, B, n, 3], accept uniforms [S, B] and the walker keys after S steps."""
    keys = np.asarray(keys, U32).copy()
    B = keys.shape[0]
    normals = np.empty((n_steps, B, n_el, 3), np.float32)
    uniforms = np.empty((n_steps, B), np.float32)
    for s in range(n_steps):
        for b in range(B):
            k_next, sub = split(keys[b])
            normals[s, b] = normal(sub, (n_el, 3))
            uniforms[s, b] = uniform(sub, ())
            keys[b] = k_next
    return normals, uniforms, keys


def init_walkers(rng_key, n_walkers: int, R, el_ion_mapping):
    """Gaussian initialisation around the mapped nuclei and the walker keys (mcmc.py:38-91)."""
    rng_r, _rng_spin, rng = split(rng_key, 3)
    n = len(el_ion_mapping)
    r0 = normal(rng_r, (n_walkers, n, 3)) + np.asarray(R, np.float32)[np.asarray(el_ion_mapping)]
    return r0.astype(np.float32), split(rng, n_walkers)
