"""Pins of the JAX random-stream restatement (oracle/threefry.py) on the CPU, by published known answers.

* Threefry-2x32-20 test vectors of the Random123 distribution (kat_vectors: zero key and counter, all-ones
  key and counter, the pi-digits key and counter).
* Values JAX's own documentation prints for jax 0.4-era threefry: split(PRNGKey(0)) =
  [[4146024105, 967050713], [2718843009, 1272950319]] ("Pseudo Random Numbers in JAX"),
  normal(PRNGKey(0), (10,)) = [-0.3721109, 0.26423115, ..., 0.5673026] (JAX Quickstart) and
  normal(PRNGKey(42)) = -0.18471177 (the random-numbers tutorial).
* erfinv32 against scipy's float64 erfinv (8 float32 ulp for |x| < 0.99, 1e-5 relative in the tails), and
  the uniform's mantissa fill against its definition on hand-picked words.
"""
import numpy as np
import pytest

from oracle import threefry as tf


@pytest.mark.parametrize("key,ctr,out", [
    ((0x00000000, 0x00000000), (0x00000000, 0x00000000), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
])
def test_threefry_random123_vectors(key, ctr, out):
    y0, y1 = tf.threefry2x32(key, np.array([ctr[0]], np.uint32), np.array([ctr[1]], np.uint32))
    assert (int(y0[0]), int(y1[0])) == out


def test_jax_documented_values():
    np.testing.assert_array_equal(tf.prng_key(0), [0, 0])
    np.testing.assert_array_equal(tf.split(tf.prng_key(0)), [[4146024105, 967050713], [2718843009, 1272950319]])
    np.testing.assert_allclose(tf.normal(tf.prng_key(0), (10,)),
                               [-0.3721109, 0.26423115, -0.18252768, -0.7368197, -0.44030377, -0.1521442,
                                -0.67135346, -0.5908641, 0.73168886, 0.5673026], rtol=0, atol=5e-8)
    assert float(tf.normal(tf.prng_key(42), ())) == pytest.approx(-0.18471177, abs=5e-8)


def test_erfinv32_and_uniform_fill():
    scipy_special = pytest.importorskip("scipy.special")
    x = np.linspace(-0.9999, 0.9999, 20001).astype(np.float32)
    ref = scipy_special.erfinv(x.astype(np.float64))
    got = tf.erfinv32(x).astype(np.float64)
    # the single-precision approximation is good to a few ulp in the central branch and to ~1e-5 relative
    # in the tails, where w = -log1p(-x^2) amplifies the rounding of x^2
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    central = np.abs(x) < 0.99
    assert np.all(np.abs(got - ref)[central] <= 8 * ulp[central])
    assert np.all(np.abs(got - ref) <= 1e-5 * np.abs(ref) + 8 * ulp)
    words = np.array([0, 1 << 9, 0xFFFFFFFF, 0x80000000], np.uint32)
    np.testing.assert_array_equal(tf._unit(words), np.array([0.0, 2.0 ** -23, 1 - 2.0 ** -23, 0.5], np.float32))


def test_mcmc_stream_shares_one_subkey_per_step():
    keys = tf.split(tf.prng_key(7), 3)
    normals, uniforms, out = tf.mcmc_draws(keys, n_el=2, n_steps=2)
    k = keys[1]
    nxt, sub = tf.split(k)
    np.testing.assert_array_equal(normals[0, 1], tf.normal(sub, (2, 3)))
    assert uniforms[0, 1] == tf.uniform(sub, ())
    nxt2, sub2 = tf.split(nxt)
    assert uniforms[1, 1] == tf.uniform(sub2, ())
    np.testing.assert_array_equal(out[1], nxt2)
