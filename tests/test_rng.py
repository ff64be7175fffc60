"""The step's random stream: Threefry-2x32-20 pinned by the published
known-answer vectors; the device split / uniform / normal against the CPU
restatement of jax's original scheme (oracle/threefry.py)."""
import numpy as np
import pytest

from oracle import threefry as tf

KAT = [  # (key, counter) -> output, Random123 / Salmon et al. known-answer vectors
    ((0x00000000, 0x00000000), (0x00000000, 0x00000000), (0x6B200159, 0x99BA4EFE)),
    ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF), (0x1CB996FC, 0xBB002BE7)),
    ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3), (0xC4923A9C, 0x483DF7A0)),
]


@pytest.mark.parametrize("key,ctr,want", KAT)
def test_threefry_known_answers(key, ctr, want):
    a, b = tf.threefry2x32(key[0], key[1], np.uint32([ctr[0]]), np.uint32([ctr[1]]))
    assert (int(a[0]), int(b[0])) == want


def test_key_split_and_bit_layout():
    key = tf.prng_key(2)
    assert key.tolist() == [0, 2]
    assert tf.prng_key((7 << 32) | 5).tolist() == [7, 5]
    nk, sk = tf.split(key)
    # split = threefry over counts (0, 2) and (1, 3): new key takes the first
    # outputs, the subkey the second
    a, b = tf.threefry2x32(0, 2, np.uint32([0, 1]), np.uint32([2, 3]))
    assert nk.tolist() == a.tolist() and sk.tolist() == b.tolist()
    # 64-bit values: pair i = (i, n + i), value = (A << 32) | B
    bits = tf.random_bits64(sk, 3)
    a, b = tf.threefry2x32(sk[0], sk[1], np.uint32([0, 1, 2]), np.uint32([3, 4, 5]))
    assert bits.tolist() == [(int(x) << 32) | int(y) for x, y in zip(a, b)]


# Outputs printed in jax's own documentation (original Threefry scheme, x64 off):
# the "Pseudo random numbers" tutorial splits PRNGKey(0) and draws
# normal(PRNGKey(42)); the quickstart draws normal(PRNGKey(0), (10,)).
JAX_SPLIT_KEY0 = ([4146024105, 967050713], [2718843009, 1272950319])
JAX_NORMAL_KEY0_10 = [-0.3721109, 0.26423115, -0.18252768, -0.7368197, -0.44030377,
                      -0.1521442, -0.67135346, -0.5908641, 0.73168886, 0.5673026]
JAX_NORMAL_KEY42 = -0.18471177


def test_split_matches_published_jax_keys():
    nk, sk = tf.split(tf.prng_key(0))
    assert (nk.tolist(), sk.tolist()) == JAX_SPLIT_KEY0


def test_float32_normals_match_published_jax_values():
    # the values are printed to float32 precision and erfinv here is not XLA's
    # polynomial: agreement to a few float32 ulp (2^-24 relative each)
    x = tf.normal(tf.prng_key(0), (10,), np.float32)
    assert x.dtype == np.float32
    np.testing.assert_allclose(x, JAX_NORMAL_KEY0_10, rtol=5e-7, atol=0)
    y = tf.normal(tf.prng_key(42), (), np.float32)
    np.testing.assert_allclose(y, JAX_NORMAL_KEY42, rtol=5e-7, atol=0)
    # a planted layout mistake (pairs (2i, 2i + 1) instead of (i, n/2 + i))
    # moves every value off the published ones
    c = np.arange(10, dtype=np.uint32)
    a, b = tf.threefry2x32(0, 0, c[0::2], c[1::2])
    wrong = np.stack([a, b], 1).ravel()
    f = ((wrong >> np.uint32(9)) | np.uint32(0x3F800000)).view(np.float32) - np.float32(1.0)
    assert not np.allclose(f * 2 - 1, tf.uniform(tf.prng_key(0), (10,), -1.0, 1.0, np.float32),
                           rtol=1e-3)


def test_uniform_and_normal_ranges():
    key = tf.split(tf.prng_key(11))[1]
    u = tf.uniform(key, (200000,))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 4 * (1 / 12 / u.size) ** 0.5
    x = tf.normal(key, (200000,))
    assert np.all(np.isfinite(x))
    assert abs(x.mean()) < 4 / u.size ** 0.5 and abs(x.var() - 1) < 4 * (2 / u.size) ** 0.5
    # bits -> [0, 1): the top 52 bits as mantissa
    bits = np.uint64([0, 0xFFFFFFFFFFFFFFFF, 1 << 63])
    f = ((bits >> np.uint64(12)) | np.uint64(0x3FF0000000000000)).view(np.float64) - 1.0
    assert f.tolist() == [0.0, 1.0 - 2.0 ** -52, 0.5]


@pytest.mark.gpu
def test_device_threefry_matches(gpu_device):
    import torch
    from ad_afqmc_prototype_amd import kernels
    rng = np.random.default_rng(0)
    c0 = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    c1 = rng.integers(0, 2 ** 32, 4096, dtype=np.uint64).astype(np.uint32)
    for (k0, k1), ctr, want in KAT:
        c0[:1], c1[:1] = ctr[0], ctr[1]
        o0, o1 = kernels.threefry2x32(k0, k1, torch.tensor(c0.view(np.int32), device=gpu_device),
                                      torch.tensor(c1.view(np.int32), device=gpu_device))
        o0 = o0.cpu().numpy().view(np.uint32)
        o1 = o1.cpu().numpy().view(np.uint32)
        assert (int(o0[0]), int(o1[0])) == want
        r0, r1 = tf.threefry2x32(k0, k1, c0, c1)
        np.testing.assert_array_equal(o0, r0)
        np.testing.assert_array_equal(o1, r1)


@pytest.mark.gpu
def test_device_stream_matches_oracle(gpu_device):
    """split and uniform bit for bit; normals to a few ulp (the erfinv differs
    from scipy's in the last bits)."""
    import torch
    from ad_afqmc_prototype_amd import kernels
    key = kernels.rng_key(5, gpu_device)
    assert key.cpu().numpy().view(np.uint32).tolist() == tf.prng_key(5).tolist()
    nk, sk = kernels.rng_split(key)
    rnk, rsk = tf.split(tf.prng_key(5))
    assert nk.cpu().numpy().view(np.uint32).tolist() == rnk.tolist()
    assert sk.cpu().numpy().view(np.uint32).tolist() == rsk.tolist()
    u = kernels.rng_uniform(sk).item()
    assert u == tf.uniform(rsk)
    shape = (64, 1001)
    x = kernels.rng_normal(sk, shape).cpu().numpy()
    ref = tf.normal(rsk, shape)
    err = np.abs(x - ref) / np.maximum(np.abs(ref), 1e-300)
    assert err.max() < 4e-15, err.max()
    # the tails: uniforms within 2^-40 of -1 or 1 (erfc branch) from many keys
    worst = 0.0
    for seed in range(8):
        k = tf.split(tf.prng_key(100 + seed))[1]
        ref = tf.normal(k, (1 << 18,))
        dk = torch.tensor(k.view(np.int32), device=gpu_device)
        got = kernels.rng_normal(dk, (1 << 18,)).cpu().numpy()
        worst = max(worst, float((np.abs(got - ref) / np.abs(ref)).max()))
    assert worst < 4e-15, worst
