"""Pins of the LFSR88 restatement (oracle) -- CPU only.

LFSR88 is L'Ecuyer's taus88 as written in brainpy/_src/math/tifunc.py:47-180.
The reference holds no golden vectors for it (its test only plots
histograms, math/tests/test_tifunc.py:17-122), so it is pinned by what the
mathematics fixes: each Tausworthe component of taus88 has maximal period
(2^31-1, 2^29-1, 2^28-1) only with the published shift/mask constants, and
by an independent numpy restatement.
"""
import numpy as np
import pytest

from oracle import np_ref, orc


@pytest.mark.parametrize("comp,period", [(2, 2 ** 28 - 1), (1, 2 ** 29 - 1), (0, 2 ** 31 - 1)])
def test_component_periods(comp, period):
    # a wrong shift or mask constant breaks the maximal period
    assert orc.lfsr88_component_period(comp, 987654321, period + 16) == period


@pytest.mark.parametrize("seed", [0, 1, 42, 1234, 2 ** 32 - 20])
def test_stream_matches_numpy(seed):
    np.testing.assert_array_equal(orc.lfsr88_stream(seed, 257), np_ref.stream(seed, 257))


def test_stream_golden():
    g = np.load("tests/golden/lfsr88_streams.npz")
    for seed in (0, 1, 42, 1234):
        np.testing.assert_array_equal(orc.lfsr88_stream(seed, 64), g[f"seed{seed}"])


def test_key_wraps_u32():
    # lfsr88_key adds 1, 7, 15 with u32 wrap: seeds 2^32-1 and -1 coincide
    np.testing.assert_array_equal(orc.lfsr88_stream(2 ** 32 - 1, 8), orc.lfsr88_stream(-1, 8))


def test_randint_range_and_mean():
    clen = 19999
    r0 = orc.lfsr88_randint_stream(7, 0, clen - 1, 200000)
    r1 = orc.lfsr88_randint_stream(7, 1, clen, 200000)
    assert r0.min() >= 0 and r0.max() <= clen - 1
    assert r1.min() >= 1 and r1.max() <= clen
    # same draws, shifted by one (both are bits % clen + low)
    np.testing.assert_array_equal(r1, r0 + 1)
    assert abs(r1.mean() - (clen + 1) / 2) < 0.01 * clen


def test_uniform_is_scaled_bits():
    # lfsr88_uniform(k, 0, 1) == f32(bits) * 2^-32 exactly
    bits = orc.lfsr88_stream(99, 1000)
    u = orc.lfsr88_uniform_stream(99, 0.0, 1.0, 1000)
    np.testing.assert_array_equal(u, bits.astype(np.float32) * np.float32(2.0 ** -32))
    lo, hi = np.float32(0.3), np.float32(0.9)
    u2 = orc.lfsr88_uniform_stream(99, lo, hi, 1000)
    np.testing.assert_array_equal(u2, (u * (hi - lo)) + lo)


def test_normal_moments():
    z = orc.lfsr88_normal_stream(5, 0.5, 0.1, 200000).astype(np.float64)
    assert abs(z.mean() - 0.5) < 2e-3
    assert abs(z.std() - 0.1) < 2e-3
    # numpy twin (libm vs numpy transcendental differences only)
    k = np_ref.key(np.array([5]))
    out = []
    for _ in range(64):
        k, v = np_ref.normal(k, 0.5, 0.1)
        out.append(v[0])
    np.testing.assert_allclose(z[:64], np.array(out, np.float64), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("seed", [5, 42, 2024])
def test_normal_is_box_muller_sine_of_consecutive_uniforms(seed):
    # lfsr88_normal (tifunc.py:89-133): u1 then u2 from the (pinned) uniform
    # stream, u1 floored at 1e-10, z = sqrt(-2 ln u1) * sin(2 pi u2),
    # mu + sigma * z.  Recomputed here in FP64 from the uniform stream; a
    # sine/cosine or u1/u2 swap keeps the moments but fails this.
    mu, sigma, count = 0.5, 0.1, 4000
    u = orc.lfsr88_uniform_stream(seed, 0.0, 1.0, 2 * count).astype(np.float64)
    u1, u2 = np.maximum(u[0::2], 1e-10), u[1::2]
    ref = mu + sigma * np.sqrt(-2.0 * np.log(u1)) * np.sin(2.0 * np.pi * u2)
    z = orc.lfsr88_normal_stream(seed, mu, sigma, count).astype(np.float64)
    assert np.max(np.abs(z - ref)) <= 2e-6 * (abs(mu) + 5 * sigma)
    cos = mu + sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    swap = mu + sigma * np.sqrt(-2.0 * np.log(np.maximum(u2, 1e-10))) * np.sin(2.0 * np.pi * u[0::2])
    assert np.max(np.abs(z - cos)) > 0.1 and np.max(np.abs(z - swap)) > 0.1
