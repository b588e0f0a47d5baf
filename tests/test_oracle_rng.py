"""Pins of the oracle's random streams (CPU only).

* std::mt19937: the C++ standard's required value (the 10000th output of a
  default-constructed mt19937 is 4123659995) and NumPy's legacy MT19937
  seeding (init_genrand), which is the same generator, on other seeds;
* the reference's draws (pyrdsim.pyx:71-85: a fresh normal_distribution /
  uniform_real_distribution object per draw) restated from the raw 32-bit
  words: generate_canonical<double, 53> = (w1 + w2 2^32) 2^-64 and the
  Marsaglia polar method returning y * sqrt(-2 ln r2 / r2), the x half being
  discarded with the distribution object;
* Philox4x32-10: Random123's published known-answer vectors;
* the engine's noise definition (Box-Muller on Philox words) restated here.
"""

import json
import math
import os

import numpy as np
import pytest

from oracle import mt_raw, mt_sequence, philox, philox_streams

GOLDEN = os.path.join(os.path.dirname(__file__), 'golden')


def test_mt19937_standard_known_answer():
    assert int(mt_raw(5489, 10000)[-1]) == 4123659995


@pytest.mark.parametrize('seed', [1, 2001, 123456789, 2**31 + 5])
def test_mt19937_matches_numpy_legacy_seeding(seed):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    ref = bg.random_raw(2000).astype(np.uint32)
    assert np.array_equal(mt_raw(seed, 2000), ref)


def _canonical(words, k):
    w1, w2 = float(words[k]), float(words[k + 1])
    v = (w1 + w2 * 4294967296.0) / 18446744073709551616.0
    if v >= 1.0:
        v = math.nextafter(1.0, 0.0)
    return v, k + 2


def _restated_draws(seed, kinds):
    words = mt_raw(seed, 20 * len(kinds) + 100)
    k = 0
    out = []
    for kind in kinds:
        if kind == 1:
            u, k = _canonical(words, k)
            out.append(u)
            continue
        while True:
            a, k = _canonical(words, k)
            b, k = _canonical(words, k)
            x = 2.0 * a - 1.0
            y = 2.0 * b - 1.0
            r2 = x * x + y * y
            if not (r2 > 1.0 or r2 == 0.0):
                break
        out.append(y * math.sqrt(-2 * math.log(r2) / r2))
    return np.array(out)


@pytest.mark.parametrize('seed', [7, 2001])
def test_reference_draw_sequence(seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = (rng.random(600) < 0.3).astype(np.int8)   # 0 normal, 1 uniform
    got = mt_sequence(seed, kinds)
    want = _restated_draws(seed, kinds)
    assert np.array_equal(got, want)


def test_reference_normals_are_standard():
    got = mt_sequence(11, np.zeros(200000, dtype=np.int8))
    assert abs(got.mean()) < 0.01
    assert abs(got.std() - 1.0) < 0.01


def test_philox_known_answers():
    with open(os.path.join(GOLDEN, 'philox_kat.json')) as f:
        kat = json.load(f)['philox4x32_10']
    for v in kat:
        ctr = [int(c, 16) for c in v['ctr']]
        key = [int(c, 16) for c in v['key']]
        out = [int(c, 16) for c in v['out']]
        assert [int(w) for w in philox(ctr, key)] == out


def test_engine_noise_definition():
    """Normals: counter (i, step, 0, 0) -> Box-Muller of an odd 24-bit uniform
    and a 24-bit angle; uniforms: counter (i, step, 0, 1) -> two 53-bit values."""
    seed, step, n = 0x1234_5678_9ABC, 77, 50
    nm, u = philox_streams(seed, step, n)
    key = [seed & 0xffffffff, seed >> 32]
    for i in range(n):
        w = [int(x) for x in philox([i, step, 0, 0], key)]
        ua = np.float32(((w[0] >> 9) * 2 + 1) * 2.0 ** -24)
        ub = np.float32((w[1] >> 8) * 2.0 ** -24)
        r = math.sqrt(-2.0 * math.log(float(ua)))
        th = 2.0 * math.pi * float(ub)
        assert abs(nm[i, 0] - r * math.cos(th)) < 1e-5
        assert abs(nm[i, 1] - r * math.sin(th)) < 1e-5
        w = [int(x) for x in philox([i, step, 0, 1], key)]
        assert u[i, 0] == ((w[1] << 32 | w[0]) >> 11) * 2.0 ** -53
        assert u[i, 1] == ((w[3] << 32 | w[2]) >> 11) * 2.0 ** -53


def test_engine_noise_statistics():
    nm, u = philox_streams(99, 3, 200000)
    assert abs(nm.mean()) < 0.01 and abs(nm.std() - 1) < 0.01
    assert abs(np.corrcoef(nm[:, 0], nm[:, 1])[0, 1]) < 0.01
    assert abs(u.mean() - 0.5) < 0.005 and u.min() >= 0 and u.max() < 1
