"""Pins of the oracle's three-dimensional path (CPU only): the reference
loops over N_dim in the update (pyrdsim.pyx:2244), the boundaries
(pyrdsim.pyx:2269-2292), distances (pyrdsim.pyx:1625-1647) and the cell test
(pyrdsim.pyx:1540).  Pair drifts and neighbour sets in 3-D are pinned by brute
force in test_oracle_physics.py; here: closed-form single steps on the third
axis, the third normal of the engine's Philox noise, free diffusion, and the
bond and angle forces of a non-planar trimer against the gradient of their
energy.
"""

import math

import numpy as np
import pytest

from oracle import Oracle, philox


def _one(x, L, periodic, tau=(1.0,), rng='injected', seed=0, dt=0.01):
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    return Oracle(x, np.zeros(len(x), dtype=np.int32), L, periodic, dt, list(tau), [],
                  rng=rng, seed=seed)


def test_free_step_third_axis_wrap_and_reflection():
    dt = 0.01
    npre = math.sqrt(2 * dt)
    L = [10.0, 10.0, 10.0]
    cases = [
        ([5.0, 5.0, 5.0], [0.3, -0.2, 0.7], [1, 1, 1],
         [5.0 + npre * 0.3, 5.0 - npre * 0.2, 5.0 + npre * 0.7]),
        ([5.0, 5.0, 9.99], [0.0, 0.0, 1.0], [1, 1, 1], [5.0, 5.0, (9.99 + npre) % 10.0]),
        ([5.0, 5.0, 0.01], [0.0, 0.0, -1.0], [1, 1, 1], [5.0, 5.0, (0.01 - npre) % 10.0]),
        ([5.0, 5.0, 0.01], [0.0, 0.0, -1.0], [1, 1, -1], [5.0, 5.0, -(0.01 - npre)]),
        ([5.0, 5.0, 9.99], [0.0, 0.0, 1.0], [1, 1, -1], [5.0, 5.0, 10 - ((9.99 + npre) - 10)]),
    ]
    for x0, xi, per, want in cases:
        o = _one(x0, L, per)
        o.set_injected(np.array([[xi]]), np.zeros((1, 1)), np.zeros((1, 1)), first_step=1)
        o.init()
        o.run(1, 1)
        x, _, _ = o.state()
        assert np.allclose(x[0], want, rtol=0, atol=1e-15), (x0, per, x[0], want)


def test_third_axis_out_of_bounds():
    o = _one([5.0, 5.0, 5.0], [10.0, 10.0, 10.0], [1, 1, -1])
    o.set_injected(np.array([[[0.0, 0.0, 300.0]]]), np.zeros((1, 1)), np.zeros((1, 1)),
                   first_step=1)
    o.init()
    with pytest.raises(RuntimeError, match='out of bounds'):
        o.run(1, 1)


def _box_muller(w):
    ua = np.float32(((int(w[0]) >> 9) * 2 + 1) * 2.0 ** -24)
    ub = np.float32((int(w[1]) >> 8) * 2.0 ** -24)
    r = math.sqrt(-2.0 * math.log(float(ua)))
    th = 2.0 * math.pi * float(ub)
    return r * math.cos(th), r * math.sin(th)


def test_philox_third_normal_is_stream_two():
    """x and y take the two normals of counter (i, step, 0, 0); z takes the
    first normal of counter (i, step, 0, 2) (stream 1 is the reaction
    uniforms)."""
    seed, n, dt, step = 0xBEEF_1234_5, 40, 1e-3, 1
    rng = np.random.Generator(np.random.PCG64(4))
    x0 = rng.random((n, 3)) * 100.0
    o = Oracle(x0, np.zeros(n, dtype=np.int32), [100.0] * 3, [1, 1, 1], dt, [1.0], [],
               rng='philox', seed=seed)
    o.init()
    o.run(step, 1)
    x, _, _ = o.state()
    xi = (x - x0) / math.sqrt(2 * dt)
    key = [seed & 0xffffffff, seed >> 32]
    for i in range(n):
        nx, ny = _box_muller(philox([i, step, 0, 0], key))
        nz, _ = _box_muller(philox([i, step, 0, 2], key))
        assert np.allclose(xi[i], [nx, ny, nz], rtol=0, atol=1e-5), (i, xi[i], nx, ny, nz)


def test_free_diffusion_msd_3d():
    """<|dx|^2> = 2 d D t with d = 3 and D = tau_by_tauD = 1.5; each axis
    carries a third of it."""
    n, steps, dt = 4000, 200, 1e-3
    rng = np.random.Generator(np.random.PCG64(2))
    x0 = rng.random((n, 3)) * 1000.0
    o = Oracle(x0, np.zeros(n, dtype=np.int32), [1000.0] * 3, [1, 1, 1], dt, [1.5], [],
               rng='philox', seed=9)
    o.init()
    o.run(1, steps)
    x, _, _ = o.state()
    d = x - x0
    d -= 1000.0 * np.round(d / 1000.0)
    per_axis = (d ** 2).mean(axis=0) / (2 * 1.5 * steps * dt)
    assert np.all(np.abs(per_axis - 1) < 0.06), per_axis


def _energy(x, bonds, angles):
    u = 0.0
    for i, j, r0, k in bonds:
        u += 0.5 * k * (np.linalg.norm(x[i] - x[j]) - r0) ** 2
    for i1, i2, i3, theta0, k0 in angles:
        a = x[i1] - x[i2]
        b = x[i3] - x[i2]
        h = a @ b / np.linalg.norm(a) / np.linalg.norm(b)
        u += 0.5 * k0 * (np.arccos(np.clip(h, -1, 1)) - theta0) ** 2
    return u


def test_trimer_bond_and_angle_forces_are_minus_gradient_3d():
    """Bonds (pyx:1303-1320) and an angle (pyx:1322-1407) in a non-planar
    4-particle molecule: the drift is -grad of sum (k/2)(d - r0)^2 +
    (k0/2)(theta - theta0)^2, and it sums to zero."""
    x = np.array([[10.0, 10.0, 10.0], [11.1, 10.2, 9.8], [10.3, 11.2, 10.6],
                  [9.4, 9.9, 11.05]])
    bonds = [(0, 1, 1.0, 30.0), (0, 2, 1.1, 20.0), (0, 3, 0.9, 25.0)]
    angles = [(1, 0, 2, 2 * np.pi / 3, 7.0), (2, 0, 3, 1.7, 4.0)]
    o = Oracle(x, np.zeros(4, dtype=np.int32), [50.0] * 3, [1, 1, 1], 1e-4, [1.0], [],
               molecules=np.zeros(4, dtype=np.int32), bonds=bonds, angles=angles)
    o.init()
    a = o.state()[2]
    g = np.zeros_like(x)
    h = 1e-6
    for p in range(4):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[p, k] += h
            xm[p, k] -= h
            g[p, k] = (_energy(xp, bonds, angles) - _energy(xm, bonds, angles)) / (2 * h)
    assert np.abs(a + g).max() < 1e-6 * (np.abs(g).max() + 1), (a, -g)
    assert np.abs(a.sum(axis=0)).max() < 1e-11
