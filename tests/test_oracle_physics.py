"""Pins of the oracle's force laws, neighbour sets, integrator, boundaries and
reactions (CPU only).

Force laws: f must be dU/dr of the potential each law stands for
(pyrdsim.pyx:1068-1145); checked by central differences of U written here,
plus special values.  Neighbour sets: brute force.  Integrator/boundaries:
closed-form single steps (free motion, periodic wrap, reflection, frozen
species).  Reactions: hand-built neighbourhoods with injected uniforms, and
two-state kinetics.
"""

import math

import numpy as np
import pytest

from helpers import brute_pairs, oracle_for, small
from oracle import Oracle, force_law

R0 = 2.0 ** (1.0 / 6.0)


def U_lj(r, u0, r0, n):
    s = (r0 / r) ** n
    return u0 * (s * s - 2 * s)


LAWS = [
    (dict(type='lj', pair=[0, 0], u0=1.3, r0=1.1, n=6, rc=3.0), lambda r: U_lj(r, 1.3, 1.1, 6)),
    (dict(type='lj', pair=[0, 0], u0=0.2, r0=2.25, n=10, rc=5.0), lambda r: U_lj(r, 0.2, 2.25, 10)),
    (dict(type='lj', pair=[0, 0], u0=0.7, r0=1.0, n=5, rc=3.0), lambda r: U_lj(r, 0.7, 1.0, 5)),
    (dict(type='tanh', pair=[0, 0], u0=3.5, r0=2.15, a=0.05, rc=3.15),
     lambda r: 3.5 / 2 * math.tanh((r - 2.15) / 0.05)),
    (dict(type='yukawa', pair=[0, 0], u0=10.0, r0=0.2, rc=2.0),
     lambda r: 10.0 * math.exp(-r / 0.2) / r),
    (dict(type='harmonic', pair=[0, 0], u0=4.0, rc=3.0), lambda r: 4.0 * r * r / 2),
]


@pytest.mark.parametrize('law,U', LAWS)
def test_force_is_derivative_of_potential(law, U):
    for r in np.linspace(0.6, 3.0, 97):
        h = 1e-6 * r
        fd = (U(r + h) - U(r - h)) / (2 * h)
        f = force_law(law, r)
        assert abs(f - fd) <= 1e-6 * (abs(fd) + abs(f)) + 1e-8, (law['type'], r, f, fd)


def test_special_values():
    lj = dict(type='lj', pair=[0, 0], u0=1.7, r0=1.3, n=6, rc=3)
    assert abs(force_law(lj, 1.3)) < 1e-14
    assert abs(U_lj(1.3, 1.7, 1.3, 6) + 1.7) < 1e-14
    th = dict(type='tanh', pair=[0, 0], u0=5.0, r0=0.3, a=0.05, rc=0.6)
    assert abs(force_law(th, 0.3) - 5.0 / (2 * 0.05)) < 1e-12
    pw = dict(type='piecewise_linear', pair=[0, 0], u0=2.0, r0=1.5, dr=0.5)
    assert force_law(pw, 0.9) == 0.0 and force_law(pw, 1.2) == 2.0 / 0.5
    wall = dict(type='wall', particle_type=1, direction=1, u0=3.0, r0=1.0, a=0.5)
    for y in (0.2, 1.0, 2.7):
        h = 1e-6
        U = lambda v: 3.0 / 2 * math.tanh((v - 1.0) / 0.5)
        assert abs(force_law(wall, y) - (U(y + h) - U(y - h)) / (2 * h)) < 1e-7


def _drift_brute(s):
    """Drift of an atomic system by brute force over all pairs: every pair
    term of the model within its cutoff, force on i is +f(r) (x_j - x_i)/r."""
    x, L, per = s['x'], s['L'], s['periodic']
    types = s['types']
    a = np.zeros_like(x)
    inter = s['model']['interactions']
    rmax = max(i['rc'] for i in inter if i['type'] != 'wall')
    ii, jj, rr = brute_pairs(x, L, per, rmax)
    for it in inter:
        if it['type'] == 'wall':
            sel = types == it['particle_type']
            for k in np.nonzero(sel)[0]:
                a[k, it['direction']] -= force_law(it, x[k, it['direction']])
            continue
        lo, hi = it['pair']
        for i, j, r in zip(ii, jj, rr):
            if (min(types[i], types[j]), max(types[i], types[j])) != (lo, hi) or r > it['rc']:
                continue
            d = x[j] - x[i]
            for k in range(len(L)):
                if per[k] > 0:
                    d[k] -= L[k] * round(d[k] / L[k])
            f = force_law(it, r)
            a[i] += f * d / r
            a[j] -= f * d / r
    return a


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall', 'lj3', 'wall3'])
def test_drift_matches_brute_force(name):
    s = small(name, n=400) if name != 'lj1' else small('lj1', n=400)
    if name in ('wall', 'wall3'):
        s['types'][::3] = 1
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    b = _drift_brute(s)
    assert np.abs(a - b).max() <= 1e-10 * (np.abs(b).max() + 1)


@pytest.mark.parametrize('name', ['lj1', 'wall', 'lj3', 'wall3'])
def test_exact_neighbour_list_is_all_pairs_within_range(name):
    s = small(name, n=300)
    o = oracle_for(s)
    o.init()
    got = {tuple(p) for p in o.nl_pairs()}
    i, j, _ = brute_pairs(s['x'], s['L'], s['periodic'], 2.5)
    assert got == set(zip(i.tolist(), j.tolist()))


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall'])
def test_faithful_and_exact_modes_agree(name):
    """With every type pair a partner and no pair crossing the skin in a few
    steps, the reference's Verlet list and the exact search give the same
    forces and the same trajectory (same summation order, bitwise)."""
    s = small(name, n=500)
    n = len(s['x'])
    rng = np.random.Generator(np.random.PCG64(3))
    steps = 25
    normals = rng.standard_normal((steps, n, 2))
    u1 = rng.random((steps, n)) * 1e-3
    u2 = rng.random((steps, n))
    out = []
    for kw in (dict(nl='exact'), dict(nl='faithful', rs=5.0, ratio=2, nl_freq=10)):
        o = oracle_for(s, rng='injected', **kw)
        o.set_injected(normals, u1, u2, first_step=1)
        o.init()
        o.run(1, steps)
        out.append(o.state())
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def _one_particle(x, L, periodic, tau=(1.0,), types=(0,)):
    x = np.atleast_2d(np.asarray(x, dtype=np.float64))
    return Oracle(x, np.asarray(types), L, periodic, 0.01, list(tau), [], rng='injected')


def test_free_step_wrap_and_reflection():
    dt = 0.01
    npre = math.sqrt(2 * dt)
    cases = [
        # x0, normal, L, periodic, expected
        ([5.0, 5.0], [0.3, -0.2], [10, 10], [1, 1], [5.0 + npre * 0.3, 5.0 - npre * 0.2]),
        ([9.99, 0.01], [1.0, -1.0], [10, 10], [1, 1],
         [(9.99 + npre * 1.0) % 10.0, (0.01 - npre * 1.0) % 10.0]),
        ([0.01, 9.99], [-1.0, 1.0], [10, 10], [-1, -1],
         [-(0.01 - npre * 1.0), 10 - ((9.99 + npre) - 10)]),
    ]
    for x0, xi, L, per, want in cases:
        o = _one_particle(x0, L, per)
        o.set_injected(np.array([[xi]]), np.zeros((1, 1)), np.zeros((1, 1)), first_step=1)
        o.init()
        o.run(1, 1)
        x, _, _ = o.state()
        assert np.allclose(x[0], want, rtol=0, atol=1e-15), (x0, x[0], want)


def test_frozen_species_and_out_of_bounds():
    o = _one_particle([5.0, 5.0], [10, 10], [1, 1], tau=(-1.0,))
    o.set_injected(np.array([[[3.0, 3.0]]]), np.zeros((1, 1)), np.zeros((1, 1)), first_step=1)
    o.init()
    o.run(1, 1)
    assert np.array_equal(o.state()[0][0], [5.0, 5.0])
    o = _one_particle([5.0, 5.0], [10, 10], [-1, -1])
    o.set_injected(np.array([[[200.0, 0.0]]]), np.zeros((1, 1)), np.zeros((1, 1)), first_step=1)
    o.init()
    with pytest.raises(RuntimeError, match='out of bounds'):
        o.run(1, 1)


def test_free_diffusion_msd():
    """<|dx|^2> = 2 d D t for free particles (D = tau_by_tauD = 2 here)."""
    n, steps, dt = 4000, 200, 1e-3
    rng = np.random.Generator(np.random.PCG64(1))
    x0 = rng.random((n, 2)) * 1000.0
    o = Oracle(x0, np.zeros(n, dtype=np.int32), [1000.0, 1000.0], [1, 1], dt, [2.0], [],
               rng='philox', seed=5)
    o.init()
    o.run(1, steps)
    x, _, _ = o.state()
    d = x - x0
    d -= 1000.0 * np.round(d / 1000.0)
    msd = (d ** 2).sum(axis=1).mean()
    assert abs(msd / (2 * 2 * 2.0 * steps * dt) - 1) < 0.05


def test_catalysed_rate_and_decision():
    """Three mutually close reactant/catalysts (total rate 2k each) and a
    non-catalyst: p = 2k dt; injected u1 just below fires, just above does not."""
    k, r0, dt = 0.5, 1.0, 0.01
    x = np.array([[5.0, 5.0], [5.5, 5.0], [5.0, 5.8], [6.2, 5.0]])
    types = np.array([1, 1, 1, 0])
    rx = [dict(type='locally_catalyzed', rate=k, r0=r0, reactant=1, catalyst=1, product=0)]
    for u1, fired in ((2 * k * dt - 1e-12, True), (2 * k * dt + 1e-12, False)):
        o = Oracle(x, types, [20, 20], [1, 1], dt, [-1.0, -1.0], [], rx, rng='injected')
        us = np.full((1, 4), 0.999)
        us[0, 0] = u1
        o.set_injected(np.zeros((1, 4, 2)), us, np.zeros((1, 4)), first_step=1)
        o.init()
        o.run(1, 1)
        rates, total, _ = o.reaction_info()
        # particles 0, 1, 2 are mutually within r0 (0.5, 0.8, 0.94); particle 3
        # (type 0) is neither reactant nor catalyst
        assert total[0] == k + k and total[1] == k + k and total[2] == k + k
        assert total[3] == 0.0
        assert (o.state()[1][0] == 0) == fired


def test_inhibited_rate_clamped():
    k, rd, r0 = 1.0, 0.4, 1.0
    x = np.array([[5.0, 5.0], [5.5, 5.0], [5.0, 5.5], [4.5, 5.0]])
    types = np.array([0, 1, 1, 2])
    rx = [dict(type='locally_inhibited', rate=k, rate_decrease=rd, r0=r0, reactant=0,
               inhibitor=1, inhibitor2=2, product=3)]
    o = Oracle(x, types, [20, 20], [1, 1], 0.01, [-1.0] * 4, [], rx, rng='injected')
    o.set_injected(np.zeros((1, 4, 2)), np.ones((1, 4)), np.zeros((1, 4)), first_step=1)
    o.init()
    o.run(1, 1)
    _, total, _ = o.reaction_info()
    assert total[0] == max(0.0, k - rd - rd - rd)


def test_two_state_kinetics():
    """Linear 0 -> 1 at k_a and 1 -> 0 at k_d: the activated fraction relaxes
    to k_a / (k_a + k_d)."""
    n, ka, kd, dt = 20000, 0.3, 0.7, 0.01
    rng = np.random.Generator(np.random.PCG64(2))
    x0 = rng.random((n, 2)) * 1000
    rx = [dict(type='linear', rate=ka, reactant=0, product=1),
          dict(type='linear', rate=kd, reactant=1, product=0)]
    o = Oracle(x0, np.zeros(n, dtype=np.int32), [1000, 1000], [1, 1], dt, [-1.0, -1.0], [],
               rx, rng='philox', seed=9)
    o.init()
    o.run(1, 1000)
    frac = (o.state()[1] == 1).mean()
    assert abs(frac - ka / (ka + kd)) < 0.02


def _trimer_drift(x, angles, L=50.0):
    """Oracle drift of three particles with only angle terms."""
    from oracle import Oracle
    x = np.asarray(x, dtype=np.float64)
    o = Oracle(x, np.zeros(len(x), dtype=np.int32), [L, L], [1, 1], 1e-4, [1.0], [],
               molecules=np.zeros(len(x), dtype=np.int32), angles=angles)
    o.init()
    return o.state()[2]


def _angle_energy(x, i1, i2, i3, theta0, k0):
    u = x[i1] - x[i2]
    v = x[i3] - x[i2]
    h = u @ v / np.sqrt(u @ u) / np.sqrt(v @ v)
    return 0.5 * k0 * (np.arccos(np.clip(h, -1, 1)) - theta0) ** 2


@pytest.mark.parametrize('theta0,bend', [(2 * np.pi / 3, 0.3), (np.pi, 0.2)])
def test_angle_force_is_minus_gradient(theta0, bend):
    """The angle force (pyx:2052-2065) is -dU/dx of U = (k0/2)(theta - theta0)^2
    (the GROMACS harmonic angle), for an ordinary angle and, on the series
    branch, for a straight one (theta0 = pi)."""
    k0 = 7.0
    phi = theta0 - bend
    x = np.array([[10.0 + np.cos(0.4), 10.0 + np.sin(0.4)], [10.0, 10.0],
                  [10.0 + 1.3 * np.cos(0.4 + phi), 10.0 + 1.3 * np.sin(0.4 + phi)]])
    a = _trimer_drift(x, [(0, 1, 2, theta0, k0)])
    g = np.zeros_like(x)
    h = 1e-6
    for p in range(3):
        for k in range(2):
            xp, xm = x.copy(), x.copy()
            xp[p, k] += h
            xm[p, k] -= h
            g[p, k] = (_angle_energy(xp, 0, 1, 2, theta0, k0) -
                       _angle_energy(xm, 0, 1, 2, theta0, k0)) / (2 * h)
    assert np.abs(a + g).max() < 1e-6 * (np.abs(g).max() + 1), (a, -g)
    assert np.abs(a.sum(axis=0)).max() < 1e-12


def test_angle_pi_flag_binds_late():
    """The angle closures read angle_is_pi when they are called, not when they
    are made (pyx:1334-1338 vs 1378): every angle uses the value set for the
    last angle defined.  A 120-degree angle followed by a straight one is
    therefore evaluated with the series formula of pyx:1385-1386 when
    theta >= theta0 - 0.1."""
    k0, theta0 = 7.0, 2 * np.pi / 3
    phi = theta0 + 0.05
    tri = np.array([[10.0 + np.cos(0.4), 10.0 + np.sin(0.4)], [10.0, 10.0],
                    [10.0 + 1.3 * np.cos(0.4 + phi), 10.0 + 1.3 * np.sin(0.4 + phi)]])
    line = np.array([[20.0, 20.0], [21.0, 20.0], [22.0, 20.02]])
    x = np.concatenate([tri, line])
    straight = (3, 4, 5, np.pi, k0)
    bent = (0, 1, 2, theta0, k0)
    a_last = _trimer_drift(x, [bent, straight])   # flag of the straight angle
    a_first = _trimer_drift(x, [straight, bent])  # flag of the bent angle
    u, v = tri[0] - tri[1], tri[2] - tri[1]
    d1, d2 = np.linalg.norm(u), np.linalg.norm(v)
    u, v = u / d1, v / d2
    hh = u @ v
    t = np.arccos(hh) - theta0
    series = -k0 / (1 - t ** 2 / 6 + t ** 4 / 120)
    exact = k0 * t / np.sqrt(1 - hh ** 2)
    assert np.allclose(a_last[0], series / d1 * (v - hh * u), rtol=1e-12, atol=1e-12)
    assert np.allclose(a_first[0], exact / d1 * (v - hh * u), rtol=1e-12, atol=1e-12)
    assert not np.allclose(series, exact)
