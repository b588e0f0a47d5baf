"""Angle potentials (pyrdsim.pyx:1322-1407, 2052-2065) on the GPU against the
oracle.

Three molecular systems:
* triangles: bead 0 bonded to beads 1 and 2, and an angle (0, 1, 2) at vertex
  1 with theta0 = 60 degrees.  Its leg (1, 2) is no bond, so only the angle
  makes bead 1 a bonded partner of bead 2 for the catalysed reaction
  (pyrdsim.pyx:1405-1406);
* chains whose last angle is straight: bent (120 degree) and straight
  trimers; every angle then takes the series branch when
  theta >= theta0 - 0.1, because the reference's closures read the flag of
  the last angle defined;
* the same chains with a bent angle last: every angle uses the exact formula.
Forces: FP64 within 1e-10, mixed and compensated within 1e-5 of the
per-particle force scale; injected FP64 trajectories within 1e-12 with
identical reaction events, in cell and list mode.
"""

import os

import numpy as np
import pytest

from helpers import engine_for, force_scale, near_cutoff, oracle_for
from pyrdsim_amd import _parse, workloads
from test_gpu_parity import _injected

pytestmark = pytest.mark.gpu

LJ_R0 = 2.0 ** (1.0 / 6.0)
DEG = np.pi / 180.0


def _model():
    wca = lambda a, b: dict(type='lj', pair=[a, b], u0=1.0, r0=LJ_R0, n=6, rc=LJ_R0)
    inter = [wca(0, 0), wca(0, 1), wca(1, 1),
             dict(type='tanh', pair=[1, 1], u0=2.0, r0=1.3, a=0.2, rc=2.0)]
    rx = [dict(type='linear', rate=0.5, reactant=0, product=1),
          dict(type='locally_catalyzed', rate=3.0, r0=1.0, reactant=1, catalyst=1, product=0)]
    return dict(interactions=inter, reactions=rx, tau_by_tauD=[1.0, 1.2])


def _molecules(kind, n_mol, seed, last_pi=True):
    """Positions, types, molecule ids, bonds and angles of n_mol molecules of
    three beads (angles as (i1, i2 vertex, i3, theta0 [rad], k0))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    L = workloads.square_box(n_mol, 0.08)
    centres = workloads.rsa_positions(n_mol, L, 2.6, rng)
    x = np.zeros((3 * n_mol, 2))
    mol = np.zeros((3 * n_mol, 3), dtype=np.intc)
    bi, bj, br, bu = [], [], [], []
    ai, aj, ak, at, a0 = [], [], [], [], []
    k0 = 0.05 / DEG ** 2  # 0.05 kT/deg^2 as written in a bonded_interactions file
    for m in range(n_mol):
        b = 3 * m
        phi = rng.uniform(0, 2 * np.pi)
        if kind == 'triangles':
            # equilateral triangle of side 0.8 around the centre, jittered
            for k in range(3):
                a = phi + 2 * np.pi * k / 3
                x[b + k] = centres[m] + 0.46 * np.array([np.cos(a), np.sin(a)])
            x[b:b + 3] += rng.uniform(-0.05, 0.05, (3, 2))
            for i, j in ((0, 1), (0, 2)):
                bi.append(b + i), bj.append(b + j), br.append(0.8), bu.append(200.0)
            theta0 = 60 * DEG
        else:
            # chain 0-1-2 with bond lengths 0.8 and a bend near theta0
            straight = last_pi if m == n_mol - 1 else m % 2 == 1
            theta0 = np.pi if straight else 120 * DEG
            theta = np.pi - rng.uniform(0, 0.25) if straight else theta0 + rng.uniform(-0.15, 0.15)
            d1 = np.array([np.cos(phi), np.sin(phi)])
            d2 = np.array([np.cos(phi + theta), np.sin(phi + theta)])
            x[b + 1] = centres[m]
            x[b] = centres[m] + 0.8 * d1
            x[b + 2] = centres[m] + 0.8 * d2
            for i, j in ((0, 1), (1, 2)):
                bi.append(b + i), bj.append(b + j), br.append(0.8), bu.append(200.0)
        ai.append(b), aj.append(b + 1), ak.append(b + 2), at.append(theta0), a0.append(k0)
        mol[b:b + 3, 0] = m
        mol[b:b + 3, 1] = np.arange(3)
    x %= L
    types = rng.integers(0, 2, 3 * n_mol).astype(np.intc)
    bonds = tuple(np.asarray(v) for v in (bi, bj, br, bu))
    angles = (np.asarray(ai), np.asarray(aj), np.asarray(ak), np.asarray(at), np.asarray(a0))
    return dict(name=kind, x=x, types=types, molecule_ids=mol, L=L,
                periodic=np.array([1.0, 1.0]), dt=1e-4, model=_model(), bonds=bonds,
                angles=angles)


def _system(name, n_mol=600, seed=51):
    if name == 'triangles':
        return _molecules('triangles', n_mol, seed)
    return _molecules('chains', n_mol, seed, last_pi=(name == 'chains_last_pi'))


NAMES = ['triangles', 'chains_last_pi', 'chains_last_bent']


def test_systems_cover_both_angle_branches():
    """The chains hold straight angles on both sides of theta0 - 0.1, and bent
    angles on both sides of it, so both formulas are exercised."""
    s = _system('chains_last_pi')
    i1, i2, i3, th0, _ = s['angles']
    u = s['x'][i1] - s['x'][i2]
    v = s['x'][i3] - s['x'][i2]
    u -= s['L'] * np.round(u / s['L'])
    v -= s['L'] * np.round(v / s['L'])
    th = np.arccos(np.sum(u * v, 1) / np.linalg.norm(u, axis=1) / np.linalg.norm(v, axis=1))
    series = th >= th0 - 0.1
    for straight in (True, False):
        sel = np.isclose(th0, np.pi) == straight
        assert series[sel].any() and (~series[sel]).any()
    assert np.isclose(th0[-1], np.pi)
    assert not np.isclose(_system('chains_last_bent')['angles'][3][-1], np.pi)


@pytest.mark.parametrize('name', NAMES)
@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-10), ('mixed', 1e-5),
                                            ('compensated', 1e-5)])
@pytest.mark.parametrize('skin', [0.0, 0.4])
def test_angle_forces_match_oracle(gpu, name, precision, rtol, skin):
    if precision == 'compensated' and skin:
        pytest.skip('compensated positions run with the cell search')
    s = _system(name)
    eng = engine_for(s, precision=precision, skin=skin)
    f = eng.forces()
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    err = np.abs(f - a).max(axis=1)
    bad = (err > rtol * scale) & ~near_cutoff(s)
    assert not bad.any(), (name, precision, skin, int(bad.sum()), err.max())


@pytest.mark.parametrize('name', NAMES)
@pytest.mark.parametrize('skin', [0.0, 0.2])
def test_angle_injected_trajectory(gpu, name, skin):
    s = _system(name, seed=52)
    steps = 40
    normals, u1, u2 = _injected(len(s['x']), steps, 53)
    u1 *= 1e-3
    eng = engine_for(s, precision='fp64', rng='injected', skin=skin)
    eng.set_injected(normals, u1, u2, first_step=1)
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    for chunk in (1, 14, 25):
        eng.run(chunk)
        st = eng.sync()
        o.run(o.step + 1 if o.step else 1, chunk)
        xg, tg = eng.state()
        xo, to, _ = o.state()
        d = xg - xo
        d -= s['L'] * np.round(d / s['L'])
        assert np.abs(d).max() < 1e-12, (name, skin, chunk, np.abs(d).max())
        assert np.array_equal(tg, to), (name, skin, chunk, int((tg != to).sum()))
    _, _, ev = o.reaction_info()
    assert st.reaction_events == ev and ev > 50


def test_angle_leg_is_a_bonded_catalyst(gpu):
    """Angles of zero strength change nothing but the bonded partners: with
    them the triangles' beads 1 and 2 catalyse each other, without them they
    do not, and the engine's types follow the oracle's in both cases."""
    s = _system('triangles', seed=54)
    i1, i2, i3, th0, k0 = s['angles']
    weak = (i1, i2, i3, th0, np.zeros_like(k0))
    steps = 30
    normals, u1, u2 = _injected(len(s['x']), steps, 55)
    u1 *= 1e-3
    types = []
    for angles in (weak, None):
        t = dict(s, angles=angles)
        e = engine_for(t, precision='fp64', rng='injected')
        e.set_injected(normals, u1, u2, first_step=1)
        e.run(steps)
        e.sync()
        o = oracle_for(t, rng='injected')
        o.set_injected(normals, u1, u2, first_step=1)
        o.init()
        o.run(1, steps)
        xg, tg = e.state()
        xo, to, _ = o.state()
        assert np.array_equal(tg, to)
        types.append(tg)
    assert (types[0] != types[1]).any()


def test_facade_reads_angles_from_files(gpu, tmp_path):
    """A bonded_interactions.txt with [ angles ] (degrees, kT/deg^2) runs
    through pyrdsim_amd.system and equals the engine built from the arrays."""
    from pyrdsim_amd.system import system
    from test_gpu_facade import _write_inputs
    s = _system('chains_last_pi', n_mol=200, seed=56)
    # all molecules of one molecule_id: the angle of the last molecule decides
    s['angles'][3][:] = np.pi
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        _write_inputs(tmp_path, s, n_steps=30, stride=30, n_out=30)
        with open('bonded_interactions.txt', 'w') as f:
            f.write('[ molecule ]\nmolecule_id = 0\nN_atoms = 3\n[ bonds ]\n'
                    '0 1 0.8 200.0\n1 2 0.8 200.0\n[ angles ]\n0 1 2 180. 0.05\n')
        _parse.write_initial_conditions('initial_conditions.txt', s['x'], s['types'],
                                        s['molecule_ids'])
        out = system({'seed': 9}).simulate(verbose=False)
    finally:
        os.chdir(cwd)
    s['model']['tau_by_tauD'] = [1.0, 1.0]
    eng = engine_for(s, precision='fp64', seed=9)
    eng.run(30)
    x, t = eng.state()
    d = out['x'] - x
    d -= s['L'] * np.round(d / s['L'])
    assert np.abs(d).max() < 1e-9
    assert np.array_equal(out['particle_types'], t)


def test_reference_four_patch_file_forces(gpu, tmp_path):
    """The reference's bonded_interactions/4.txt mixes 90 and 180 degree
    angles with a 180 degree one last, so every angle takes the straight-angle
    flag.  Molecules placed from that file's minimum-energy cross, jittered:
    bond and angle lists built by the facade, forces against the oracle."""
    from pyrdsim_amd import initial_conditions as icg
    from pyrdsim_amd.system import system
    ex = os.path.join(os.path.dirname(__file__), 'golden', 'examples', 'patchy')
    base = dict(parameters_filename=str(tmp_path / 'none.txt'), n_dim=2,
                l=np.array([70., 70.]), dt=1e-5, n_steps=10, n_out=10,
                interactions=[dict(type='lj', pair=[0, 0], u0=1.0, r0=2.5, n=10, rc=5.0),
                              dict(type='tanh', pair=[2, 2], u0=20.0, r0=0.2, a=0.1, rc=1.0)],
                reactions=[], tau_by_tauD=[1.0, 4.5, 4.5],
                bonded_interactions_filename=os.path.join(ex, 'bonded_interactions_4.txt'))
    sim = system(dict(base, seed=3))
    rng = np.random.Generator(np.random.PCG64(61))
    ref = icg.reference_molecule(sim.molecule_interactions[0], 5, 2, rng)
    out = icg.place_molecules({0: 100}, {0: [0, 1, 2, 1, 2]}, {0: ref}, np.array([70., 70.]),
                              [1, 1], 2.6, rng)
    sim._set_molecules(out)
    x = sim.x + rng.uniform(-0.08, 0.08, sim.x.shape)
    bonds, angles = sim._bonds()
    assert len(angles[0]) == 6 * 100 and np.isclose(angles[3][-1], np.pi)
    s = dict(name='patchy4', x=x % 70.0, types=sim.particle_types,
             molecule_ids=sim.molecule_ids, L=np.array([70., 70.]),
             periodic=np.array([1.0, 1.0]), dt=1e-5,
             model=dict(interactions=base['interactions'], reactions=[],
                        tau_by_tauD=[1.0, 4.5, 4.5]),
             bonds=tuple(np.asarray(v) for v in bonds),
             angles=tuple(np.asarray(v) for v in angles))
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    for precision, rtol in (('fp64', 1e-10), ('mixed', 1e-5)):
        f = engine_for(s, precision=precision).forces()
        err = np.abs(f - a).max(axis=1)
        assert not ((err > rtol * scale) & ~near_cutoff(s)).any(), (precision, err.max())
