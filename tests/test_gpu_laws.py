"""Every pair law and reaction kind of the reference on the GPU against the
oracle (pyrdsim.pyx:1037-1154 pair laws, 2070-2222 reactions).

A three-species model uses the general term table (not the register LJ path):
* LJ with n = 6 and a tanh well on the same type pair (two terms summed, as
  in the reference's one-species example);
* Yukawa, harmonic and piecewise linear (its cutoff is r0), each added to a
  WCA core on its pair; LJ with odd n = 5;
* linear, locally catalysed and locally inhibited reactions (two inhibitor
  types, rate_decrease).
Forces: FP64 within 1e-10 and mixed within 1e-5 of the per-particle force
scale; injected FP64 trajectories within 1e-12 with identical reaction
events, in cell and list mode.
"""

import numpy as np
import pytest

from helpers import brute_pairs, engine_for, force_scale, near_cutoff, oracle_for
from pyrdsim_amd import workloads
from test_gpu_parity import _injected

pytestmark = pytest.mark.gpu

LJ_R0 = 2.0 ** (1.0 / 6.0)
PW_R0, PW_DR = 1.8, 0.6


def _model():
    # every type pair has the same repulsive core, so that a reaction never
    # leaves two particles overlapping under their new types' laws; the other
    # laws are added as further terms on the pairs
    wca = lambda a, b: dict(type='lj', pair=[a, b], u0=1.0, r0=LJ_R0, n=6, rc=LJ_R0)
    inter = [
        dict(type='lj', pair=[0, 0], u0=1.0, r0=LJ_R0, n=6, rc=2.5),
        dict(type='tanh', pair=[0, 0], u0=1.0, r0=1.5, a=0.3, rc=2.5),
        wca(0, 1),
        dict(type='yukawa', pair=[0, 1], u0=5.0, r0=0.5, rc=2.0),
        wca(0, 2),
        wca(1, 1),
        dict(type='harmonic', pair=[1, 1], u0=0.5, rc=1.5),
        wca(1, 2),
        dict(type='piecewise_linear', pair=[1, 2], u0=0.5, r0=PW_R0, dr=PW_DR, rc=PW_R0),
        dict(type='lj', pair=[2, 2], u0=1.0, r0=LJ_R0, n=5, rc=2.2),
    ]
    rx = [
        dict(type='linear', rate=0.2, reactant=0, product=1),
        dict(type='locally_catalyzed', rate=0.5, r0=1.5, reactant=1, catalyst=0, product=2),
        dict(type='locally_inhibited', rate=2.0, rate_decrease=0.5, r0=1.6, reactant=2,
             inhibitor=1, inhibitor2=0, product=0),
    ]
    return dict(interactions=inter, reactions=rx, tau_by_tauD=[1.0, 0.8, 1.2])


def _system(n=2000, seed=31):
    rng = np.random.Generator(np.random.PCG64(seed))
    L = workloads.square_box(n, 0.5)
    x = workloads.rsa_positions(n, L, 0.9, rng)
    t = rng.integers(0, 3, n).astype(np.intc)
    return dict(name='laws', x=x, types=t, molecule_ids=workloads._atoms(x, t), L=L,
                periodic=np.array([1.0, 1.0]), dt=1e-4, model=_model(), bonds=None)


def _flagged(s):
    """near a cutoff or reaction radius, or near the piecewise-linear step"""
    flag = near_cutoff(s)
    i, j, r = brute_pairs(s['x'], s['L'], s['periodic'], PW_R0)
    near = np.abs(r - (PW_R0 - PW_DR)) < 1e-6
    flag[i[near]] = True
    flag[j[near]] = True
    return flag


@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-10), ('mixed', 1e-5),
                                            ('compensated', 1e-5)])
@pytest.mark.parametrize('skin', [0.0, 0.4])
def test_all_pair_laws_forces(gpu, precision, rtol, skin):
    if precision == 'compensated' and skin:
        pytest.skip('compensated positions run with the cell search')
    s = _system()
    eng = engine_for(s, precision=precision, skin=skin)
    f = eng.forces()
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    err = np.abs(f - a).max(axis=1)
    bad = (err > rtol * scale) & ~_flagged(s)
    assert not bad.any(), (precision, skin, int(bad.sum()))


@pytest.mark.parametrize('skin', [0.0, 0.2])
@pytest.mark.parametrize('u1_scale', [1e-3, 3e-3])
def test_all_reaction_kinds_trajectory(gpu, skin, u1_scale):
    """Every reaction kind against the oracle, event for event.  With u1 up
    to 3e-3 the draws of the catalysed (p = dt * 0.5 * count) and inhibited
    (p = dt * (2 - 0.5 * count)) reactants fall in all three of: below p (the
    particle reacts), between p and the bound from its stencil's particle
    count (it is walked and does not react), above the bound (it is not
    walked: no count could make it react)."""
    s = _system(seed=32)
    steps = 40
    normals, u1, u2 = _injected(len(s['x']), steps, 33)
    u1 *= u1_scale
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
        assert np.abs(d).max() < 1e-12, (skin, chunk, np.abs(d).max())
        assert np.array_equal(tg, to), (skin, chunk, int((tg != to).sum()))
    _, _, ev = o.reaction_info()
    assert st.reaction_events == ev and ev > 50
