"""Where the engine's exact neighbour search departs from the reference's
type-gated Verlet list, measured on the patchy-colloid model (CPU only).

The reference lists a pair at a refresh only if the two types are partners
(pyrdsim.pyx:1529-1534).  In the patchy model the inactive patch type 1 has
no partners (examples/patchy-colloids/starting-files/nonbonded_interactions.txt:1-18),
so a patch that activates (1 -> 2) between refreshes stays invisible to the
tanh attraction of other active patches and to catalysis until the next
refresh, neighbor_list_update_frequency (10) steps at most.  The engine finds
every pair within range every step, as the oracle's exact mode does.

The test runs the oracle in both modes in lockstep with the same injected
streams (reference grid: rs = 2 rc, cell_length_ratio 2, refresh every 10
steps), copies the exact state into the faithful one after every step (so
each step's difference starts from identical states), and checks:
* forces and reaction outcomes differ only at patches activated since the
  last refresh, or within the tanh range (forces) or catalysis radius (events)
  of such a patch;
* they do differ: the divergence is real and is counted.
"""

import numpy as np

from helpers import min_image, oracle_for, small

TANH_RC = 0.6      # tanh pair cutoff of the patchy model
CAT_R0 = 0.3       # catalysis radius of the patchy model
NL_FREQ = 10


def _near(x, L, periodic, mol, src, r):
    """Particles within r of any particle in src (other molecules), and src."""
    hit = np.zeros(len(x), dtype=bool)
    hit[src] = True
    for i in src:
        d = min_image(x - x[i], L, periodic)
        close = ((d ** 2).sum(axis=1) <= r * r) & (mol != mol[i])
        hit |= close
    return hit


def run_lockstep(n_mol=192, steps=300, seed=5):
    s = small('patchy', n_mol=n_mol)
    rng = np.random.Generator(np.random.PCG64(seed))
    patch = s['types'] != 0
    s['types'][patch] = np.where(rng.random(patch.sum()) < 0.5, 2, 1)
    n = len(s['x'])
    normals = rng.standard_normal((steps, n, 2))
    # activation (dt k_a = 5e-7) about once per 80 steps per inactive patch;
    # catalysed deactivation at dt k_d = 5e-6 per active neighbour
    u1 = rng.random((steps, n)) * 4e-5
    u2 = rng.random((steps, n))
    fa = oracle_for(s, rng='injected', nl='faithful', ratio=2, nl_freq=NL_FREQ)
    ex = oracle_for(s, rng='injected', nl='exact')
    for o in (fa, ex):
        o.set_injected(normals, u1, u2, first_step=1)
        o.init()
    mol = s['molecule_ids'][:, 0]
    t_refresh = s['types'].copy()
    stats = dict(particle_steps=0, force_diff=0, type_diff=0, steps_with_diff=0,
                 new_actives=0, unexplained_force=0, unexplained_type=0)
    for k in range(1, steps + 1):
        _, t_before, _ = ex.state()
        if k % NL_FREQ == 0:
            t_refresh = t_before.copy()   # this step refreshes the faithful list
        fa.run(k, 1)
        ex.run(k, 1)
        xe, te, ae = ex.state()
        xf, tf, af = fa.state()
        assert np.array_equal(xe, xf)     # same positions: the drift of step k-1 was equal
        fd = (ae != af).any(axis=1)
        td = te != tf
        # patches active now (before or after this step's reactions) that were
        # not active at the last refresh
        new_pre = np.nonzero((t_before == 2) & (t_refresh != 2))[0]
        new_post = np.nonzero(((te == 2) | (t_before == 2)) & (t_refresh != 2))[0]
        stats['new_actives'] += int(((te == 2) & (t_before != 2)).sum())
        stats['particle_steps'] += n
        stats['force_diff'] += int(fd.sum())
        stats['type_diff'] += int(td.sum())
        stats['steps_with_diff'] += int(fd.any() or td.any())
        if fd.any():
            ok = _near(xe, s['L'], s['periodic'], mol, new_post, TANH_RC)
            stats['unexplained_force'] += int((fd & ~ok).sum())
        if td.any():
            ok = _near(xe, s['L'], s['periodic'], mol, new_pre, CAT_R0)
            stats['unexplained_type'] += int((td & ~ok).sum())
        # continue both from the exact state
        fa.set_state(xe, te, ae)
    return stats


def test_type_gating_divergence_is_confined_to_new_actives():
    st = run_lockstep()
    print('\ntype-gating divergence (patchy, 768 beads, 300 steps): %s' % st)
    assert st['unexplained_force'] == 0 and st['unexplained_type'] == 0, st
    # the gating bites: newly active patches next to active ones feel no force
    assert st['force_diff'] > 0
    assert st['new_actives'] > 0
