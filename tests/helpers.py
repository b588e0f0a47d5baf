"""Shared helpers of the test suite: build the engine and the CPU oracle on
the same system, and brute-force pair geometry."""

import numpy as np

from pyrdsim_amd import workloads


def small(name, **kw):
    """A small instance of one of the synthetic models (same density/model)."""
    if name == 'lj1':
        return workloads.lj1(**kw)
    if name == 'kalj':
        kw.setdefault('n', 2048)
        return workloads.kalj(**kw)
    if name == 'wall':
        kw.setdefault('n', 2048)
        return workloads.wall(**kw)
    if name == 'patchy':
        kw.setdefault('n_mol', 256)
        return workloads.patchy(**kw)
    if name == 'lj3':
        kw.setdefault('n', 2000)
        return workloads.lj3(**kw)
    if name == 'wall3':
        kw.setdefault('n', 2000)
        return workloads.wall3(**kw)
    if name == 'patchy3':
        kw.setdefault('n_mol', 400)
        return workloads.patchy3(**kw)
    raise KeyError(name)


def rates_of(model):
    return [float(r['rate']) for r in model['reactions']]


def engine_for(s, precision='fp64', rng='philox', seed=2001, graph_steps=0,
               reaction_prob='linear', tile_particles=0, tile_rows=1, skin=0.0,
               list_capacity=0):
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    m = s['model']
    model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
    rep = make_replica(0, len(s['x']), s['L'], s['periodic'], s['dt'], m['tau_by_tauD'],
                       rates, seed)
    has_mol = s['bonds'] is not None
    eng = Engine([model], [rep], precision=precision, rng=rng, reaction_prob=reaction_prob,
                 bonds=s['bonds'], has_molecules=has_mol, graph_steps=graph_steps,
                 tile_particles=tile_particles, tile_rows=tile_rows, skin=skin,
                 list_capacity=list_capacity, angles=s.get('angles'), n_dim=len(s['L']))
    eng.set_state(s['x'], s['types'], s['molecule_ids'][:, 0] if has_mol else None)
    return eng


def oracle_for(s, rng='philox', seed=2001, nl='exact', **kw):
    from oracle import Oracle
    m = s['model']
    bonds = ()
    if s['bonds'] is not None:
        bi, bj, br, bu = s['bonds']
        bonds = list(zip(bi, bj, br, bu))
    angles = ()
    if s.get('angles') is not None:
        angles = list(zip(*s['angles']))
    o = Oracle(s['x'], s['types'], s['L'], s['periodic'], s['dt'], m['tau_by_tauD'],
               m['interactions'], m['reactions'], molecules=s['molecule_ids'][:, 0],
               bonds=bonds, seed=seed, rng=rng, nl=nl, angles=angles, **kw)
    return o


def min_image(d, L, periodic):
    d = d.copy()
    for k in range(len(L)):
        if periodic[k] > 0:
            d[..., k] -= L[k] * np.round(d[..., k] / L[k])
    return d


def brute_pairs(x, L, periodic, rmax):
    """All pairs i<j within rmax: (i, j, r)."""
    n = len(x)
    ii, jj, rr = [], [], []
    for i in range(n - 1):
        d = min_image(x[i + 1:] - x[i], L, periodic)
        r = np.sqrt((d ** 2).sum(axis=1))
        k = np.nonzero(r <= rmax)[0]
        ii.append(np.full(len(k), i))
        jj.append(k + i + 1)
        rr.append(r[k])
    return np.concatenate(ii), np.concatenate(jj), np.concatenate(rr)


def force_scale(s, types=None):
    """Per particle sum over pairs of |f(r)| + |r f'(r)| of every pair term
    within its cutoff (plus k*r0 per bond) -- the magnitude against which
    summed forces are compared.  |r f'(r)| is the sensitivity of a pair force
    to a relative error of the distance, which is what FP32 rounding of the
    displacement introduces; it keeps the scale finite where f crosses zero
    (the LJ minimum, the WCA cutoff)."""
    from oracle import force_law
    x = s['x']
    types = s['types'] if types is None else types
    inter = [i for i in s['model']['interactions'] if i['type'] != 'wall']
    rmax = max(i['rc'] for i in inter)
    i, j, r = brute_pairs(x, s['L'], s['periodic'], rmax)
    mol = s['molecule_ids'][:, 0]
    scale = np.zeros(len(x))
    for it in inter:
        a, b = it['pair']
        ti, tj = types[i], types[j]
        sel = (np.minimum(ti, tj) == a) & (np.maximum(ti, tj) == b) & (r <= it['rc']) \
            & (mol[i] != mol[j])
        for p, q, rv in zip(i[sel], j[sel], r[sel]):
            h = 1e-6 * rv
            fp = (force_law(it, rv + h) - force_law(it, rv - h)) / (2 * h)
            f = abs(force_law(it, rv)) + abs(rv * fp)
            scale[p] += f
            scale[q] += f
    for it in s['model']['interactions']:
        if it['type'] == 'wall':
            # a wall's natural force unit is its largest force, u0 / (2a); far
            # from its centre 1 - tanh^2 is a cancellation whose rounding is
            # of that size times 1e-16, whatever the (vanishing) force itself
            scale[types == int(it['particle_type'])] += float(it['u0']) / (2.0 * float(it['a']))
    if s['bonds'] is not None:
        # a spring's natural force unit is k * r0 (its force at 100% strain);
        # at rest length its force is rounding noise of that size
        bi, bj, br, bu = s['bonds']
        np.add.at(scale, bi, bu * br)
        np.add.at(scale, bj, bu * br)
    if s.get('angles') is not None:
        # an angle's force at one radian from rest over a leg of length 1
        i1, i2, i3, _, k0 = s['angles']
        for p in (i1, i2, i3):
            np.add.at(scale, p, np.asarray(k0))
    return scale


def near_cutoff(s, band=1e-6):
    """Particles with some pair within `band` of a pair cutoff or reaction
    radius: their cutoff decision is rounding-dependent, so they are flagged."""
    radii = [float(i['rc']) for i in s['model']['interactions'] if i['type'] != 'wall']
    radii += [float(r['r0']) for r in s['model']['reactions'] if 'r0' in r]
    i, j, r = brute_pairs(s['x'], s['L'], s['periodic'], max(radii) + 2 * band)
    flag = np.zeros(len(s['x']), dtype=bool)
    for rc in radii:
        near = np.abs(r - rc) < band
        flag[i[near]] = True
        flag[j[near]] = True
    return flag
