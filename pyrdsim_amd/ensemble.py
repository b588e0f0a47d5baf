"""The batched ensemble: 512 independent replicas in one engine context.

Half of the replicas run the two-species LJ mixture (density 0.6,
dt = 5e-5, periodic), half the wall + reaction model (density 0.4,
dt = 1e-4, periodic in x, reflecting in y) with their own rates k_a, k_d.
Replica sizes are log-uniform over 2,048..32,768 particles; each replica has
its own Philox key.  The reference runs such ensembles as separate processes
in separate directories (examples/patchy-colloids/set_up_simulations.py:160-172);
here they share one context, one set of launches per step.

This module builds the replica plan, the seeded initial states (in worker
processes when asked), and the engine's model and replica tables.
"""

import ctypes
import multiprocessing as mp
import os

import numpy as np

from . import workloads


def replica_spec(p):
    """Box, boundaries, time step and model dict of one planned replica."""
    if p['kind'] == 'kalj':
        return dict(L=workloads.square_box(p['n'], 0.6), periodic=np.array([1.0, 1.0]),
                    dt=5e-5, model=workloads.kalj_model())
    return dict(L=workloads.square_box(p['n'], 0.4), periodic=np.array([1.0, -1.0]),
                dt=1e-4, model=workloads.wall_model(p['k_a'], p['k_d']))


def replica_state(job):
    """Seeded initial state (x, types) of one replica: job = (kind, n, k_a,
    k_d, philox seed, initial-condition seed)."""
    kind, n, k_a, k_d, _, ic_seed = job
    if kind == 'kalj':
        s = workloads.kalj(n=n, seed=ic_seed)
    else:
        s = workloads.wall(n=n, seed=ic_seed, k_a=k_a, k_d=k_d)
    return s['x'], s['types']


def states(plan, workers=1, cache=None, start_method='fork'):
    """Initial states of every planned replica, [(x, types)].  `cache` (an
    .npz path) is read when it exists and matches the plan, and written
    otherwise.  Worker processes use `start_method`; a process that has
    already initialised the GPU should use 'spawn'."""
    if cache and os.path.exists(cache):
        with np.load(cache, allow_pickle=False) as z:
            x, t, n = z['x'], z['t'], z['n']
        if len(n) == len(plan) and all(int(a) == p['n'] for a, p in zip(n, plan)):
            off = np.concatenate([[0], np.cumsum(n)])
            return [(x[off[r]:off[r + 1]], t[off[r]:off[r + 1]]) for r in range(len(plan))]
    jobs = [(p['kind'], p['n'], p['k_a'], p['k_d'], p['seed'], 1005 * 1000 + r)
            for r, p in enumerate(plan)]
    if workers > 1:
        with mp.get_context(start_method).Pool(workers) as pool:
            out = pool.map(replica_state, jobs, chunksize=4)
    else:
        out = [replica_state(j) for j in jobs]
    if cache:
        np.savez(cache, x=np.concatenate([s[0] for s in out]),
                 t=np.concatenate([s[1] for s in out]),
                 n=np.array([p['n'] for p in plan]))
    return out


def engine_inputs(plan):
    """Engine model tables (0 = LJ mixture, 1 = wall + reactions) and one
    replica record per planned replica."""
    from .engine import build_model, make_replica
    mk = workloads.kalj_model()
    mw = workloads.wall_model()
    kalj_model, _ = build_model(mk['interactions'], mk['reactions'], 2)
    wall_model, _ = build_model(mw['interactions'], mw['reactions'], 2)
    reps = []
    for p in plan:
        sp = replica_spec(p)
        if p['kind'] == 'kalj':
            reps.append(make_replica(0, p['n'], sp['L'], sp['periodic'], sp['dt'], [1.0, 1.0],
                                     [], p['seed']))
        else:
            reps.append(make_replica(1, p['n'], sp['L'], sp['periodic'], sp['dt'], [1.0, 1.0],
                                     [p['k_a'], p['k_d']], p['seed']))
    return [kalj_model, wall_model], reps


def with_model0(rep):
    """A copy of a replica record that points at model 0 (for an engine
    built over that replica's model alone)."""
    from . import _lib as L
    r = L.Replica()
    ctypes.pointer(r)[0] = rep
    r.model = 0
    return r


def stratified(plan, k, kinds=('kalj', 'wall'), max_n=None):
    """k replicas spread over the size range, alternating models: for each
    kind, k/len(kinds) replicas at evenly spaced quantiles of its sizes (only
    sizes <= max_n when given)."""
    picks = []
    m = max(1, k // len(kinds))
    for kind in kinds:
        idx = sorted((r for r, p in enumerate(plan)
                      if p['kind'] == kind and (max_n is None or p['n'] <= max_n)),
                     key=lambda r: plan[r]['n'])
        picks += [idx[int((i + 0.5) * len(idx) / m)] for i in range(m)]
    return picks
