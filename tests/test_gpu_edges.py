"""Edge cases of the step loop against the oracle (FP64, injected noise).

* a single particle (no partners: noise and walls only);
* a reflecting box one cell wide (every particle in one cell);
* a periodic box too small for the 3x3 stencil is refused;
* a dense cluster in a large box: cells far above the mean occupancy, so the
  force tiles stage in several LDS rounds, and in list mode the lists
  overflow into the stencil fallback;
* replicas of very different sizes (1, 2, 7 and 3000 particles) in one batch.
"""

import numpy as np
import pytest

from helpers import engine_for, force_scale, oracle_for
from pyrdsim_amd import workloads
from test_gpu_parity import _injected

pytestmark = pytest.mark.gpu


def _system(x, L, periodic, model, dt=1e-4, types=None):
    x = np.asarray(x, dtype=np.float64)
    t = np.zeros(len(x), dtype=np.intc) if types is None else np.asarray(types, dtype=np.intc)
    return dict(name='edge', x=x, types=t, molecule_ids=workloads._atoms(x, t),
                L=np.asarray(L, dtype=np.float64), periodic=np.asarray(periodic, dtype=np.float64),
                dt=dt, model=model, bonds=None)


def _wrap(d, s):
    return d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)


def _compare_injected(s, steps=30, seed=3, **kw):
    normals, u1, u2 = _injected(len(s['x']), steps, seed)
    eng = engine_for(s, precision='fp64', rng='injected', **kw)
    eng.set_injected(normals, u1, u2, first_step=1)
    eng.run(steps)
    eng.sync()
    xg, tg = eng.state()
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    o.run(1, steps)
    xo, to, _ = o.state()
    assert np.abs(_wrap(xg - xo, s)).max() < 1e-12
    assert np.array_equal(tg, to)
    return eng


def test_single_particle(gpu):
    s = _system([[3.0, 4.0]], [10.0, 10.0], [1, 1], workloads.lj1_model())
    _compare_injected(s)


def test_single_particle_at_a_wall(gpu):
    m = workloads.wall_model()
    s = _system([[5.0, 0.7]], [12.0, 12.0], [1, -1], m, types=[1])
    _compare_injected(s)


def test_one_cell_wide_reflecting_box(gpu):
    rng = np.random.Generator(np.random.PCG64(5))
    L = np.array([2.3, 2.2])
    x = workloads.rsa_positions(3, L, 0.95, rng)
    s = _system(x, L, [-1, -1], workloads.lj1_model())
    eng = _compare_injected(s, steps=40)
    lxy, _, _ = eng.grid()
    assert tuple(lxy[0]) == (1, 1)


def test_periodic_box_below_three_cells_is_refused(gpu):
    s = _system([[1.0, 1.0], [3.0, 3.0]], [6.0, 6.0], [1, 1], workloads.lj1_model())
    with pytest.raises(Exception, match='fewer than 3 cells'):
        engine_for(s)


@pytest.mark.parametrize('kw', [dict(), dict(tile_rows=4, tile_particles=144),
                                dict(tile_rows=8, tile_particles=64),
                                dict(skin=0.3, list_capacity=8)])
def test_dense_cluster_in_large_box(gpu, kw):
    """400 particles on a jittered square lattice of spacing 1.05 (density
    0.9) in one corner of a box of mean density 0.02: the cluster's cells hold
    about 40 times the mean occupancy the LDS staging is sized for."""
    rng = np.random.Generator(np.random.PCG64(8))
    g = np.arange(20) * 1.05
    core = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2) + 2.0
    core += rng.uniform(-0.04, 0.04, core.shape)
    L = np.array([140.0, 140.0])
    s = _system(core, L, [1, 1], workloads.lj1_model())
    eng = engine_for(s, precision='fp64', **kw)
    f = eng.forces()
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    assert (np.abs(f - a).max(axis=1) <= 1e-10 * (force_scale(s) + 1.0)).all()
    if 'skin' in kw:
        assert eng.sync().list_overflow > 0
    eng.close()
    _compare_injected(s, steps=20, **kw)


def test_ragged_batch(gpu):
    """Replicas of 1, 2, 7 and 3000 particles batched equal the same replicas
    run one at a time (same Philox streams, FP64)."""
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    rng = np.random.Generator(np.random.PCG64(12))
    systems = []
    for n in (1, 2, 7, 3000):
        L = workloads.square_box(max(n, 50), 0.5)
        systems.append(_system(workloads.rsa_positions(n, L, 0.9, rng), L, [1, 1],
                               workloads.lj1_model()))
    models, reps = [], []
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], 1)
        models.append(model)
        reps.append(make_replica(k, len(s['x']), s['L'], s['periodic'], s['dt'], [1.0], rates,
                                 400 + k))
    eng = Engine(models, reps, precision='fp64', graph_steps=6)
    eng.set_state(np.concatenate([s['x'] for s in systems]),
                  np.concatenate([s['types'] for s in systems]))
    eng.run(24)
    xb, _ = eng.state()
    lo = 0
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], 1)
        e1 = Engine([model], [make_replica(0, len(s['x']), s['L'], s['periodic'], s['dt'], [1.0],
                                           rates, 400 + k)], precision='fp64')
        e1.set_state(s['x'], s['types'])
        e1.run(24)
        x1, _ = e1.state()
        n = len(s['x'])
        assert np.abs(_wrap(xb[lo:lo + n] - x1, s)).max() < 1e-9, k
        lo += n
