"""List mode (skin > 0: Verlet lists rebuilt, with the cell sort, only after a
particle moved more than skin/2) against the oracle's exact neighbour search.

Same bars as the cell mode: forces within rtol of the per-particle force
scale, FP64 injected trajectories within 1e-12, reaction events and neighbour
counts bit-exact.  Between rebuilds the cells are those of the last rebuild;
they are then at least (range + skin) wide, which the reaction and diagnostic
paths rely on, so the trajectory tests run long enough to rebuild many times.
"""

import numpy as np
import pytest

from helpers import engine_for, force_scale, near_cutoff, oracle_for, small
from test_gpu_parity import _injected

pytestmark = pytest.mark.gpu


def _wrap(d, s):
    return d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)


def _activate(name, s):
    if name == 'wall':
        s['types'][::3] = 1
    if name == 'patchy':
        s['types'][2::4] = 2


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall', 'patchy'])
@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-10), ('mixed', 1e-5)])
@pytest.mark.parametrize('skin', [0.3, 1.0])
def test_list_forces_match_oracle(gpu, name, precision, rtol, skin):
    s = small(name)
    _activate(name, s)
    eng = engine_for(s, precision=precision, skin=skin)
    f = eng.forces()
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    err = np.abs(f - a).max(axis=1)
    bad = (err > rtol * scale) & ~near_cutoff(s)
    assert not bad.any(), (name, precision, skin, int(bad.sum()))
    st = eng.sync()
    assert st.rebuilds == 1 and st.list_overflow == 0


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall', 'patchy'])
@pytest.mark.parametrize('skin', [0.05, 0.5])
def test_list_injected_trajectory_fp64(gpu, name, skin):
    """Positions within 1e-12 of the oracle and identical reaction events over
    steps that rebuild many times (skin 0.05) or a few times (skin 0.5)."""
    s = small(name)
    steps = 60
    normals, u1, u2 = _injected(len(s['x']), steps, 78)
    if name == 'wall':
        u1 *= 1e-3
    if name == 'patchy':
        u1 *= 1e-4   # activation p = 5e-7 per draw: make events certain
    eng = engine_for(s, precision='fp64', rng='injected', skin=skin)
    eng.set_injected(normals, u1, u2, first_step=1)
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    for chunk in (1, 9, 20, 30):
        eng.run(chunk)
        st = eng.sync()
        o.run(o.step + 1 if o.step else 1, chunk)
        xg, tg = eng.state()
        xo, to, _ = o.state()
        assert np.abs(_wrap(xg - xo, s)).max() < 1e-12, (name, skin, chunk)
        assert np.array_equal(tg, to), (name, skin, chunk, int((tg != to).sum()))
    if skin == 0.05:
        assert st.rebuilds > 5, st.rebuilds
    if name in ('wall', 'patchy'):
        _, _, ev = o.reaction_info()
        assert st.reaction_events == ev and ev > 0


@pytest.mark.parametrize('name', ['kalj', 'patchy'])
def test_list_overflow_falls_back_to_stencil(gpu, name):
    """With room for 3 partners most particles overflow; they scan the stencil
    and the forces and a trajectory are unchanged."""
    s = small(name)
    _activate(name, s)
    ref = engine_for(s, precision='fp64')
    eng = engine_for(s, precision='fp64', skin=0.4, list_capacity=3)
    f0, f1 = ref.forces(), eng.forces()
    assert np.abs(f1 - f0).max() < 1e-12 * (np.abs(f0).max() + 1.0)
    st = eng.sync()
    assert st.list_overflow > len(s['x']) // 2
    normals, u1, u2 = _injected(len(s['x']), 20, 9)
    outs = []
    for kw in (dict(), dict(skin=0.4, list_capacity=3)):
        e = engine_for(s, precision='fp64', rng='injected', **kw)
        e.set_injected(normals, u1, u2, first_step=1)
        e.run(20)
        e.sync()
        outs.append(e.state())
    assert np.abs(_wrap(outs[0][0] - outs[1][0], s)).max() < 1e-12
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize('rows,per_tile', [(1, 8), (2, 56), (4, 400), (1, 400), (8, 64)])
@pytest.mark.parametrize('name', ['wall', 'patchy'])
def test_list_tile_shapes_and_lds_rounds(gpu, name, rows, per_tile):
    """List mode equals the cell mode for every tile shape, including staged
    regions larger than one LDS round (per_tile = 400), where each lane takes
    its list entries round by round; re-tiling rebuilds the lists."""
    s = small(name)
    _activate(name, s)
    ref = engine_for(s, precision='fp64')
    eng = engine_for(s, precision='fp64', skin=0.5, tile_rows=rows, tile_particles=per_tile)
    f0 = ref.forces()
    assert np.abs(eng.forces() - f0).max() < 1e-12 * (np.abs(f0).max() + 1.0)
    eng.set_tiling(1, 56)
    assert np.abs(eng.forces() - f0).max() < 1e-12 * (np.abs(f0).max() + 1.0)
    steps = 25
    normals, u1, u2 = _injected(len(s['x']), steps, 5)
    u1 *= 1e-3
    outs = []
    for kw in (dict(), dict(skin=0.1, tile_rows=rows, tile_particles=per_tile)):
        e = engine_for(s, precision='fp64', rng='injected', **kw)
        e.set_injected(normals, u1, u2, first_step=1)
        e.run(steps)
        e.sync()
        outs.append(e.state())
    assert np.abs(_wrap(outs[0][0] - outs[1][0], s)).max() < 1e-12
    assert np.array_equal(outs[0][1], outs[1][1])


def test_list_grid_cells_and_counts(gpu):
    """Cells are at least (range + skin) wide (the reference's rule with
    rs = range + skin, ratio 1); cell ids and neighbour counts are exact after
    the particles have moved."""
    from helpers import brute_pairs
    s = small('wall', n=3000, seed=4)
    skin = 0.7
    eng = engine_for(s, skin=skin, seed=8)
    lxy, cl, base = eng.grid()
    for k in range(2):
        L = s['L'][k]
        l = 1
        while L / (l + 1) >= 2.5 + skin:
            l += 1
        assert lxy[0][k] == l
    eng.run(30)
    x, _ = eng.state()
    cells = eng.cell_ids()
    lx, ly = lxy[0]
    cx = np.clip(np.floor(x[:, 0] / cl[0, 0]), 0, lx - 1).astype(int)
    cy = np.clip(np.floor(x[:, 1] / cl[0, 1]), 0, ly - 1).astype(int)
    assert np.array_equal(cells, base[0] + cy * lx + cx)
    for r in (1.5, 2.5):
        got = eng.neighbor_counts(r)
        i, j, d = brute_pairs(x, s['L'], s['periodic'], r + 1e-6)
        want = np.bincount(i[d <= r], minlength=len(x)) + np.bincount(j[d <= r], minlength=len(x))
        near = np.abs(d - r) < 1e-6
        flag = np.zeros(len(x), dtype=bool)
        flag[i[near]] = True
        flag[j[near]] = True
        assert np.array_equal(got[~flag], want[~flag])


def test_list_graph_replay_equals_eager(gpu):
    """Rebuilds requested inside a captured hipGraph run there (conditional
    kernels); the result equals eager stepping."""
    s = small('wall')
    a = engine_for(s, graph_steps=8, seed=3, skin=0.08)
    b = engine_for(s, graph_steps=0, seed=3, skin=0.08)
    for e in (a, b):
        e.run(41)
        e.sync()
    xa, ta = a.state()
    xb, tb = b.state()
    assert np.abs(_wrap(xa - xb, s)).max() < 1e-9
    assert np.array_equal(ta, tb)
    assert a.sync().rebuilds > 3


def test_list_equals_cell_mode_philox(gpu):
    """Same Philox streams, FP64: the list mode and the every-step cell mode
    agree to summation order over a run with rebuilds, batched replicas
    included."""
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    systems = [small('kalj', n=1500, seed=5), small('wall', n=1800, seed=6)]
    models, reps, xs, ts = [], [], [], []
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        models.append(model)
        reps.append(make_replica(k, len(s['x']), s['L'], s['periodic'], s['dt'],
                                 m['tau_by_tauD'], rates, 700 + k))
        xs.append(s['x'])
        ts.append(s['types'])
    out = []
    for skin in (0.0, 0.15):
        eng = Engine(models, reps, precision='fp64', graph_steps=10, skin=skin)
        eng.set_state(np.concatenate(xs), np.concatenate(ts))
        eng.run(60)
        out.append(eng.state())
        if skin:
            assert eng.sync().rebuilds > 2
    lo = 0
    for s in systems:
        n = len(s['x'])
        assert np.abs(_wrap(out[0][0][lo:lo + n] - out[1][0][lo:lo + n], s)).max() < 1e-9
        lo += n
    assert np.array_equal(out[0][1], out[1][1])
