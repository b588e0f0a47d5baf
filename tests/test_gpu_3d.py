"""The three-dimensional engine (rds_config.n_dim = 3) against the CPU oracle
on the same inputs: the reference loops over N_dim in the update
(pyrdsim.pyx:2244), boundaries (pyrdsim.pyx:2269-2292), distances
(pyrdsim.pyx:1625-1647) and the cell test (pyrdsim.pyx:1540).

Tolerances are those of the two-dimensional tests: forces per particle within
rtol * S_i (S_i the pair force scale of helpers.force_scale plus the
particle's wall/bond drift), rtol = 1e-10 in FP64 and 1e-5 with FP32 pair
arithmetic, pairs within 1e-6 of a cutoff excluded; FP64 positions after
injected steps within 1e-12 absolute; reaction outcomes and cell ids
bit-exact.
"""

import math

import numpy as np
import pytest

from helpers import brute_pairs, engine_for, force_scale, near_cutoff, oracle_for, small

pytestmark = pytest.mark.gpu

NAMES = ['lj3', 'wall3', 'patchy3']


def _prepared(name, seed=None):
    s = small(name) if seed is None else small(name, seed=seed)
    if name == 'wall3':
        s['types'][::3] = 1    # active particles: LJ well + wall terms
    if name == 'patchy3':
        s['types'][2::4] = 2   # some active patches
    return s


def _wrap(d, s):
    return d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)


@pytest.mark.parametrize('name', NAMES)
@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-10), ('mixed', 1e-5)])
def test_forces_3d_match_oracle(gpu, name, precision, rtol):
    s = _prepared(name)
    print('forces 3-D', name, precision, len(s['x']), flush=True)
    eng = engine_for(s, precision=precision)
    f = eng.forces()
    assert f.shape == (len(s['x']), 3)
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    err = np.abs(f - a).max(axis=1)
    bad = (err > rtol * scale) & ~near_cutoff(s)
    assert not bad.any(), (name, precision, int(bad.sum()),
                           float((err / (scale + 1e-300))[bad].max()))
    # the forces are not trivially zero (in wall3 the inactive particles,
    # 1.36 apart on the lattice, are beyond the 1.12 WCA range of each other)
    assert (np.abs(a).sum(axis=1) > 0).mean() > 0.3


def _injected(n, steps, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    return (rng.standard_normal((steps, n, 3)), rng.random((steps, n)), rng.random((steps, n)))


@pytest.mark.parametrize('name', NAMES)
def test_injected_trajectory_3d_fp64(gpu, name):
    """Same injected normals and uniforms: positions within 1e-12, reaction
    outcomes and event counts identical."""
    s = _prepared(name)
    steps = 40
    normals, u1, u2 = _injected(len(s['x']), steps, 91)
    if name == 'wall3':
        u1 *= 1e-3   # frequent activations
    if name == 'patchy3':
        u1 *= 1e-2
    print('injected 3-D', name, len(s['x']), flush=True)
    eng = engine_for(s, precision='fp64', rng='injected')
    eng.set_injected(normals, u1, u2, first_step=1)
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    for chunk in (1, 4, 15, 20):
        eng.run(chunk)
        st = eng.sync()
        o.run(o.step + 1 if o.step else 1, chunk)
        xg, tg = eng.state()
        xo, to, _ = o.state()
        d = _wrap(xg - xo, s)
        assert np.abs(d).max() < 1e-12, (name, chunk, np.abs(d).max())
        assert np.array_equal(tg, to), (name, chunk, int((tg != to).sum()))
    if name in ('wall3', 'patchy3'):
        _, _, ev = o.reaction_info()
        assert st.reaction_events == ev and ev > 0, (st.reaction_events, ev)


def test_philox_noise_3d_matches_oracle(gpu):
    """With no interactions the first step is x1 = x0 + npre * xi on each of
    the three axes: the engine's Philox normals (x, y from stream 0, z from
    stream 2) equal the oracle's (FP32 Box-Muller, a few ulp)."""
    s = small('lj3', n=3000)
    s['model'] = dict(interactions=[], reactions=[], tau_by_tauD=[1.0])
    s['types'][:] = 0
    eng = engine_for(s, seed=777)
    eng.run(1)
    x1, _ = eng.state()
    o = oracle_for(s, seed=777)
    o.init()
    o.run(1, 1)
    xo, _, _ = o.state()
    npre = math.sqrt(2.0 * s['dt'])
    xi_g = _wrap(x1 - s['x'], s) / npre
    xi_o = _wrap(xo - s['x'], s) / npre
    assert np.abs(xi_g - xi_o).max() < 2e-5
    # the three axes are independent standard normals
    assert abs(xi_g.mean()) < 0.05 and abs(xi_g.std() - 1) < 0.05
    c = np.corrcoef(xi_g.T)
    assert np.abs(c - np.eye(3)).max() < 0.08


def test_cells_3d_bit_exact(gpu):
    """Cells per axis follow the reference's rule (the largest l with
    L / l >= the largest pair range, pyrdsim.pyx:1439-1441) and every particle
    sits in the cell floor(x / cell length) on each axis."""
    s = small('wall3')
    eng = engine_for(s)
    lxyz, cl, base = eng.grid()
    assert lxyz.shape == (1, 3)
    for k in range(3):
        l = 1
        while s['L'][k] / (l + 1) >= 2.5:
            l += 1
        assert lxyz[0, k] == l
        assert cl[0, k] == s['L'][k] / l
    cells = eng.cell_ids()
    x = s['x']
    c = [np.minimum((x[:, k] / cl[0, k]).astype(int), lxyz[0, k] - 1) for k in range(3)]
    want = base[0] + (c[2] * lxyz[0, 1] + c[1]) * lxyz[0, 0] + c[0]
    assert np.array_equal(cells, want)


def test_batched_3d_equals_plain(gpu):
    """3-D replicas of different models and sizes stepped together equal the
    same replicas stepped alone (same Philox streams; FP64 to summation
    order), through captured graphs."""
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    systems = [small('lj3', n=1500, seed=5), small('wall3', n=1800, seed=6),
               small('lj3', n=2600, seed=7)]
    models, reps = [], []
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        models.append(model)
        reps.append(make_replica(k, len(s['x']), s['L'], s['periodic'], s['dt'],
                                 m['tau_by_tauD'], rates, 400 + k))
    eng = Engine(models, reps, precision='fp64', graph_steps=8, n_dim=3)
    eng.set_state(np.concatenate([s['x'] for s in systems]),
                  np.concatenate([s['types'] for s in systems]))
    eng.run(50)
    xb, tb = eng.state()
    lo = 0
    for k, s in enumerate(systems):
        e1 = engine_for(s, precision='fp64', seed=400 + k)
        e1.run(50)
        x1, t1 = e1.state()
        n = len(s['x'])
        d = _wrap(xb[lo:lo + n] - x1, s)
        assert np.abs(d).max() < 1e-9, (k, np.abs(d).max())
        assert np.array_equal(tb[lo:lo + n], t1)
        lo += n


def test_graph_replay_3d_equals_eager(gpu):
    s = small('wall3')
    a = engine_for(s, graph_steps=8, seed=3)
    b = engine_for(s, graph_steps=0, seed=3)
    for e in (a, b):
        e.run(37)
        e.sync()
    xa, ta = a.state()
    xb, tb = b.state()
    assert np.abs(_wrap(xa - xb, s)).max() < 1e-9
    assert np.array_equal(ta, tb)


def test_wall3_run_stays_in_the_box(gpu):
    """2000 Philox steps of the reactive wall model in mixed precision: no
    bounds error, positions inside the box, particles activate and deactivate
    (types change in both directions), and the active ones gather at the wall
    (y small)."""
    s = small('wall3')
    eng = engine_for(s, precision='mixed', graph_steps=50)
    eng.run(2000)
    st = eng.sync()
    x, t = eng.state()
    assert st.bounds_error == 0
    assert np.isfinite(x).all() and (x >= 0).all() and (x <= s['L']).all()
    assert st.reaction_events > 0 and 0 < (t == 1).sum() < len(t)


def test_3d_refuses_what_it_does_not_build(gpu):
    """Three dimensions run the cell search in FP64 or mixed precision; the
    list mode and compensated positions are refused with an error, not run."""
    s = small('lj3', n=1000)
    with pytest.raises(RuntimeError, match='three dimensions'):
        engine_for(s, skin=0.3)
    with pytest.raises(RuntimeError, match='three dimensions'):
        engine_for(s, precision='compensated')
    # a 2-D box handed to a 3-D engine is refused by make_replica's caller
    from pyrdsim_amd.engine import make_replica
    with pytest.raises(RuntimeError):
        make_replica(0, 10, [5.0], [1], 1e-3, [1.0], [], 1)


@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_analysis_3d_matches_host(gpu, precision):
    """The device analysis in three dimensions against host brute force on
    the same state (a batch of an LJ-mixture and a wall replica with active
    particles, after 30 steps): the pair histogram (all pairs and type 1
    only) and the RDF from it, neighbour counts within r, type counts and the
    cluster labels of type 1 linked at 1.5.  Integer results equal except
    where a pair lies within 1e-9 of a bin edge or of r (flagged)."""
    from pyrdsim_amd import analysis
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    from test_gpu_analysis import _host_hist, _host_labels
    systems = [small('lj3', n=2500, seed=31), small('wall3', n=2000, seed=32)]
    systems[1]['types'][::2] = 1
    models, reps = [], []
    for k, sy in enumerate(systems):
        m = sy['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        models.append(model)
        reps.append(make_replica(k, len(sy['x']), sy['L'], sy['periodic'], sy['dt'],
                                 m['tau_by_tauD'], rates, 700 + k))
    eng = Engine(models, reps, precision=precision, graph_steps=10, n_dim=3)
    eng.set_state(np.concatenate([sy['x'] for sy in systems]),
                  np.concatenate([sy['types'] for sy in systems]))
    eng.run(30)
    x, t = eng.state()
    r_max, dr = 4.0, 0.05
    nb = int(round(r_max / dr))
    hist = eng.pair_histogram(r_max, dr)
    hist1 = eng.pair_histogram(r_max, dr, types=(1,))
    counts = eng.type_counts()
    labels = eng.cluster_labels(1.5, types=(1,))
    r_probe = 2.0
    nbr = eng.neighbor_counts(r_probe)
    lo = 0
    for k, sy in enumerate(systems):
        n = len(sy['x'])
        xs, ts = x[lo:lo + n], t[lo:lo + n]
        h, flagged = _host_hist(xs, sy['L'], sy['periodic'], r_max, dr, nb)
        assert np.abs(hist[k] - h).sum() <= 2 * flagged, (k, np.abs(hist[k] - h).sum())
        assert hist[k].sum() > 1000
        act = ts == 1
        if act.any():
            h1, f1 = _host_hist(xs[act], sy['L'], sy['periodic'], r_max, dr, nb)
            assert np.abs(hist1[k] - h1).sum() <= 2 * f1
        _, g_host = analysis.rdf(xs, sy['L'], sy['periodic'], r_max, dr)
        _, g_dev = analysis.rdf_from_counts(hist[k], n, sy['L'], r_max, dr, n_dim=3)
        assert analysis.rdf_error(g_dev, g_host) < 1e-3
        assert np.array_equal(counts[k][:3], np.bincount(ts, minlength=3)[:3])
        i, j, d = brute_pairs(xs, sy['L'], sy['periodic'], r_probe + 1e-6)
        want = np.bincount(i[d <= r_probe], minlength=n) + np.bincount(j[d <= r_probe], minlength=n)
        near = np.abs(d - r_probe) < 1e-6
        flag = np.zeros(n, dtype=bool)
        flag[i[near]] = True
        flag[j[near]] = True
        assert np.array_equal(nbr[lo:lo + n][~flag], want[~flag]), k
        want_l, near_l = _host_labels(xs, ts, sy['L'], sy['periodic'], 1.5, [1])
        got = labels[lo:lo + n]
        got = np.where(got >= 0, got - lo, -1)
        if near_l == 0:
            assert np.array_equal(got, want_l), (k, int((got != want_l).sum()))
        lo += n
    eng.close()


def test_dense_cluster_3d(gpu):
    """A dense cubic cluster (8 x 8 x 8 particles 1.05 apart) and a sparse
    background in a large periodic cube: the LDS capacity follows the mean
    occupancy, so the tiles that hold the cluster stage it in several LDS
    rounds.  FP64 and mixed forces against the oracle, then 10 FP64 steps
    with injected noise: positions within 1e-12, no particle lost."""
    from pyrdsim_amd import workloads
    rng = np.random.Generator(np.random.PCG64(5))
    g = np.arange(8) * 1.05
    core = np.stack(np.meshgrid(g, g, g), -1).reshape(-1, 3) + 20.0
    core += rng.uniform(-0.04, 0.04, core.shape)
    L = 60.0
    # background: a coarse lattice (spacing 5) away from the cluster
    b = np.arange(12) * 5.0 + 2.5
    bg = np.stack(np.meshgrid(b, b, b), -1).reshape(-1, 3)
    bg = bg[~np.all((bg > 17.0) & (bg < 31.0), axis=1)]
    x = np.concatenate([core, bg])
    t = (np.arange(len(x)) % 2).astype(np.intc)
    s = dict(name='cluster3', x=x, types=t, molecule_ids=workloads._atoms(x, t),
             L=np.array([L, L, L]), periodic=np.array([1.0, 1.0, 1.0]), dt=5e-5,
             model=workloads.kalj_model(), bonds=None)
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s) + np.abs(a).max(axis=1)
    for precision, rtol in (('fp64', 1e-10), ('mixed', 1e-5)):
        eng = engine_for(s, precision=precision)
        rows, per, cap = eng.tiling()
        assert cap < len(core), cap   # the cluster's tiles stage in rounds
        f = eng.forces()
        eng.close()
        err = np.abs(f - a).max(axis=1)
        bad = (err > rtol * scale) & ~near_cutoff(s)
        assert not bad.any(), (precision, int(bad.sum()))
    steps = 10
    normals, u1, u2 = _injected(len(x), steps, 6)
    eng = engine_for(s, precision='fp64', rng='injected')
    eng.set_injected(normals, u1, u2, first_step=1)
    eng.run(steps)
    xg, tg = eng.state()
    eng.close()
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    o.run(1, steps)
    xo, to, _ = o.state()
    assert np.abs(_wrap(xg - xo, s)).max() < 1e-12
    assert np.array_equal(tg, to)


@pytest.mark.parametrize('name', ['lj3', 'patchy3'])
def test_3d_retiling_keeps_state_and_forces(gpu, name):
    """Re-tiling a 3-D context (tiles of 2 x 2 cell rows in y and z, strips
    of about tile_particles particles in x, from one cell to the whole row)
    changes only the work decomposition: same state, same FP64 forces up to
    summation order."""
    s = small(name, n_mol=400) if name == 'patchy3' else small(name, n=3000)
    eng = engine_for(s, precision='fp64', seed=4)
    eng.run(5)
    x0, t0 = eng.state()
    f0 = eng.forces()
    scale = np.abs(f0).max() + 1.0
    for per in (1, 200, 4000, 0):
        eng.set_tiling(1, per)
        x1, t1 = eng.state()
        assert np.array_equal(x0, x1) and np.array_equal(t0, t1)
        assert np.abs(eng.forces() - f0).max() <= 1e-12 * scale, per
    eng.run(3)
    assert np.isfinite(eng.state()[0]).all()


def test_facade_simulates_3d(gpu, tmp_path):
    """The drop-in facade with N_dim = 3 in parameters.txt: simulate()
    returns [N][3] frames, and a second run with the same seed repeats it."""
    import os
    from test_gpu_facade import _write_inputs
    from pyrdsim_amd.system import system
    s = small('wall3', n=1500)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        _write_inputs(str(tmp_path), s, n_steps=40, stride=5, n_out=20)
        out = system({'seed': 11}).simulate(verbose=False)
        n = len(s['x'])
        assert out['x'].shape == (n, 3)
        assert out['trajectory'].shape == (4, n, 3)
        assert (out['x'] >= 0).all() and (out['x'] <= s['L']).all()
        out2 = system({'seed': 11}).simulate(verbose=False)
        assert np.abs(_wrap(out2['x'] - out['x'], s)).max() < 1e-9
    finally:
        os.chdir(cwd)
