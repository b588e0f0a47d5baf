"""Engine (HIP, through the C ABI) against the CPU oracle on the same inputs.

Tolerances, as stated for the step loop:
* cell assignment, neighbour counts and reaction events: bit-exact, except
  pairs within 1e-6 sigma of a cutoff, which are excluded (flagged);
* forces: |F_gpu - F_oracle| <= rtol * S_i per particle, S_i = sum over the
  particle's pairs of |f(r)| + |r f'(r)| (the force scale the sum is made of,
  including its sensitivity to the distance) plus k*r0 per bond, with
  rtol = 1e-5 for FP32 pair arithmetic and 1e-10 for FP64;
* positions after injected-noise steps: FP64 engine within 1e-12 absolute.
"""

import numpy as np
import pytest

from helpers import brute_pairs, engine_for, force_scale, near_cutoff, oracle_for, small

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall', 'patchy'])
@pytest.mark.parametrize('precision,rtol', [('fp64', 1e-10), ('mixed', 1e-5), ('compensated', 1e-5)])
def test_forces_match_oracle(gpu, name, precision, rtol):
    s = small(name)
    if name == 'wall':
        s['types'][::3] = 1    # some active particles: LJ + wall terms
    if name == 'patchy':
        s['types'][2::4] = 2   # some active patches
    eng = engine_for(s, precision=precision)
    f = eng.forces()
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    scale = force_scale(s)
    err = np.abs(f - a).max(axis=1)
    # walls and bonds add to the drift outside the pair sum: include them
    extra = np.abs(a).max(axis=1)
    bad = (err > rtol * (scale + extra)) & ~near_cutoff(s)
    assert not bad.any(), (name, precision, int(bad.sum()),
                           float((err / (scale + extra + 1e-300))[bad].max()))
    # plain relative errors: norm-wise over the system, and per particle over
    # the particles whose force is at least the median magnitude (away from
    # near-cancellations of their terms)
    far = ~near_cutoff(s)
    normwise = float(np.linalg.norm((f - a)[far]) / np.linalg.norm(a[far]))
    mag = np.linalg.norm(a, axis=1)
    solid = far & (mag >= np.median(mag))
    rel = np.linalg.norm(f - a, axis=1)[solid] / mag[solid]
    print('\nforces %s %s: ||dF||/||F|| = %.2e, max |dF_i|/|F_i| = %.2e over the %d of %d '
          'particles with |F_i| >= median' % (name, precision, normwise, float(rel.max()),
                                               int(solid.sum()), len(a)))
    assert normwise <= rtol, (name, precision, normwise)


@pytest.mark.parametrize('name', ['lj1', 'wall'])
def test_cell_assignment_bit_exact(gpu, name):
    s = small(name)
    eng = engine_for(s)
    cells = eng.cell_ids()
    lxy, cl, base = eng.grid()
    lx, ly = lxy[0]
    x = s['x']
    cx = np.clip(np.floor(x[:, 0] / cl[0, 0]), 0, lx - 1).astype(int)
    cy = np.clip(np.floor(x[:, 1] / cl[0, 1]), 0, ly - 1).astype(int)
    assert np.array_equal(cells, base[0] + cy * lx + cx)
    # the grid is the reference's rule with rs = largest range, ratio 1
    for k in range(2):
        L = s['L'][k]
        rc = 2.5
        l = 1
        while L / (l + 1) >= rc:
            l += 1
        assert lxy[0][k] == l and cl[0][k] == L / l


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall'])
def test_neighbor_counts_bit_exact(gpu, name):
    s = small(name)
    eng = engine_for(s)
    for r in (1.5, 2.5):
        got = eng.neighbor_counts(r)
        i, j, d = brute_pairs(s['x'], s['L'], s['periodic'], r + 1e-6)
        want = np.bincount(i[d <= r], minlength=len(s['x'])) + \
            np.bincount(j[d <= r], minlength=len(s['x']))
        near = np.abs(d - r) < 1e-6
        flag = np.zeros(len(s['x']), dtype=bool)
        flag[i[near]] = True
        flag[j[near]] = True
        assert np.array_equal(got[~flag], want[~flag]), name


def _injected(n, steps, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    normals = rng.standard_normal((steps, n, 2))
    u1 = rng.random((steps, n))
    u2 = rng.random((steps, n))
    return normals, u1, u2


@pytest.mark.parametrize('name', ['lj1', 'kalj', 'wall', 'patchy'])
def test_injected_trajectory_fp64(gpu, name):
    """FP64 engine and oracle consume the same injected normals and uniforms:
    positions agree to 1e-12 and the reaction events are identical."""
    s = small(name)
    steps = 40
    normals, u1, u2 = _injected(len(s['x']), steps, 77)
    if name == 'wall':
        u1 *= 1e-3   # make activations frequent enough to exercise the reactions
    if name == 'patchy':
        u1 *= 1e-2
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
        d = xg - xo
        d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        assert np.abs(d).max() < 1e-12, (name, chunk, np.abs(d).max())
        assert np.array_equal(tg, to), (name, chunk, int((tg != to).sum()))
    if name in ('wall', 'patchy'):
        _, _, ev = o.reaction_info()
        assert st.reaction_events == ev and ev > 0


def test_philox_noise_matches_oracle(gpu):
    """With no interactions the first step is x1 = x0 + npre * xi: the engine's
    Philox normals equal the oracle's restatement (FP32 Box-Muller, few ulp)."""
    from oracle import philox_streams
    s = small('lj1')
    s['model'] = dict(interactions=[], reactions=[], tau_by_tauD=[1.0])
    eng = engine_for(s, seed=12345)
    eng.run(1)
    x1, _ = eng.state()
    d = x1 - s['x']
    d = d - s['L'] * np.round(d / s['L'])
    npre = np.sqrt(2.0 * s['dt'])
    nm, _ = philox_streams(12345, 1, len(s['x']))
    assert np.abs(d / npre - nm).max() < 2e-5


def test_batched_equals_plain(gpu):
    """Replicas stepped together in one launch equal the same replicas stepped
    one at a time (same Philox streams; FP64 up to summation order)."""
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    systems = [small('kalj', n=1500, seed=5), small('wall', n=1800, seed=6),
               small('kalj', n=2600, seed=7)]
    models, reps, xs, ts = [], [], [], []
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        models.append(model)
        reps.append(make_replica(k, len(s['x']), s['L'], s['periodic'], s['dt'],
                                 m['tau_by_tauD'], rates, 900 + k))
        xs.append(s['x'])
        ts.append(s['types'])
    eng = Engine(models, reps, precision='fp64', graph_steps=8)
    eng.set_state(np.concatenate(xs), np.concatenate(ts))
    eng.run(50)
    xb, tb = eng.state()
    lo = 0
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        rep = make_replica(0, len(s['x']), s['L'], s['periodic'], s['dt'],
                           m['tau_by_tauD'], rates, 900 + k)
        e1 = Engine([model], [rep], precision='fp64')
        e1.set_state(s['x'], s['types'])
        e1.run(50)
        x1, t1 = e1.state()
        n = len(s['x'])
        d = xb[lo:lo + n] - x1
        d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        assert np.abs(d).max() < 1e-9, (k, np.abs(d).max())
        assert np.array_equal(tb[lo:lo + n], t1)
        lo += n


def test_autotune_leaves_the_engine_unchanged(gpu):
    """The autotuner times steps on a clone: the tuned engine keeps its state,
    step counter and streams, and its next run equals that of an engine built
    with the chosen tiling that was never tuned."""
    from pyrdsim_amd.autotune import autotune
    s = small('wall')
    s['types'][::3] = 1
    a = engine_for(s, seed=9, graph_steps=8)
    x0, t0 = a.state()
    best, times = autotune(a, candidates=((1, 56), (2, 64), (4, 112)), steps=3)
    assert set(times) == {(1, 56), (2, 64), (4, 112)}
    x1, t1 = a.state()
    assert a.step == 0 and a.sync().steps_done == 0
    assert np.array_equal(x0, x1) and np.array_equal(t0, t1)
    assert a.tiling()[:2] == best
    b = engine_for(s, seed=9, graph_steps=8, tile_rows=best[0], tile_particles=best[1])
    for e in (a, b):
        e.run(31)
    xa, ta = a.state()
    xb, tb = b.state()
    d = xa - xb
    d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-12
    assert np.array_equal(ta, tb)


def test_graph_replay_equals_eager(gpu):
    s = small('wall')
    a = engine_for(s, graph_steps=8, seed=3)
    b = engine_for(s, graph_steps=0, seed=3)
    for e in (a, b):
        e.run(37)
        e.sync()
    xa, ta = a.state()
    xb, tb = b.state()
    d = xa - xb
    d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-9
    assert np.array_equal(ta, tb)


def test_out_of_bounds_raises(gpu):
    """A step that leaves the reflecting box raises RuntimeError like
    pyrdsim.pyx:2286-2292."""
    s = small('wall')
    s['dt'] = 50.0
    eng = engine_for(s)
    eng.run(1)
    with pytest.raises(RuntimeError, match='out of bounds'):
        eng.sync()


@pytest.mark.parametrize('rows,per_tile', [(1, 8), (2, 20), (4, 56), (1, 400), (4, 400),
                                           (8, 24), (8, 64), (8, 400)])
@pytest.mark.parametrize('name', ['kalj', 'wall', 'patchy'])
def test_tile_shapes(gpu, name, rows, per_tile):
    """Every tile shape gives the same forces and the same injected-noise
    trajectory: small strips, multi-row tiles, home rows of more than one wave
    batch and staged regions larger than one LDS round (per_tile = 400)."""
    s = small(name)
    if name == 'wall':
        s['types'][::3] = 1
    if name == 'patchy':
        s['types'][2::4] = 2
    ref = engine_for(s, precision='fp64')
    eng = engine_for(s, precision='fp64', tile_rows=rows, tile_particles=per_tile)
    f0, f1 = ref.forces(), eng.forces()
    scale = np.abs(f0).max() + 1.0
    assert np.abs(f1 - f0).max() < 1e-12 * scale
    steps = 12
    normals, u1, u2 = _injected(len(s['x']), steps, 5)
    u1 *= 1e-3
    outs = []
    for e in (ref, eng):
        e.close()
    for kw in (dict(), dict(tile_rows=rows, tile_particles=per_tile)):
        e = engine_for(s, precision='fp64', rng='injected', **kw)
        e.set_injected(normals, u1, u2, first_step=1)
        e.run(steps)
        e.sync()
        outs.append(e.state())
    d = outs[0][0] - outs[1][0]
    d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-12
    assert np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize('rows,per_tile', [(1, 56), (2, 56), (4, 112), (8, 64)])
@pytest.mark.parametrize('precision', ['fp64', 'mixed'])
def test_no_reads_of_unstaged_lds(gpu, rows, per_tile, precision):
    """Regression: with the LDS of the GPU filled with NaN bits before the
    launch, forces and one step stay finite and equal to the oracle's -- no
    kernel reads LDS it did not stage (a masked extra candidate once did)."""
    s = small('kalj', n=4000, seed=11)
    eng = engine_for(s, precision=precision, tile_rows=rows, tile_particles=per_tile)
    o = oracle_for(s)
    o.init()
    _, _, a = o.state()
    for _ in range(3):
        eng.debug_fill_lds()
        f = eng.forces()
        assert np.isfinite(f).all()
        err = np.abs(f - a).max(axis=1)
        tol = (1e-10 if precision == 'fp64' else 1e-5) * (force_scale(s) + np.abs(a).max(axis=1))
        assert ((err <= tol) | near_cutoff(s)).all()
    eng.debug_fill_lds()
    eng.run(1)
    eng.sync()
    x, _ = eng.state()
    assert np.isfinite(x).all()


def test_catalysis_counts_bonded_partners(gpu):
    """The reference counts catalysts over its neighbour list, which holds the
    bonded partners of the same molecule (pyrdsim.pyx:1513) but no other
    same-molecule pair.  With the catalysis radius above the patch-patch bond
    length (0.87 sigma), bonded active patches count: events equal the
    oracle's, step by step."""
    s = small('patchy')
    s['types'][1::4] = 2
    s['types'][2::4] = 2
    for rx in s['model']['reactions']:
        if rx['type'] == 'locally_catalyzed':
            rx['r0'] = 1.0
            rx['rate'] = 500.0
    steps = 20
    normals, u1, u2 = _injected(len(s['x']), steps, 41)
    eng = engine_for(s, precision='fp64', rng='injected')
    eng.set_injected(normals, u1, u2, first_step=1)
    o = oracle_for(s, rng='injected')
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    eng.run(steps)
    st = eng.sync()
    o.run(1, steps)
    xg, tg = eng.state()
    xo, to, _ = o.state()
    assert np.array_equal(tg, to), int((tg != to).sum())
    _, _, ev = o.reaction_info()
    assert st.reaction_events == ev and ev > 0
