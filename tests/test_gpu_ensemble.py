"""The batched ensemble as a tested configuration, and the 1 - exp(-k dt)
reaction probability against the oracle.

Ensemble: 512 replicas, half the LJ mixture and half the wall + reaction
model, N log-uniform over 2,048..32,768 (sum N = 5,789,872), per-replica rates
and Philox keys, in one engine context:
* mixed precision (FP32 pair forces, FP64 positions), hipGraph replay and the
  autotuned tiling, 250 steps: every replica finite and inside its box, types
  within the model, device type counts equal to the host's;
* pair forces of the whole batched context against the oracle on a stratified
  subset of replicas of both models (N <= 3,000): per particle within 1e-5 of
  the force scale, and norm-wise ||F - F_oracle|| / ||F_oracle|| <= 1e-5;
* FP64 batched steps with injected noise and uniforms: the subset's positions
  within 1e-12 of the oracle's and its reaction events identical.

Reaction probability (pyrdsim.pyx:2214): the reference fires with p = dt*k;
reaction_probability='exp' uses p = 1 - exp(-dt*k).  Engine and oracle agree
event for event in both forms, and uniforms planted between the two
probabilities fire in the linear form only.
"""

import numpy as np
import pytest

from helpers import engine_for, force_scale, near_cutoff, oracle_for, small
from pyrdsim_amd import ensemble, workloads

pytestmark = pytest.mark.gpu

SUBSET_MAX_N = 3000


@pytest.fixture(scope='module')
def ens():
    plan = workloads.ensemble_plan(n_rep=512)
    st = ensemble.states(plan, workers=8, start_method='spawn')
    models, reps = ensemble.engine_inputs(plan)
    n = np.array([p['n'] for p in plan])
    off = np.concatenate([[0], np.cumsum(n)])
    x0 = np.concatenate([s[0] for s in st])
    t0 = np.concatenate([s[1] for s in st]).astype(np.int32)
    subset = ensemble.stratified(plan, 6, max_n=SUBSET_MAX_N)
    return dict(plan=plan, states=st, models=models, reps=reps, off=off, x0=x0, t0=t0,
                subset=subset)


def _system(e, r, x, t):
    """The oracle-side system dict of replica r with state (x, t)."""
    sp = ensemble.replica_spec(e['plan'][r])
    n = len(x)
    mol = np.zeros((n, 3), dtype=np.intc)
    mol[:, 0] = np.arange(n)
    return dict(x=x, types=t, L=sp['L'], periodic=sp['periodic'], dt=sp['dt'],
                model=sp['model'], bonds=None, molecule_ids=mol)


def _with_actives(e, frac, seed):
    """The initial types with a fraction of every wall replica's particles
    active (type 1), so that forces and reactions of both states occur."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = e['t0'].copy()
    for r, p in enumerate(e['plan']):
        if p['kind'] == 'wall':
            lo, hi = e['off'][r], e['off'][r + 1]
            t[lo:hi] = (rng.random(hi - lo) < frac).astype(np.int32)
    return t


def test_full_ensemble_mixed_graphs_autotuned(gpu, ens):
    from pyrdsim_amd.autotune import autotune
    from pyrdsim_amd.engine import Engine
    e = ens
    eng = Engine(e['models'], e['reps'], precision='mixed', graph_steps=50)
    eng.set_state(e['x0'], e['t0'])
    best, times = autotune(eng)
    assert len(times) > 1 and best in times
    assert eng.step == 0
    eng.run(250)
    st = eng.sync()
    assert st.steps_done == 250 and st.integrity == 0
    x, t = eng.state()
    counts = eng.type_counts()
    eng.close()
    assert np.isfinite(x).all()
    for r, p in enumerate(e['plan']):
        lo, hi = e['off'][r], e['off'][r + 1]
        L = e['reps'][r].L
        xr, tr = x[lo:hi], t[lo:hi]
        assert (xr >= 0).all() and (xr[:, 0] <= L[0]).all() and (xr[:, 1] <= L[1]).all(), r
        if p['kind'] == 'kalj':
            assert np.array_equal(tr, e['t0'][lo:hi]), r
        else:
            assert set(np.unique(tr)) <= {0, 1}, r
        want = np.bincount(tr, minlength=counts.shape[1])
        assert np.array_equal(counts[r], want), r
    # wall replicas activate at rates k_a in [0.01, 1]: some have by now
    assert st.reaction_events > 0


@pytest.mark.parametrize('precision', ['mixed', 'fp64'])
def test_batched_forces_match_oracle(gpu, ens, precision):
    from pyrdsim_amd.engine import Engine
    e = ens
    t = _with_actives(e, 0.3, 11)
    eng = Engine(e['models'], e['reps'], precision=precision, tile_rows=4, tile_particles=112)
    eng.set_state(e['x0'], t)
    f = eng.forces()
    eng.close()
    rtol = 1e-5 if precision == 'mixed' else 1e-10
    report = []
    for r in e['subset']:
        lo, hi = e['off'][r], e['off'][r + 1]
        s = _system(e, r, e['x0'][lo:hi], t[lo:hi])
        o = oracle_for(s)
        o.init()
        _, _, a = o.state()
        fr = f[lo:hi]
        scale = force_scale(s) + np.abs(a).max(axis=1)
        err = np.abs(fr - a).max(axis=1)
        bad = (err > rtol * scale) & ~near_cutoff(s)
        assert not bad.any(), (r, precision, int(bad.sum()))
        normwise = float(np.linalg.norm(fr - a) / np.linalg.norm(a))
        report.append((e['plan'][r]['kind'], e['plan'][r]['n'], normwise))
        assert normwise <= rtol, (r, precision, normwise)
    print('\nbatched %s forces, ||dF||/||F|| per replica: %s' % (precision, report))


def test_batched_fp64_injected_steps_match_oracle(gpu, ens):
    from pyrdsim_amd.engine import Engine
    e = ens
    steps = 4
    t = _with_actives(e, 0.4, 12)
    N = len(e['x0'])
    rng = np.random.Generator(np.random.PCG64(2024))
    normals = rng.standard_normal((steps, N, 2))
    u1 = rng.random((steps, N))
    u2 = rng.random((steps, N))
    # frequent reactions in the wall replicas
    for r, p in enumerate(e['plan']):
        if p['kind'] == 'wall':
            u1[:, e['off'][r]:e['off'][r + 1]] *= 1e-3
    eng = Engine(e['models'], e['reps'], precision='fp64', rng='injected', tile_rows=4,
                 tile_particles=112)
    eng.set_state(e['x0'], t)
    eng.set_injected(normals, u1, u2, first_step=1)
    eng.run(steps)
    eng.sync()
    xg, tg = eng.state()
    eng.close()
    events = 0
    for r in e['subset']:
        lo, hi = e['off'][r], e['off'][r + 1]
        s = _system(e, r, e['x0'][lo:hi], t[lo:hi])
        o = oracle_for(s, rng='injected')
        o.set_injected(np.ascontiguousarray(normals[:, lo:hi]), np.ascontiguousarray(u1[:, lo:hi]),
                       np.ascontiguousarray(u2[:, lo:hi]), first_step=1)
        o.init()
        o.run(1, steps)
        xo, to, _ = o.state()
        d = xg[lo:hi] - xo
        d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        assert np.abs(d).max() < 1e-12, (r, float(np.abs(d).max()))
        assert np.array_equal(tg[lo:hi], to), (r, int((tg[lo:hi] != to).sum()))
        events += o.reaction_info()[2]
    assert events > 0


def _planted(s, steps, seed, k_lin, u1_max=2e-4):
    """Injected streams for a wall system whose inactive particles react at
    rate k_lin: at step 1 the uniform of every third inactive particle lies
    halfway between 1 - exp(-dt k) and dt k, where the two forms disagree."""
    n = len(s['x'])
    rng = np.random.Generator(np.random.PCG64(seed))
    normals = rng.standard_normal((steps, n, 2))
    u1 = rng.random((steps, n)) * u1_max
    u2 = rng.random((steps, n))
    p_lin = s['dt'] * k_lin
    p_exp = -np.expm1(-s['dt'] * k_lin)
    planted = np.nonzero(s['types'] == 0)[0][::3]
    u1[0, planted] = 0.5 * (p_lin + p_exp)
    return normals, u1, u2, planted


@pytest.mark.parametrize('form', ['linear', 'exp'])
@pytest.mark.parametrize('u1_max', [2e-4, 5e-3])
def test_reaction_probability_forms_match_oracle(gpu, form, u1_max):
    """Both reaction forms against the oracle, event for event.  With u1 up
    to 5e-3 the active particles' draws (p = dt * k_d * count, at most about
    dt * k_d * 22 = 2.2e-3 for the 22 or so particles of a stencil) fall below
    p, between p and that bound, and above it (no walk)."""
    s = small('wall')
    s['types'][::4] = 1
    steps = 12
    normals, u1, u2, _ = _planted(s, steps, 31, 0.1, u1_max)
    eng = engine_for(s, precision='fp64', rng='injected', reaction_prob=form)
    eng.set_injected(normals, u1, u2, first_step=1)
    o = oracle_for(s, rng='injected', reaction_prob=form)
    o.set_injected(normals, u1, u2, first_step=1)
    o.init()
    for chunk in (1, 3, 8):
        eng.run(chunk)
        st = eng.sync()
        o.run(o.step + 1 if o.step else 1, chunk)
        xg, tg = eng.state()
        xo, to, _ = o.state()
        d = xg - xo
        d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        assert np.abs(d).max() < 1e-12, (form, chunk)
        assert np.array_equal(tg, to), (form, chunk, int((tg != to).sum()))
    assert st.reaction_events == o.reaction_info()[2] > 0


def test_planted_uniforms_separate_the_forms(gpu):
    """Uniforms between 1 - exp(-dt k) and dt k fire in the linear form only."""
    s = small('wall')
    normals, u1, u2, planted = _planted(s, 1, 32, 0.1)
    types = {}
    for form in ('linear', 'exp'):
        eng = engine_for(s, precision='fp64', rng='injected', reaction_prob=form)
        eng.set_injected(normals, u1, u2, first_step=1)
        eng.run(1)
        eng.sync()
        types[form] = eng.state()[1]
        eng.close()
    diff = np.nonzero(types['linear'] != types['exp'])[0]
    assert np.array_equal(diff, planted)
    assert (types['linear'][planted] == 1).all() and (types['exp'][planted] == 0).all()


def test_event_counts_of_the_two_forms_on_the_wall_model(gpu):
    """The wall + reaction model at full size (N = 65,536, k_a = 0.1,
    k_d = 1.0 per active neighbour) over 20,000 steps with the same Philox
    streams in both forms: the event counts differ by far less than their
    statistical spread (the forms differ by (dt k)^2 / 2 per particle-step)."""
    s = workloads.wall()
    out = {}
    for form in ('linear', 'exp'):
        eng = engine_for(s, precision='mixed', rng='philox', seed=2003, graph_steps=50,
                         reaction_prob=form)
        eng.run(20000)
        st = eng.sync()
        out[form] = (int(st.reaction_events), eng.state()[1])
        eng.close()
    ev_lin, ev_exp = out['linear'][0], out['exp'][0]
    differ = int((out['linear'][1] != out['exp'][1]).sum())
    print('\nwall model, 20000 steps: events linear %d, exp %d, final types differing %d'
          % (ev_lin, ev_exp, differ))
    assert ev_lin > 0 and ev_exp > 0
    assert abs(ev_lin - ev_exp) <= 5 * np.sqrt(ev_lin)
