"""Precision modes and tuning on the GPU.

* FP32 pair forces with FP64 positions ('mixed') and with compensated FP32
  positions ('compensated') reproduce the FP64 path's radial distribution
  function within 1% (L1, relative) and its activated fraction, averaged over
  the second half of a run, with the same Philox streams;
* against an FP64 trajectory with the same injected noise, compensated
  positions stay as close as FP64 positions do (both limited by the FP32 pair
  arithmetic), far below the per-step rounding that plain FP32 coordinates
  would carry (half an ulp of the box length);
* re-tiling a live context (what the autotuner does) changes nothing but the
  work decomposition.
"""

import numpy as np
import pytest

from helpers import engine_for, small
from pyrdsim_amd import analysis
from pyrdsim_amd.autotune import autotune

pytestmark = pytest.mark.gpu


def _cluster_distribution(x, t, s, kmax=64):
    """Share of the type-1 particles in clusters of each size (linked at
    1.5 sigma), sizes above kmax pooled."""
    sizes = analysis.cluster_sizes(x, t, s['L'], s['periodic'], active=1, r_link=1.5)
    h = np.zeros(kmax + 1)
    for k in sizes:
        h[min(int(k), kmax)] += k
    return h / max(1.0, h.sum())


def _averaged(eng, s, steps, every):
    gs, acts, cls = [], [], []
    for k in range(steps // every):
        eng.run(every)
        if k >= (steps // every) // 2:
            x, t = eng.state()
            _, g = analysis.rdf(x, s['L'], s['periodic'])
            gs.append(g)
            acts.append(analysis.activated_fraction(t))
            cls.append(_cluster_distribution(x, t, s))
    return np.mean(gs, axis=0), float(np.mean(acts)), np.mean(cls, axis=0)


@pytest.mark.parametrize('name', ['kalj', 'wall', 'patchy'])
def test_mixed_precision_rdf_within_one_percent(gpu, name):
    """RDF within 1%, activated fraction within 0.01 and the cluster-size
    distribution within 0.05 (L1) of the FP64 run, averaged over the second
    half of the run."""
    s = small(name, n_mol=1500, seed=21) if name == 'patchy' else small(name, n=6000, seed=21)
    steps, every = 4000, 200
    res = {}
    for prec in ('fp64', 'mixed', 'compensated'):
        eng = engine_for(s, precision=prec, seed=77, graph_steps=20)
        res[prec] = _averaged(eng, s, steps, every)
        eng.close()
    for prec in ('mixed', 'compensated'):
        err = analysis.rdf_error(res[prec][0], res['fp64'][0])
        assert err <= 0.01, (prec, err)
        assert abs(res[prec][1] - res['fp64'][1]) <= 0.01, prec
        assert np.abs(res[prec][2] - res['fp64'][2]).sum() <= 0.05, prec


@pytest.mark.parametrize('name', ['kalj', 'wall', 'patchy'])
def test_full_size_configurations(gpu, name):
    """The three single-system configurations at their full sizes (LJ
    mixture N = 262,144 at density 0.6; wall + reactions N = 65,536 at density
    0.4; patchy colloids, 16,384 molecules of four bonded beads at packing
    fraction 0.3): with the same Philox streams, FP32-force runs (FP64 or
    compensated positions) stay within 2e-5 sigma of the FP64 run after 100
    steps (the seeded states hold close pairs whose forces reach 1e3), keep
    every particle in the box, and make the same reaction events up to flips
    at pair distances within the runs' difference of r0."""
    from pyrdsim_amd import workloads
    s = {'kalj': workloads.kalj, 'wall': workloads.wall, 'patchy': workloads.patchy}[name]()
    if name == 'wall':
        s['types'][::5] = 1
    if name == 'patchy':
        s['types'][2::4] = 2   # some active patches: tanh attraction and catalysis
    out, ev = {}, {}
    for prec in ('fp64', 'mixed', 'compensated'):
        eng = engine_for(s, precision=prec, seed=9, graph_steps=25)
        eng.run(100)
        ev[prec] = eng.sync().reaction_events
        out[prec] = eng.state()
        eng.close()
    L = s['L']
    for prec in ('mixed', 'compensated'):
        x, t = out[prec]
        assert np.isfinite(x).all() and (x >= 0).all() and (x <= L).all()
        # a reaction decision can flip where a pair distance lies within the
        # runs' 1e-8 difference of r0; a flipped type changes its force law, so
        # positions are compared away from such particles
        flip = t != out['fp64'][1]
        assert flip.sum() <= 20 and abs(ev[prec] - ev['fp64']) <= 20, (prec, int(flip.sum()))
        keep = np.ones(len(x), dtype=bool)
        if flip.any():
            from scipy.spatial import cKDTree
            box = np.where(s['periodic'] > 0, L, 4 * L + 100)
            tree = cKDTree(np.mod(out['fp64'][0], L) % L, boxsize=box)
            for nb in tree.query_ball_point(np.mod(out['fp64'][0][flip], L) % L, 4.0):
                keep[nb] = False
        d = x - out['fp64'][0]
        d = d - L * np.round(d / L) * (s['periodic'] > 0)
        print(name, prec, 'max deviation from FP64:', float(np.abs(d[keep]).max()),
              'flipped types:', int(flip.sum()))
        assert np.abs(d[keep]).max() < 2e-5, (prec, np.abs(d[keep]).max())


@pytest.mark.parametrize('name', ['kalj', 'wall', 'patchy'])
def test_compensated_positions_track_fp64(gpu, name):
    from test_gpu_parity import _injected
    s = small(name)
    steps = 200
    normals, u1, u2 = _injected(len(s['x']), steps, 31)
    if name == 'wall':
        u1 *= 1e-3
    out = {}
    for prec in ('fp64', 'mixed', 'compensated'):
        eng = engine_for(s, precision=prec, rng='injected')
        eng.set_injected(normals, u1, u2, first_step=1)
        eng.run(steps)
        eng.sync()
        out[prec] = eng.state()
        eng.close()
    dev = {}
    for prec in ('mixed', 'compensated'):
        d = out[prec][0] - out['fp64'][0]
        d = d - s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        dev[prec] = float(np.abs(d).max())
        assert np.array_equal(out[prec][1], out['fp64'][1]), prec
    # plain FP32 coordinates round every step: a random walk of rounding errors
    # of about half an ulp of the box length
    plain = float(np.spacing(np.float32(max(s['L'])))) / 2 * np.sqrt(steps / 3.0)
    print(name, dev, 'plain FP32 coordinates, expected:', plain)
    assert dev['compensated'] <= 3 * dev['mixed'] + 1e-12, dev
    assert dev['compensated'] < 1e-2 * plain, (dev, plain)


def test_retiling_keeps_state_and_forces(gpu):
    s = small('wall', n=5000, seed=3)
    s['types'][::4] = 1
    eng = engine_for(s, precision='fp64', seed=4)
    eng.run(10)
    x0, t0 = eng.state()
    f0 = eng.forces()
    for rows, per in ((2, 40), (4, 96), (1, 8)):
        eng.set_tiling(rows, per)
        assert eng.tiling()[:2] == (rows, per)
        x1, t1 = eng.state()
        assert np.array_equal(x0, x1) and np.array_equal(t0, t1)
        assert np.abs(eng.forces() - f0).max() <= 1e-12 * (np.abs(f0).max() + 1)
    best, times = autotune(eng, steps=2)
    assert best in times and eng.tiling()[:2] == best
