"""Analysis on the device (pair histogram for the RDF, type counts, cluster
labels) against host brute force on the same state, for a batch of a
two-species LJ replica and two wall-model replicas with active particles, in
cell mode, list mode and with compensated positions.

Bars: integer results equal, except where a pair lies within 1e-9 (relative)
of a bin edge or of the link length, which is flagged; the RDF from the
device counts matches analysis.rdf to within those flagged pairs.
"""

import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components

from helpers import brute_pairs, small
from pyrdsim_amd import analysis

pytestmark = pytest.mark.gpu


def _ensemble(precision='mixed', skin=0.0):
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    systems = [small('kalj', n=1500, seed=21), small('wall', n=1800, seed=22),
               small('wall', n=900, seed=23)]
    systems[1]['types'][::2] = 1
    systems[2]['types'][::3] = 1
    models, reps = [], []
    for k, s in enumerate(systems):
        m = s['model']
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        models.append(model)
        reps.append(make_replica(k, len(s['x']), s['L'], s['periodic'], s['dt'],
                                 m['tau_by_tauD'], rates, 900 + k))
    eng = Engine(models, reps, precision=precision, graph_steps=10, skin=skin)
    eng.set_state(np.concatenate([s['x'] for s in systems]),
                  np.concatenate([s['types'] for s in systems]))
    eng.run(50)
    eng.sync()
    return eng, systems


def _host_hist(x, L, periodic, r_max, dr, nb):
    i, j, r = brute_pairs(x, L, periodic, r_max)
    q = r / dr
    h = np.bincount(np.floor(q).astype(int)[q < nb], minlength=nb)[:nb]
    flagged = int((np.abs(q - np.round(q)) < 1e-9 * np.maximum(q, 1.0)).sum())
    return h, flagged


def _host_labels(x, types, L, periodic, r_link, active):
    sel = np.flatnonzero(np.isin(types, active))
    i, j, r = brute_pairs(x[sel], L, periodic, r_link + 1e-6)
    near = int((np.abs(r - r_link) < 1e-9 * r_link).sum())
    keep = r <= r_link
    n = len(sel)
    m = coo_matrix((np.ones(keep.sum()), (i[keep], j[keep])), shape=(n, n))
    _, comp = connected_components(m, directed=False)
    lab = np.full(len(x), -1)
    first = np.full(comp.max() + 1 if n else 0, np.iinfo(np.int64).max)
    np.minimum.at(first, comp, sel)
    lab[sel] = first[comp]
    return lab, near


@pytest.mark.parametrize('precision,skin', [('fp64', 0.0), ('mixed', 0.0), ('mixed', 0.3),
                                            ('compensated', 0.0)])
def test_device_analysis_matches_host(gpu, precision, skin):
    eng, systems = _ensemble(precision, skin)
    x, t = eng.state()
    r_max, dr = 5.0, 0.02
    nb = int(round(r_max / dr))
    hist = eng.pair_histogram(r_max, dr)
    hist1 = eng.pair_histogram(r_max, dr, types=(1,))
    counts = eng.type_counts()
    labels = eng.cluster_labels(1.5, types=(1,))
    lo = 0
    for k, s in enumerate(systems):
        n = len(s['x'])
        xs, ts = x[lo:lo + n], t[lo:lo + n]
        h, flagged = _host_hist(xs, s['L'], s['periodic'], r_max, dr, nb)
        assert np.abs(hist[k] - h).sum() <= 2 * flagged, (k, np.abs(hist[k] - h).sum())
        assert hist[k].sum() > 1000
        act = ts == 1
        h1, f1 = _host_hist(xs[act], s['L'], s['periodic'], r_max, dr, nb)
        assert np.abs(hist1[k] - h1).sum() <= 2 * f1
        _, g_host = analysis.rdf(xs, s['L'], s['periodic'], r_max, dr)
        _, g_dev = analysis.rdf_from_counts(hist[k], n, s['L'], r_max, dr)
        assert analysis.rdf_error(g_dev, g_host) < 1e-3
        assert np.array_equal(counts[k][:3], np.bincount(ts, minlength=3)[:3])
        want, near = _host_labels(xs, ts, s['L'], s['periodic'], 1.5, [1])
        got = labels[lo:lo + n]
        got = np.where(got >= 0, got - lo, -1)  # labels are global particle indices
        if near == 0:
            assert np.array_equal(got, want), (k, int((got != want).sum()))
        if act.any():
            assert np.array_equal(np.sort(analysis.cluster_sizes_from_labels(got)),
                                  np.sort(analysis.cluster_sizes(xs, ts, s['L'],
                                                                 s['periodic'], 1, 1.5)))
        lo += n


def test_analysis_does_not_change_the_trajectory(gpu):
    """Analysis between steps leaves the run as it was: list mode keeps its
    cells (no extra rebuild), and positions agree to the summation order
    that the atomic ranks within a cell leave open."""
    outs, rebuilds = [], []
    for analyse in (False, True):
        eng, _ = _ensemble('fp64', 0.2)
        if analyse:
            eng.pair_histogram(4.0, 0.05)
            eng.cluster_labels(1.5, types=(1,))
            eng.type_counts()
        eng.run(30)
        outs.append(eng.state())
        rebuilds.append(eng.sync().rebuilds)
    assert np.abs(outs[0][0] - outs[1][0]).max() < 1e-9
    assert np.array_equal(outs[0][1], outs[1][1])
    assert rebuilds[0] == rebuilds[1]


def test_analysis_radius_limits(gpu):
    eng, systems = _ensemble('mixed')
    half = min(float(s['L'].min()) for s in systems) / 2
    with pytest.raises(RuntimeError, match='half the box length'):
        eng.pair_histogram(half * 1.01, 0.1)
    with pytest.raises(RuntimeError, match='half the box length'):
        eng.cluster_labels(half * 1.01)
