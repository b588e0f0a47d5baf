"""Observables used to compare precision modes (CPU only)."""

import numpy as np

from pyrdsim_amd import analysis


def test_ideal_gas_rdf_is_one():
    rng = np.random.Generator(np.random.PCG64(0))
    L = np.array([200.0, 200.0])
    x = rng.random((40000, 2)) * L
    r, g = analysis.rdf(x, L, [1, 1], r_max=5.0, dr=0.1)
    assert np.abs(g[5:] - 1).max() < 0.05
    # reflecting box: the normalisation ignores the walls, so g dips slightly
    r, g2 = analysis.rdf(x, L, [1, -1], r_max=5.0, dr=0.1)
    assert np.abs(g2[5:] - 1).max() < 0.08


def test_square_lattice_rdf_peaks():
    n = 30
    g1 = np.arange(n) * 1.0
    x = np.stack(np.meshgrid(g1, g1), -1).reshape(-1, 2) + 0.5
    L = [float(n), float(n)]
    r, g = analysis.rdf(x, L, [1, 1], r_max=2.1, dr=0.05)
    nz = r[g > 0]
    # shells at 1, sqrt(2), 2 with 4, 4, 4 neighbours
    assert np.allclose(sorted(set(np.round(nz, 2))), [1.02, 1.42, 2.02], atol=0.03)
    assert analysis.rdf_error(g, g) == 0.0


def test_activated_fraction_and_clusters():
    x = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [10.0, 10.0], [10.5, 10.0], [19.8, 5.0],
                  [0.2, 5.0]])
    t = np.array([1, 1, 1, 1, 1, 1, 0])
    assert analysis.activated_fraction(t) == 6 / 7
    sizes = analysis.cluster_sizes(x, t, [20.0, 20.0], [1, 1], r_link=1.5)
    # chain of 3, a pair, and one isolated (19.8, 5) whose periodic image is near
    # (0.2, 5), which is inactive
    assert sorted(sizes.tolist()) == [1, 2, 3]


def test_ideal_gas_rdf_is_one_in_three_dimensions():
    rng = np.random.Generator(np.random.PCG64(1))
    L = np.array([40.0, 40.0, 40.0])
    x = rng.random((40000, 3)) * L
    r, g = analysis.rdf(x, L, [1, 1, 1], r_max=4.0, dr=0.2)
    assert np.abs(g[3:] - 1).max() < 0.05
    # the same normalisation from a pair histogram
    counts = np.round(g * 40000 * (40000 / L.prod()) * analysis._shell(
        np.arange(0.0, 4.0 + 0.1, 0.2), 3) / 2.0)
    _, g2 = analysis.rdf_from_counts(counts, 40000, L, 4.0, 0.2, n_dim=3)
    assert np.allclose(g2, g, rtol=1e-6, atol=1e-6)


def test_two_dimensional_points_with_three_entry_box():
    # replica records carry three box lengths; the coordinates' columns decide
    rng = np.random.Generator(np.random.PCG64(2))
    x = rng.random((20000, 2)) * 100.0
    _, g = analysis.rdf(x, [100.0, 100.0, 0.0], [1, 1, 0], r_max=3.0, dr=0.1)
    _, g2 = analysis.rdf(x, [100.0, 100.0], [1, 1], r_max=3.0, dr=0.1)
    assert np.array_equal(g, g2)
    sizes = analysis.cluster_sizes(x, np.ones(len(x), int), [100.0, 100.0, 0.0], [1, 1, 0],
                                   r_link=0.3)
    assert sizes.sum() == len(x)
