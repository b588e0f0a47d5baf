"""Observables of the step loop's output: radial distribution function,
activated fraction and cluster-size histogram.

The reference computes none of these (it only writes trajectories); they are
how the engine's precision modes are compared with the FP64 path:
* RDF: pair histogram to r_max with bin dr, normalised by the ideal-gas
  shell count N rho V_shell / 2 per particle pair, V_shell = pi ((r+dr)^2 - r^2)
  in two dimensions and 4/3 pi ((r+dr)^3 - r^3) in three;
* RDF error: || g - g_ref ||_1 / || g_ref ||_1 over the bins;
* activated fraction: share of particles of a given type;
* clusters: connected components of the given type with links shorter than
  r_link (minimum image).
"""

import numpy as np
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree


def _geometry(x, L, periodic):
    """The coordinates' dimension d (their number of columns) and the first d
    box lengths and periodic flags (a caller may pass three-entry arrays for
    a two-dimensional system)."""
    x = np.asarray(x, dtype=np.float64)
    d = x.shape[1] if x.ndim == 2 else 2
    return x, d, np.asarray(L, dtype=np.float64)[:d], np.asarray(periodic)[:d]


def _shell(edges, d):
    """Volume of the shells between consecutive edges in d dimensions."""
    if d == 3:
        return 4.0 / 3.0 * np.pi * (edges[1:] ** 3 - edges[:-1] ** 3)
    return np.pi * (edges[1:] ** 2 - edges[:-1] ** 2)


def _tree(x, L, periodic):
    """KD-tree with periodic axes; a reflecting axis gets a box much larger
    than the system, so that it never wraps."""
    x, d, L, periodic = _geometry(x, L, periodic)
    box = np.where(periodic > 0, L, 4.0 * L + 100.0)
    xx = x.copy()
    for k in range(d):
        if periodic[k] > 0:
            xx[:, k] = np.mod(xx[:, k], L[k])
            xx[:, k][xx[:, k] >= L[k]] = 0.0
    return cKDTree(xx, boxsize=box)


def rdf(x, L, periodic, r_max=5.0, dr=0.02, select=None):
    """g(r) at bin centres; `select` (bool mask) restricts to a species."""
    x, nd, L, periodic = _geometry(x, L, periodic)
    if select is not None:
        x = x[select]
    n = len(x)
    edges = np.arange(0.0, r_max + dr / 2, dr)
    if n < 2:
        return 0.5 * (edges[1:] + edges[:-1]), np.zeros(len(edges) - 1)
    tree = _tree(x, L, periodic)
    pairs = tree.query_pairs(r_max, output_type='ndarray')
    d = x[pairs[:, 1]] - x[pairs[:, 0]]
    for k in range(nd):
        if periodic[k] > 0:
            d[:, k] -= L[k] * np.round(d[:, k] / L[k])
    r = np.sqrt((d ** 2).sum(axis=1))
    counts, _ = np.histogram(r, bins=edges)
    rho = n / float(np.prod(L))
    g = 2.0 * counts / (n * rho * _shell(edges, nd))
    return 0.5 * (edges[1:] + edges[:-1]), g


def rdf_from_counts(counts, n, L, r_max=5.0, dr=0.02, n_dim=2):
    """g(r) at bin centres from a pair histogram (the engine's
    pair_histogram), with the normalisation of rdf() in n_dim dimensions."""
    edges = np.arange(0.0, r_max + dr / 2, dr)
    counts = np.asarray(counts, dtype=np.float64)[:len(edges) - 1]
    if n < 2:
        return 0.5 * (edges[1:] + edges[:-1]), np.zeros(len(edges) - 1)
    rho = n / float(np.prod(np.asarray(L, dtype=np.float64)[:n_dim]))
    return 0.5 * (edges[1:] + edges[:-1]), 2.0 * counts / (n * rho * _shell(edges, n_dim))


def cluster_sizes_from_labels(labels):
    """Cluster sizes from per-particle cluster labels (-1 = not in a cluster)."""
    lab = np.asarray(labels)
    lab = lab[lab >= 0]
    if lab.size == 0:
        return np.zeros(0, dtype=int)
    return np.unique(lab, return_counts=True)[1]


def rdf_error(g, g_ref):
    return float(np.abs(g - g_ref).sum() / np.abs(g_ref).sum())


def activated_fraction(types, active=1):
    return float(np.mean(np.asarray(types) == active))


def cluster_sizes(x, types, L, periodic, active=1, r_link=1.5):
    """Sizes of connected clusters of `active` particles (links < r_link)."""
    sel = np.asarray(types) == active
    xa = np.asarray(x)[sel]
    n = len(xa)
    if n == 0:
        return np.zeros(0, dtype=int)
    pairs = _tree(xa, L, periodic).query_pairs(r_link, output_type='ndarray')
    m = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(n, n))
    _, lab = connected_components(m, directed=False)
    return np.bincount(lab)
