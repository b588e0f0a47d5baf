"""Seeded initial conditions for molecular systems: the counterpart of the
reference's ``create_initial_conditions`` (pyrdsim.pyx:2463-2828).

The reference inserts molecules one at a time: it picks the best of 1000
random rigid placements of a reference geometry (largest distance between the
new molecule's first atom and the placed first atoms), and after each
insertion it runs a whole ``simulate()`` to relax the system.  Its random
numbers come from unseeded ``np.random.default_rng()`` calls, and without
reference geometries it minimises the bonded energy with CMA-ES.

Here, everything is driven by one seeded ``numpy.random.Generator``:
* reference geometries are read from files (``np.genfromtxt``, as the
  reference does) or found by minimising the reference's bonded energy
  (pyrdsim.pyx:2496-2527) from several seeded random starts with SciPy's
  BFGS (the ``cma`` package the reference uses is not available);
* each molecule is a rigid copy of its geometry, rotated uniformly at random
  (a uniform angle in 2-D, a Haar-random rotation in 3-D), moved into the box and translated uniformly, as in
  ``place_configuration_randomly`` (pyrdsim.pyx:2576-2611), and accepted only
  if every atom is at least ``min_distance`` from every placed atom (periodic
  axes by minimum image);
* the relaxation runs once, on the GPU engine, after all molecules are placed.
"""

import itertools

import numpy as np


def bonded_energy(x, interactions, smooth=False):
    """The reference's objective for a reference molecule (pyrdsim.pyx:2509-2527):
    sum of u0 (d - r0)^2 over bonds and u0 (theta - theta0)^2 over angles, with
    the cosine rounded to 7 decimals before acos.  smooth=True clips the
    cosine to [-1, 1] instead of rounding it, which keeps the energy
    differentiable for the minimiser."""
    v = 0.0
    for it in interactions:
        p = it['participants']
        if it['type'] == 'bond':
            d = np.linalg.norm(x[p[0]] - x[p[1]])
            v += it['u0'] * (d - it['r0']) ** 2
        elif it['type'] == 'angle':
            a = x[p[1]] - x[p[0]]
            b = x[p[1]] - x[p[2]]
            h = np.sum(a * b) / (np.linalg.norm(a) * np.linalg.norm(b))
            h = np.clip(h, -1.0, 1.0) if smooth else np.round(h, 7)
            v += it['u0'] * (np.arccos(h) - it['theta0']) ** 2
    return v


def reference_molecule(interactions, n_atoms, n_dim, rng, n_tries=10):
    """Lowest-energy geometry of one molecule over n_tries seeded starts
    (pyrdsim.pyx:2463-2574): the first atom stays at the origin, the others
    start uniform in the unit cube."""
    from scipy.optimize import minimize
    if n_atoms < 2:
        return np.zeros((n_atoms, n_dim))

    def f(flat, smooth=True):
        x = np.concatenate([np.zeros((1, n_dim)), flat.reshape(n_atoms - 1, n_dim)])
        return bonded_energy(x, interactions, smooth)

    best, best_e = None, np.inf
    for _ in range(n_tries):
        x0 = rng.random((n_atoms - 1) * n_dim)
        res = minimize(f, x0, method='BFGS', options=dict(maxiter=5000, gtol=1e-9))
        e = f(res.x, smooth=False)
        if e < best_e:
            best_e = e
            best = np.concatenate([np.zeros((1, n_dim)), res.x.reshape(n_atoms - 1, n_dim)])
    return best


def random_rotation(d, rng):
    """A rotation drawn uniformly (Haar measure) from SO(d): in two dimensions
    a uniform angle; in general the Q factor of a Gaussian matrix with the
    signs of R's diagonal folded in (Mezzadri, "How to generate random
    matrices from the classical compact groups", Notices AMS 54 (2007)),
    its determinant made +1."""
    if d == 2:
        a = rng.random() * 2 * np.pi
        c, s = np.cos(a), np.sin(a)
        return np.array([[c, -s], [s, c]])
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def place_randomly(config, L, rng):
    """A rigid copy of config rotated uniformly at random, shifted so that its
    minimum is at 0 on each axis and translated by U(0, L - extent)
    (pyrdsim.pyx:2576-2611), in two or three dimensions."""
    config = np.asarray(config, dtype=float)
    d = config.shape[1]
    out = config @ random_rotation(d, rng).T
    out -= np.minimum(out.min(axis=0), 0.0)
    ext = out.max(axis=0)
    if np.any(ext > L):
        raise RuntimeError('a molecule is larger than the box')
    return out + rng.random(d) * (L - ext)


class _Grid:
    """Atoms placed so far, hashed into bins of side >= min_distance."""

    def __init__(self, L, periodic, h):
        self.L = np.asarray(L, dtype=float)
        self.per = np.asarray(periodic)[:len(self.L)] > 0
        self.n = np.maximum((self.L // h).astype(int), 1)
        self.h = self.L / self.n
        self.bins = {}
        self.pts = []
        self.offsets = list(itertools.product((-1, 0, 1), repeat=len(self.L)))

    def _bin(self, p):
        return tuple(np.minimum((p // self.h).astype(int), self.n - 1))

    def min_distance(self, atoms):
        best = np.inf
        for p in atoms:
            b = self._bin(p)
            for off in self.offsets:
                q = [b[k] + off[k] for k in range(len(b))]
                skip = False
                for k in range(len(q)):
                    if q[k] < 0 or q[k] >= self.n[k]:
                        if not self.per[k]:
                            skip = True
                        q[k] %= self.n[k]
                if skip:
                    continue
                idx = self.bins.get(tuple(q))
                if not idx:
                    continue
                d = np.asarray([self.pts[i] for i in idx]) - p
                d -= self.L * np.round(d / self.L) * self.per
                best = min(best, float(np.sqrt((d * d).sum(axis=1)).min()))
        return best

    def add(self, atoms):
        for p in atoms:
            self.bins.setdefault(self._bin(p), []).append(len(self.pts))
            self.pts.append(p)


def place_molecules(counts, particle_types, refs, L, periodic, min_distance, rng,
                    n_tries=1000):
    """Output array of the reference's layout (pyrdsim.pyx:2711-2712): one row
    per atom, (molecule number, index in molecule, molecule id, particle type,
    x, y[, z]), molecules in a seeded random order of their ids."""
    order = np.concatenate([np.full(int(n), int(mid)) for mid, n in counts.items()])
    rng.shuffle(order)
    total = int(sum(len(refs[m]) for m in order))
    out = np.zeros((total, 4 + len(L)))
    grid = _Grid(L, periodic, max(min_distance, 1e-9))
    row = 0
    for i, mid in enumerate(order):
        ref = refs[mid]
        for _ in range(n_tries):
            cand = place_randomly(ref, L, rng)
            if grid.min_distance(cand) >= min_distance:
                break
        else:
            raise RuntimeError('Could not place molecule {0}'.format(i + 1))
        grid.add(cand)
        n = len(ref)
        out[row:row + n, 0] = i
        out[row:row + n, 1] = np.arange(n)
        out[row:row + n, 2] = mid
        out[row:row + n, 3] = particle_types[mid]
        out[row:row + n, 4:] = cand
        row += n
    return out


def min_intermolecular_distance(x, molecule, L, periodic, r_max):
    """Smallest distance between atoms of different molecules (minimum image
    on periodic axes), the reference's min_dist (pyrdsim.pyx:1640-1650), if
    it is below r_max; otherwise inf."""
    grid = _Grid(L, periodic, r_max)
    best = np.inf
    starts = np.flatnonzero(np.r_[True, molecule[1:] != molecule[:-1]])
    ends = np.r_[starts[1:], len(x)]
    for a, b in zip(starts, ends):
        best = min(best, grid.min_distance(x[a:b]))
        grid.add(x[a:b])
    return best if best < r_max else np.inf
