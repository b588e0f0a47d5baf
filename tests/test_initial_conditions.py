"""Seeded initial conditions (the counterpart of pyrdsim.pyx:2463-2828), CPU
parts: the reference-molecule minimiser against the reference's own
geometry file, rigid random placement, the distance checks, and the
facade's argument errors."""

import os

import numpy as np
import pytest

from pyrdsim_amd import _parse
from pyrdsim_amd import initial_conditions as icg
from pyrdsim_amd.system import system

EX = os.path.join(os.path.dirname(__file__), 'golden', 'examples', 'patchy')


def _pdist(x):
    d = x[:, None, :] - x[None, :, :]
    return np.sort(np.sqrt((d * d).sum(-1)).ravel())


def _brute_min(x, mol, L, periodic):
    d = x[:, None, :] - x[None, :, :]
    per = np.asarray(periodic) > 0
    d -= L * np.round(d / L) * per
    r = np.sqrt((d * d).sum(-1))
    r[mol[:, None] == mol[None, :]] = np.inf
    return r.min()


def test_reference_molecule_reproduces_the_reference_geometry():
    """Minimising the bonded energy of bonded_interactions_3.txt (three
    patches bonded at 1.0 to a core, 120 degrees apart) gives the geometry of
    the reference's reference_configuration_3.txt up to a rotation."""
    _, inter = _parse.load_bonded_interactions(os.path.join(EX, 'bonded_interactions_3.txt'))
    ref = np.genfromtxt(os.path.join(EX, 'reference_configuration_3.txt'))
    rng = np.random.Generator(np.random.PCG64(3))
    x = icg.reference_molecule(inter[0], 4, 2, rng)
    assert np.allclose(x[0], 0.0)
    assert np.abs(_pdist(x) - _pdist(ref)).max() < 1e-4
    assert icg.bonded_energy(x, inter[0]) < 1e-6
    # the file's geometry is (to its 5 decimals) a zero of the same energy
    assert icg.bonded_energy(ref, inter[0]) < 1e-6


def test_reference_molecule_of_four_patches_is_a_cross():
    """bonded_interactions_4.txt (four patches at 1.0, 90 and 180 degree
    angles): the minimum is a square cross, distances 1 (x4), sqrt 2 (x4) and
    2 (x2) between distinct atoms."""
    _, inter = _parse.load_bonded_interactions(os.path.join(EX, 'bonded_interactions_4.txt'))
    x = icg.reference_molecule(inter[0], 5, 2, np.random.Generator(np.random.PCG64(4)))
    want = np.sort(np.r_[np.zeros(5), np.ones(8), np.full(8, np.sqrt(2)), np.full(4, 2.0)])
    assert np.abs(_pdist(x) - want).max() < 1e-4


def test_bonded_energy_closed_forms():
    """u0 (d - r0)^2 for a bond and u0 (theta - theta0)^2 for an angle, the
    reference's objective (pyrdsim.pyx:2509-2527)."""
    bond = [dict(type='bond', participants=[0, 1], r0=1.0, u0=400.0)]
    x = np.array([[0.0, 0.0], [1.3, 0.0]])
    assert np.isclose(icg.bonded_energy(x, bond), 400.0 * 0.09)
    ang = [dict(type='angle', participants=[1, 0, 2], theta0=np.pi / 2, u0=5.0)]
    x = np.array([[0.0, 0.0], [2.0, 0.0], [np.cos(2.0), np.sin(2.0)]])
    assert np.isclose(icg.bonded_energy(x, ang), 5.0 * (2.0 - np.pi / 2) ** 2)


@pytest.mark.parametrize('periodic', [[1, 1], [1, -1]])
def test_rigid_placement_with_minimum_distance(periodic):
    ref3 = np.genfromtxt(os.path.join(EX, 'reference_configuration_3.txt'))
    refs = {0: ref3, 1: np.array([[0.0, 0.0], [0.9, 0.0]])}
    types = {0: [0, 1, 1, 1], 1: [0, 2]}
    L = np.array([30.0, 25.0])
    rng = np.random.Generator(np.random.PCG64(11))
    out = icg.place_molecules({0: 20, 1: 15}, types, refs, L, periodic, 1.0, rng)
    assert out.shape == (20 * 4 + 15 * 2, 6)
    mol = out[:, 0].astype(int)
    assert np.array_equal(np.unique(mol), np.arange(35))
    assert np.all(np.diff(mol) >= 0)
    for m in range(35):
        rows = out[mol == m]
        mid = int(rows[0, 2])
        assert np.array_equal(rows[:, 1], np.arange(len(refs[mid])))
        assert np.array_equal(rows[:, 3], types[mid])
        # a rigid copy: the same pairwise distances
        assert np.abs(_pdist(rows[:, 4:]) - _pdist(refs[mid])).max() < 1e-12
    assert np.sum(out[out[:, 1] == 0, 2] == 0) == 20
    x = out[:, 4:]
    assert (x >= 0).all() and (x <= L).all()
    assert _brute_min(x, mol, L, periodic) >= 1.0
    # seeded
    again = icg.place_molecules({0: 20, 1: 15}, types, refs, L, periodic, 1.0,
                                np.random.Generator(np.random.PCG64(11)))
    assert np.array_equal(out, again)


@pytest.mark.parametrize('periodic', [[1, 1], [-1, 1], [-1, -1]])
def test_min_intermolecular_distance_matches_brute_force(periodic):
    rng = np.random.Generator(np.random.PCG64(5))
    L = np.array([12.0, 9.0])
    x = rng.random((120, 2)) * L
    mol = np.repeat(np.arange(40), 3)
    want = _brute_min(x, mol, L, periodic)
    assert icg.min_intermolecular_distance(x, mol, L, periodic, 2.0) == pytest.approx(want,
                                                                                      abs=1e-14)
    assert icg.min_intermolecular_distance(x, mol, L, periodic, want * 0.99) == np.inf


def test_placement_refuses_an_impossible_packing():
    refs = {0: np.array([[0.0, 0.0]])}
    rng = np.random.Generator(np.random.PCG64(1))
    with pytest.raises(RuntimeError, match='Could not place molecule'):
        icg.place_molecules({0: 50}, {0: [0]}, refs, np.array([5.0, 5.0]), [1, 1], 2.0, rng,
                            n_tries=50)


def test_reference_configuration_atom_count_is_checked(tmp_path):
    base = dict(parameters_filename=str(tmp_path / 'none.txt'), n_dim=2,
                l=np.array([50., 50.]), dt=1e-5, n_steps=10, n_out=10, interactions=[],
                reactions=[], tau_by_tauD=[1.0, 1.0],
                bonded_interactions_filename=os.path.join(EX, 'bonded_interactions_4.txt'))
    s = system(dict(base, seed=1))
    with pytest.raises(RuntimeError, match='has 4 atoms, but interaction parameters'):
        s.create_initial_conditions({0: 3}, {0: [0, 1, 1, 1, 1]},
                                    reference_configurations={
                                        0: os.path.join(EX, 'reference_configuration_3.txt')})


def test_random_rotations_3d_are_uniform_proper_rotations():
    """Haar-random rotations: orthonormal with determinant +1, and the image
    of a fixed axis is uniform on the sphere (mean ~ 0, second moment I/3)."""
    rng = np.random.Generator(np.random.PCG64(3))
    v = []
    for _ in range(4000):
        q = icg.random_rotation(3, rng)
        assert np.abs(q @ q.T - np.eye(3)).max() < 1e-12
        assert abs(np.linalg.det(q) - 1.0) < 1e-12
        v.append(q @ np.array([0.0, 0.0, 1.0]))
    v = np.array(v)
    assert np.abs(v.mean(axis=0)).max() < 0.05
    assert np.abs(v.T @ v / len(v) - np.eye(3) / 3).max() < 0.03


@pytest.mark.parametrize('periodic', [[1, 1, 1], [1, -1, 1]])
def test_rigid_placement_3d(periodic):
    """Three-dimensional placement: rows (molecule, index, id, type, x, y, z),
    rigid copies of a non-planar reference, inside the box, every pair of
    molecules at least the threshold apart (minimum image where periodic)."""
    ref = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    types = {0: [0, 1, 1, 1]}
    L = np.array([12.0, 10.0, 11.0])
    rng = np.random.Generator(np.random.PCG64(21))
    out = icg.place_molecules({0: 40}, types, {0: ref}, L, periodic, 1.0, rng)
    assert out.shape == (160, 7)
    mol = out[:, 0].astype(int)
    for m in range(40):
        rows = out[mol == m]
        assert np.abs(_pdist(rows[:, 4:]) - _pdist(ref)).max() < 1e-12
        # proper rotation: the chirality of the reference is kept
        e = rows[1:, 4:] - rows[0, 4:]
        assert np.linalg.det(e) > 0
    x = out[:, 4:]
    assert (x >= 0).all() and (x <= L).all()
    assert _brute_min(x, mol, L, periodic) >= 1.0
