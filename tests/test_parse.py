"""The facade's parsers on the reference's own example files (CPU only).

Expected values are read off the fixture files by hand; the rules exercised
are those of pyrdsim.pyx:590-1034 (case folding, '#'/';' comments, sections,
int/float arrays, casts, errors) and pyx:1790-1934 (initial conditions and the
state text file).
"""

import os
import shutil

import numpy as np
import pytest

from pyrdsim_amd import _parse
from pyrdsim_amd.system import system

EX = os.path.join(os.path.dirname(__file__), 'golden', 'examples')

NB_TYPES = {'type': str, 'u0': float, 'r0': float, 'rc': float, 'dr': float, 'a': float,
            'n': int, 'particle_type': int, 'direction': int, 'position': float}
RX_TYPES = {'type': str, 'rate': float, 'reactant': int, 'product': int, 'r0': float,
            'catalyst': int, 'rate_decrease': float, 'inhibitor': int, 'inhibitor2': int}
P_TYPES = {'N_dim': int, 'dt': float, 'N_steps': int, 'stride': int, 'N_out': int,
           'bonded_interactions_filename': str, 'nonbonded_interactions_filename': str,
           'reactions_filename': str, 'initial_conditions_filename': str,
           'trajectory_filename': str, 'neighbor_list_update_frequency': int, 'rs': float,
           'cell_length_ratio': int}


def test_parameters_file():
    p = _parse.load_parameters_file(os.path.join(EX, 'lj_wall', 'parameters.txt'), P_TYPES,
                                    float_arrays=['L', 'periodic', 'tau_by_tauD'])[0]
    assert p['n_dim'] == 2 and p['dt'] == 5e-05 and p['n_steps'] == 20000000
    assert p['stride'] == 1000 and p['n_out'] == 10000
    assert np.array_equal(p['l'], [50.0, 15.0]) and np.array_equal(p['periodic'], [1, -1])
    assert np.array_equal(p['tau_by_taud'], [1.0, 1.0])
    assert p['nonbonded_interactions_filename'] == 'nonbonded_interactions.txt'
    assert p['rs'] == 7.0 and p['cell_length_ratio'] == 2
    assert p['neighbor_list_update_frequency'] == 20


def test_interaction_sections():
    it = _parse.load_parameters_file(os.path.join(EX, 'lj_wall', 'nonbonded_interactions.txt'),
                                     NB_TYPES, int_arrays=['pair'],
                                     section_keyword='[interaction]')
    assert [d['type'] for d in it] == ['lj', 'lj', 'lj', 'tanh', 'wall']
    assert np.array_equal(it[1]['pair'], [0, 1]) and it[1]['n'] == 10 and it[1]['u0'] == 0.2
    assert it[3]['a'] == 0.05 and it[3]['rc'] == 3.15 and it[3]['r0'] == 2.15
    w = it[4]
    assert w['particle_type'] == 1 and w['direction'] == 1 and w['u0'] == 10.0
    assert w['a'] == 0.1 and w['r0'] == 4.0


def test_reaction_sections_and_empty_file():
    rx = _parse.load_parameters_file(os.path.join(EX, 'lj_wall', 'reactions.txt'), RX_TYPES,
                                     section_keyword='[reaction]')
    assert rx[0] == {'type': 'linear', 'rate': 0.1, 'reactant': 0, 'product': 1}
    assert rx[1]['type'] == 'locally_catalyzed' and rx[1]['r0'] == 3.0
    assert rx[1]['catalyst'] == 1 and rx[1]['product'] == 0 and rx[1]['rate'] == 0.02
    empty = _parse.load_parameters_file(os.path.join(EX, 'lj_one_species', 'reactions.txt'),
                                        RX_TYPES, section_keyword='[reaction]')
    assert empty == []


def test_bonded_interactions():
    par, inter = _parse.load_bonded_interactions(
        os.path.join(EX, 'patchy', 'bonded_interactions_3.txt'))
    assert par == {0: {'molecule_id': 0, 'n_atoms': 4}}
    bonds = [d for d in inter[0] if d['type'] == 'bond']
    angles = [d for d in inter[0] if d['type'] == 'angle']
    assert [b['participants'] for b in bonds] == [[0, 1], [0, 2], [0, 3]]
    assert all(b['r0'] == 1.0 and b['u0'] == 400.0 for b in bonds)
    assert [a['participants'] for a in angles] == [[1, 0, 2], [2, 0, 3], [3, 0, 1]]
    assert abs(angles[0]['theta0'] - 2 * np.pi / 3) < 1e-15
    assert abs(angles[0]['u0'] - 0.171934 * (180 / np.pi) ** 2) < 1e-12


def test_parser_errors(tmp_path):
    f = tmp_path / 'bad.txt'
    f.write_text('[interaction]\ntype = lj\nbogus = 3\n')
    with pytest.raises(RuntimeError, match='Unrecognized entry'):
        _parse.load_parameters_file(str(f), NB_TYPES, section_keyword='[interaction]')
    f.write_text('u0 = 1\n[interaction]\ntype = lj\n')
    with pytest.raises(RuntimeError, match='outside of section'):
        _parse.load_parameters_file(str(f), NB_TYPES, section_keyword='[interaction]')
    f.write_text('[interaction]\nu0 = 1 = 2\n')
    with pytest.raises(RuntimeError, match='not of the form'):
        _parse.load_parameters_file(str(f), NB_TYPES, section_keyword='[interaction]')


def _write_ic(path, rows):
    with open(path, 'w') as f:
        for r in rows:
            f.write(' '.join(str(v) for v in r) + '\n')


def test_initial_conditions_and_state_file(tmp_path):
    ic = tmp_path / 'ic.txt'
    _write_ic(ic, [(0, 0, 0, 0, 1.5, 2.25), (0, 1, 0, 1, 1.0, 2.0), (1, 0, 0, 1, 3.0, 4.0),
                   ('# a comment',), (2, 0, 0, 0, 5.123456789, 6.0)])
    d = _parse.load_initial_conditions(str(ic), 2, 2)
    assert d['n_particles'] == 4 and d['n_molecules'] == 3
    assert np.array_equal(d['first_index_of_molecule'], [0, 2, 3])
    assert d['x'].dtype == np.float64 and d['particle_types'].dtype == np.intc
    text = _parse.format_state(d['molecule_ids'], d['particle_types'], d['x'])
    assert text.splitlines()[3] == '2   0   0   0   5.12346   6.00000   '
    with pytest.raises(RuntimeError, match='friction coefficients'):
        _parse.load_initial_conditions(str(ic), 2, 1)
    with pytest.raises(RuntimeError, match='inconsistent'):
        _parse.load_initial_conditions(str(ic), 3, 2)


def test_system_construction_from_example_files(tmp_path):
    for f in ('parameters.txt', 'nonbonded_interactions.txt', 'reactions.txt',
              'bonded_interactions.txt'):
        shutil.copy(os.path.join(EX, 'lj_wall', f), tmp_path / f)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        s = system({'seed': 17})
    finally:
        os.chdir(cwd)
    assert s.N_dim == 2 and s.N_species == 2 and s.N_traj == 10
    assert s.seed == 17 and s.rs == 7.0 and s.neighbor_list_update_frequency == 20
    assert s.max_rc == 5.0   # the wall's rc counts too (pyrdsim.pyx:1063-1064)
    assert s.interactions[4]['pair'] == [1, 1]
    assert np.allclose(s.force_prefacs, 5e-5) and np.allclose(s.noise_prefacs, np.sqrt(1e-4))
    assert s.interaction_partners == {0: {0, 1}, 1: {0, 1}}
    assert len(s.reactions) == 2


def test_system_defaults_and_errors(tmp_path):
    base = dict(parameters_filename=str(tmp_path / 'none.txt'), n_dim=2, l=np.array([10., 10.]),
                dt=0.1, n_steps=10, n_out=10, interactions=[], reactions=[],
                bonded_interactions_filename=os.path.join(EX, 'lj_one_species',
                                                          'bonded_interactions.txt'))
    s = system(dict(base, tau_by_tauD=[1.0, -1.0], seed=2 ** 63 + 5))
    assert s.stride == 1 and np.array_equal(s.periodic, [1, 1])
    assert s.seed == (2 ** 63 + 5) % 9223372036854775807
    assert s.force_prefacs[1] == 0 and s.noise_prefacs[1] == 0
    assert s.rs == 0.0 and s.N_traj == 10
    with pytest.raises(KeyError):
        system({k: v for k, v in base.items() if k != 'dt'})
    with pytest.raises(RuntimeError, match='box edge'):
        system(dict(base, l=np.array([1.0, 2.0, 3.0])))
    with pytest.raises(RuntimeError, match='Periodicity'):
        system(dict(base, periodic=np.array([1.0])))
    with pytest.raises(RuntimeError, match='not recognized'):
        system(dict(base, interactions=[dict(type='morse', pair=[0, 0], u0=1, rc=1)]))
    # periodic = 0 wraps positions but not distances in the reference: refused
    with pytest.raises(RuntimeError, match='holds a 0'):
        system(dict(base, periodic=np.array([1.0, 0.0])))
    from pyrdsim_amd.engine import make_replica
    with pytest.raises(RuntimeError, match='periodic'):
        make_replica(0, 10, [10.0, 10.0], [0.0, 1.0], 0.1, [1.0], [], 1)
    r = make_replica(0, 10, [10.0, 10.0], [2.0, -0.5], 0.1, [1.0], [], 1)
    assert list(r.periodic)[:2] == [1, 0]
    r = make_replica(0, 10, [10.0, 10.0, 8.0], [2.0, -0.5, 1.0], 0.1, [1.0], [], 1)
    assert list(r.periodic) == [1, 0, 1] and list(r.L) == [10.0, 10.0, 8.0]


def test_bond_and_angle_lists_of_molecules(tmp_path):
    """The facade's bond and angle arrays: participants offset by each
    molecule's first index, in the reference's definition order
    (pyrdsim.pyx:1287-1407); angles in rad and per rad^2."""
    ic = tmp_path / 'ic.txt'
    x = np.arange(16, dtype=float).reshape(8, 2) + 1.0
    mol = np.array([[m, k, 0] for m in range(2) for k in range(4)], dtype=np.intc)
    _parse.write_initial_conditions(str(ic), x, np.array([0, 1, 1, 1] * 2), mol)
    base = dict(parameters_filename=str(tmp_path / 'none.txt'), n_dim=2,
                l=np.array([20., 20.]), dt=1e-5, n_steps=10, n_out=10, interactions=[],
                reactions=[], tau_by_tauD=[1.0, 1.0], initial_conditions_filename=str(ic),
                bonded_interactions_filename=os.path.join(EX, 'patchy',
                                                          'bonded_interactions_3.txt'))
    s = system(dict(base, seed=1))
    s.load_initial_conditions_from_file()
    bonds, angles = s._bonds()
    assert bonds[0] == [0, 0, 0, 4, 4, 4] and bonds[1] == [1, 2, 3, 5, 6, 7]
    assert bonds[2] == [1.0] * 6 and bonds[3] == [400.0] * 6
    assert angles[0] == [1, 2, 3, 5, 6, 7]
    assert angles[1] == [0, 0, 0, 4, 4, 4]
    assert angles[2] == [2, 3, 1, 6, 7, 5]
    assert np.allclose(angles[3], 2 * np.pi / 3, rtol=0, atol=1e-15)
    assert np.allclose(angles[4], 0.171934 * (180 / np.pi) ** 2, rtol=1e-15)
