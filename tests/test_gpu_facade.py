"""End to end through the drop-in facade `pyrdsim_amd.system` on the GPU:
reference-format input files in a directory, `simulate()`, the result dict,
the ring buffer of frames, the trajectory file, resume with append=True, the
state text file and the reference's error behaviour (pyrdsim.pyx:2329-2456,
1917-1934)."""

import os

import numpy as np
import pytest

from pyrdsim_amd import _parse, trajectory, workloads
from pyrdsim_amd.system import system

pytestmark = pytest.mark.gpu


def _write_inputs(d, s, n_steps, stride, n_out, extra=''):
    L = s['L']
    per = s['periodic']
    with open(os.path.join(d, 'parameters.txt'), 'w') as f:
        f.write('N_dim = %d\nL = %s\nperiodic = %s\n'
                % (len(L), ' '.join(repr(float(v)) for v in L),
                   ' '.join('%d' % int(v) for v in per)))
        f.write('dt = %r\nN_steps = %d\nstride = %d\nN_out = %d\n'
                % (float(s['dt']), n_steps, stride, n_out))
        f.write('tau_by_tauD = %s\n' % ' '.join('1.0' for _ in s['model']['tau_by_tauD']))
        f.write(extra)
    with open(os.path.join(d, 'nonbonded_interactions.txt'), 'w') as f:
        for it in s['model']['interactions']:
            f.write('[ interaction ]\n')
            for k, v in it.items():
                if k == 'pair':
                    v = '%d %d' % tuple(v)
                f.write('%s = %s\n' % (k, v))
    with open(os.path.join(d, 'reactions.txt'), 'w') as f:
        for rx in s['model']['reactions']:
            f.write('[ reaction ]\n')
            for k, v in rx.items():
                f.write('%s = %s\n' % (k, v))
    with open(os.path.join(d, 'bonded_interactions.txt'), 'w') as f:
        f.write('[ molecule ]\nmolecule_id = 0\nN_atoms = 1\n[ bonds ]\n[ angles ]\n')
    _parse.write_initial_conditions(os.path.join(d, 'initial_conditions.txt'), s['x'],
                                    s['types'])


@pytest.fixture
def workdir(tmp_path):
    cwd = os.getcwd()
    os.chdir(tmp_path)
    yield tmp_path
    os.chdir(cwd)


@pytest.mark.parametrize('name', ['lj1', 'wall'])
def test_simulate_result_and_frames(gpu, workdir, name):
    s = workloads.lj1() if name == 'lj1' else workloads.wall(n=2048)
    _write_inputs(workdir, s, n_steps=60, stride=5, n_out=30)
    sim = system({'seed': 123})
    out = sim.simulate(verbose=False)
    assert set(out) == {'x', 'particle_types', 'trajectory', 'trajectory_types',
                        'molecule_ids', 'first_index_of_molecule'}
    n = len(s['x'])
    assert out['x'].shape == (n, 2) and out['x'].dtype == np.float64
    assert out['particle_types'].dtype == np.intc
    assert out['trajectory'].shape == (6, n, 2)
    assert out['trajectory_types'].dtype == np.int64
    # ring buffer of N_traj = 6 frames: step 60 is the last frame written
    assert np.array_equal(out['trajectory'][-1], out['x'])
    assert np.isfinite(out['x']).all()
    assert (out['x'] >= 0).all() and (out['x'] <= s['L']).all()
    assert sim.N_steps_completed == 60
    # two flushes happened: the first wrote one frame, the second six
    data = trajectory.load('./trajectory.h5')
    assert data is not None
    # the same run again (same seed) gives the same trajectory: Philox streams
    # are counter-based
    sim2 = system({'seed': 123})
    out2 = sim2.simulate(verbose=False)
    d = out2['x'] - out['x']
    d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-9
    # engine options in the constructor dict: Verlet lists and compensated
    # positions give the same run to rounding
    for opts in ({'skin': 0.4}, {'precision': 'compensated'}):
        sim3 = system(dict(seed=123, **opts))
        out3 = sim3.simulate(verbose=False)
        d = out3['x'] - out['x']
        d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
        assert np.abs(d).max() < 1e-6, opts
        assert np.array_equal(out3['particle_types'], out['particle_types']), opts


def test_resume_with_append(gpu, workdir):
    s = workloads.lj1(n=512)
    _write_inputs(workdir, s, n_steps=40, stride=10, n_out=20)
    a = system({'seed': 5})
    a.simulate(verbose=False)
    b = system({'seed': 5, 'n_steps': 80})
    out = b.simulate(verbose=False, append=True)
    assert b.N_steps_completed == 80
    assert out['trajectory'].shape == (2, 512, 2)


def test_resume_is_continuous(gpu, workdir):
    """40 steps, then 40 more resumed with append=True from the trajectory
    file, equal one uninterrupted 80-step run: FP64 positions within 1e-12
    and identical types.  The resumed run starts from the last frame and the
    step count the file records (pyrdsim.pyx:1703-1723, 2379, 2410), and the
    Philox counters continue from that step."""
    s = workloads.wall(n=1024)
    s['types'][::3] = 1
    _write_inputs(workdir, s, n_steps=40, stride=1, n_out=40)
    system({'seed': 7}).simulate(verbose=False)
    b = system({'seed': 7, 'n_steps': 80})
    out_b = b.simulate(verbose=False, append=True)
    assert b.N_steps_completed == 80
    c = system({'seed': 7, 'n_steps': 80})
    out_c = c.simulate(verbose=False)
    assert c.N_steps_completed == 80
    d = out_b['x'] - out_c['x']
    d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-12, float(np.abs(d).max())
    assert np.array_equal(out_b['particle_types'], out_c['particle_types'])
    # the last frames of the two runs' ring buffers are the same step too
    d = out_b['trajectory'][-1] - out_c['trajectory'][-1]
    d -= s['L'] * np.round(d / s['L']) * (s['periodic'] > 0)
    assert np.abs(d).max() < 1e-12
    # something happened in between: reactions and motion
    assert (out_c['particle_types'] != s['types']).any()


def test_state_file_and_errors(gpu, workdir):
    s = workloads.lj1(n=256)
    _write_inputs(workdir, s, n_steps=10, stride=5, n_out=10)
    sim = system({'seed': 1})
    sim.simulate(verbose=False)
    sim.write_current_state_to_text_file('state.txt')
    lines = open('state.txt').read().splitlines()
    assert len(lines) == 256 and len(lines[0].split()) == 6
    # N_out < stride: N_traj = 0 and the first output step divides by zero
    _write_inputs(workdir, s, n_steps=10, stride=5, n_out=3)
    with pytest.raises(ZeroDivisionError):
        system({'seed': 1}).simulate(verbose=False)
    # a time step that throws particles out of a reflecting box
    w = workloads.wall(n=1024)
    _write_inputs(workdir, dict(w, dt=50.0), n_steps=5, stride=1, n_out=5)
    with pytest.raises(RuntimeError, match='out of bounds'):
        system({'seed': 1}).simulate(verbose=False)


def test_random_force_off_is_deterministic_descent(gpu, workdir):
    """random_force=False: no noise (the reference still draws it and multiplies
    by 0): pure drift, independent of the seed."""
    s = workloads.lj1(n=256)
    _write_inputs(workdir, s, n_steps=20, stride=20, n_out=20)
    x1 = system({'seed': 1}).simulate(verbose=False, random_force=False)['x']
    x2 = system({'seed': 99}).simulate(verbose=False, random_force=False)['x']
    # equal up to the order of the force summation within a cell
    assert np.abs(x1 - x2).max() < 1e-12
