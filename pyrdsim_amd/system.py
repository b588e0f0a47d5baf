"""Drop-in for ``pyrdsim.system``: the reference's simulation object, with its
step loop running on the MI355X engine.

Same constructor dict, same four input files and initial-condition file, same
``simulate(verbose, random_force, append)`` result dict and the same error
types as pyrdsim/pyrdsim.pyx.  Differences, all deliberate:

* the random stream is Philox4x32-10 keyed by ``seed`` (counter = particle,
  step), not one sequential ``std::mt19937``; it is statistically equivalent
  and makes a trajectory independent of launch order;
* every pair within its cutoff is found every step (cell list rebuilt each
  step), where the reference uses a Verlet list refreshed every
  ``neighbor_list_update_frequency`` steps, type-gated at refresh time;
  ``rs``, ``cell_length_ratio`` and ``neighbor_list_update_frequency`` are
  accepted and kept but do not change results;
* the trajectory file is HDF5 through h5py when it is importable, else an
  ``.npz`` file with the same datasets.

Engine options accepted in the constructor dict (not in the files, whose
parser keeps the reference's key table): ``precision`` ('fp64' default,
'mixed' = FP32 pair forces with FP64 positions, or 'compensated' = FP32 pair
forces with float-pair positions), ``reaction_probability`` ('linear' = the
reference's p = dt*k, or 'exp' = 1 - exp(-dt*k)), ``graph_steps`` (steps per
captured hipGraph, default 32), ``skin`` (0 default: cell search every step;
> 0: Verlet lists rebuilt after a particle moved more than skin/2).
"""

import os
import time

import numpy as np

from . import _parse
from .engine import Engine, build_model, make_replica

MAX_LONG = 9223372036854775807


class system:
    """The simulation object (pyrdsim.pyx:111-2456)."""

    # -- 1. initialization + processing of parameters ----------------------
    def __init__(self, parameters={}):
        """pyrdsim.pyx:265-304."""
        self.set_default_parameters()
        self.set_system_parameters()
        try:
            self.parameters_filename = parameters['parameters_filename']
        except KeyError:
            pass
        try:
            loaded = self.load_parameters_file(
                filename=self.parameters_filename,
                section_keyword=self.parameters_section_keyword,
                datatypes=self.parameters_datatypes,
                int_arrays=self.parameters_int_arrays,
                float_arrays=self.parameters_float_arrays)
            loaded = loaded[0] if len(loaded) > 0 else {}
        except FileNotFoundError:
            loaded = {}
        for key, value in parameters.items():
            loaded[key.lower()] = value
        self.set_parameters(parameters=loaded)

    def set_default_parameters(self):
        """pyrdsim.pyx:307-335."""
        self.parameters_filename = 'parameters.txt'
        self.bint_filename = './bonded_interactions.txt'
        self.nbint_filename = './nonbonded_interactions.txt'
        self.reactions_filename = './reactions.txt'
        self.initial_conditions_filename = './initial_conditions.txt'
        self.trajectory_filename = './trajectory.h5'
        self.max_rc = 0.
        self.interaction_partners = {}
        self.N_species = 0
        self.stride = 1
        self.neighbor_list_update_frequency = 10
        self.comment_chars = set(['#', ';'])
        self.neighbor_list_rc_to_cell_length_ratio = 2
        self.N_reactions = 0
        self.reaction_rates = np.array([[]], dtype=np.float64)
        self.total_reaction_rates = np.array([], dtype=np.float64)
        self.precision = 'fp64'
        self.reaction_probability = 'linear'
        self.graph_steps = 32
        self.skin = 0.0

    def set_system_parameters(self):
        """pyrdsim.pyx:338-400."""
        self.angle_pi_approximation_threshold = 0.1
        self.parameters_section_keyword = ''
        self.parameters_datatypes = {
            'N_dim': int, 'dt': float, 'N_steps': int, 'stride': int, 'N_out': int,
            'bonded_interactions_filename': str, 'nonbonded_interactions_filename': str,
            'reactions_filename': str, 'initial_conditions_filename': str,
            'trajectory_filename': str, 'neighbor_list_update_frequency': int,
            'rs': float, 'cell_length_ratio': int}
        self.parameters_int_arrays = []
        self.parameters_float_arrays = ['L', 'periodic', 'tau_by_tauD']
        self.nbint_section_keyword = '[interaction]'
        self.nbint_datatypes = {'type': str, 'u0': float, 'r0': float, 'rc': float,
                                'dr': float, 'a': float, 'n': int, 'particle_type': int,
                                'direction': int, 'position': float}
        self.nbint_int_arrays = ['pair']
        self.nbint_float_arrays = []
        self.reactions_section_keyword = '[reaction]'
        self.reactions_datatypes = {'type': str, 'rate': float, 'reactant': int,
                                    'product': int, 'r0': float, 'catalyst': int,
                                    'rate_decrease': float, 'inhibitor': int,
                                    'inhibitor2': int}
        self.reactions_int_arrays = []
        self.reactions_float_arrays = []

    def initialize_random_number_generator(self, supplied_seed=-1):
        """pyrdsim.pyx:87-108: seed = supplied % (2^63 - 1), or pid + time."""
        if supplied_seed < 0:
            self.seed = (abs(os.getpid()) + int(time.time() * 1000)) % MAX_LONG
        else:
            self.seed = supplied_seed % MAX_LONG

    def set_parameters(self, parameters):
        """pyrdsim.pyx:406-588."""
        try:
            self.comment_chars = parameters['comment_chars']
        except KeyError:
            pass
        try:
            self.initialize_random_number_generator(supplied_seed=parameters['seed'])
        except KeyError:
            self.initialize_random_number_generator()
        self.N_dim = parameters['n_dim']
        self.L = parameters['l']
        if len(self.L) != self.N_dim:
            raise RuntimeError(
                "Number of box edge lenghts {0}".format(len(self.L))
                + " is inconsistent with N_dim = {0}.".format(self.N_dim)
                + " Please provide one length per dimension.")
        try:
            self.periodic = parameters['periodic']
        except KeyError:
            self.periodic = np.ones(self.N_dim, dtype=np.float64)
        if len(self.periodic) != self.N_dim:
            raise RuntimeError(
                "Periodicity definition {0}".format(self.periodic)
                + " is inconsistent with n_dim = {0}. ".format(self.N_dim)
                + "Please provide one periodicity value per spatial dimension.")
        if any(not (float(p) > 0 or float(p) < 0) for p in self.periodic):
            # the reference wraps positions unless periodic < 0
            # (pyrdsim.pyx:2274) but takes minimum images only where
            # periodic > 0 (pyrdsim.pyx:1629): a 0 mixes the two
            raise RuntimeError(
                "Periodicity definition {0} holds a 0 (or NaN): use a positive ".format(
                    self.periodic)
                + "value for a periodic axis and a negative one for a reflecting axis.")
        self.dt = parameters['dt']
        self.N_steps = parameters['n_steps']
        self.N_out = parameters['n_out']
        try:
            self.stride = parameters['stride']
        except KeyError:
            pass
        self.N_traj = self.N_out // self.stride
        try:
            self.tau_by_tauD = np.array(parameters['tau_by_taud'], dtype=np.float64)
        except KeyError:
            if self.N_species == 0:
                self.tau_by_tauD = np.array([1.], dtype=np.float64)
        self.N_species = len(self.tau_by_tauD)
        mask = (self.tau_by_tauD < 0)
        self.noise_prefacs = np.ones_like(self.tau_by_tauD, dtype=np.float64)
        self.force_prefacs = np.ones_like(self.tau_by_tauD, dtype=np.float64)
        self.noise_prefacs[mask] = 0.
        self.noise_prefacs[~mask] = np.sqrt(2. * self.dt * self.tau_by_tauD[~mask])
        self.force_prefacs[mask] = 0.
        self.force_prefacs[~mask] = self.dt * self.tau_by_tauD[~mask]
        for key, attr in (('bonded_interactions_filename', 'bint_filename'),
                          ('nonbonded_interactions_filename', 'nbint_filename'),
                          ('reactions_filename', 'reactions_filename'),
                          ('initial_conditions_filename', 'initial_conditions_filename'),
                          ('trajectory_filename', 'trajectory_filename')):
            if key in parameters:
                setattr(self, attr, parameters[key])
        try:
            self.interactions = parameters['interactions']
        except KeyError:
            self.interactions = self.load_parameters_file(
                filename=self.nbint_filename, section_keyword=self.nbint_section_keyword,
                datatypes=self.nbint_datatypes, int_arrays=self.nbint_int_arrays,
                float_arrays=self.nbint_float_arrays)
        self.create_nonbonded_interaction_functions()
        self.add_interaction_partners_for_nonbonded_interactions()
        try:
            self.reactions = parameters['reactions']
        except KeyError:
            self.reactions = self.load_parameters_file(
                filename=self.reactions_filename,
                section_keyword=self.reactions_section_keyword,
                datatypes=self.reactions_datatypes, int_arrays=self.reactions_int_arrays,
                float_arrays=self.reactions_float_arrays)
        self.N_reactions = len(self.reactions)
        self.add_interaction_partners_for_reactions()
        self.load_bonded_interactions()
        if 'cell_length_ratio' in parameters:
            self.neighbor_list_rc_to_cell_length_ratio = parameters['cell_length_ratio']
        self.rs = parameters['rs'] if 'rs' in parameters else 2 * self.max_rc
        if 'neighbor_list_update_frequency' in parameters:
            self.neighbor_list_update_frequency = parameters['neighbor_list_update_frequency']
        for key in ('precision', 'reaction_probability', 'graph_steps', 'skin'):
            if key in parameters:
                setattr(self, key, parameters[key])

    def preprocess_line(self, line):
        """pyrdsim.pyx:590-614."""
        return _parse.preprocess_line(line, tuple(self.comment_chars))

    def load_parameters_file(self, filename, datatypes, int_arrays=[], float_arrays=[],
                             section_keyword=''):
        """pyrdsim.pyx:616-821."""
        return _parse.load_parameters_file(filename, datatypes, int_arrays, float_arrays,
                                           section_keyword, tuple(self.comment_chars))

    def load_bonded_interactions(self):
        """pyrdsim.pyx:823-1034."""
        self.molecule_parameters, self.molecule_interactions = \
            _parse.load_bonded_interactions(self.bint_filename, tuple(self.comment_chars))

    def create_nonbonded_interaction_functions(self):
        """pyrdsim.pyx:1037-1154: normalises the dicts, sets max_rc, and checks
        the interaction types (the force laws themselves run in the engine)."""
        for i, interaction in enumerate(self.interactions):
            interaction = {k.lower(): v for k, v in interaction.items()}
            interaction['type'] = interaction['type'].lower()
            t = interaction['type']
            if t == 'piecewise_linear':
                interaction['rc'] = interaction['r0']
            if interaction['rc'] > self.max_rc:
                self.max_rc = interaction['rc']
            if t not in ('lj', 'tanh', 'yukawa', 'harmonic', 'piecewise_linear', 'wall'):
                raise RuntimeError("Interaction {0} not recognized".format(t))
            if t == 'wall':
                interaction['pair'] = [interaction['particle_type'], interaction['direction']]
            self.interactions[i] = interaction

    def add_interaction_partners_for_nonbonded_interactions(self):
        """pyrdsim.pyx:1157-1190."""
        for interaction in self.interactions:
            if interaction['type'] != 'wall':
                self.add_pair_to_interaction_partner_list(*interaction['pair'])

    def add_interaction_partners_for_reactions(self):
        """pyrdsim.pyx:1192-1234."""
        for reaction in self.reactions:
            if reaction['type'] == 'locally_catalyzed':
                self.add_pair_to_interaction_partner_list(int(reaction['reactant']),
                                                          int(reaction['catalyst']))
            elif reaction['type'] == 'locally_inhibited':
                self.add_pair_to_interaction_partner_list(int(reaction['reactant']),
                                                          int(reaction['inhibitor']))
                try:
                    self.add_pair_to_interaction_partner_list(int(reaction['reactant']),
                                                              int(reaction['inhibitor2']))
                except KeyError:
                    pass

    def add_pair_to_interaction_partner_list(self, i0, i1):
        """pyrdsim.pyx:1237-1263."""
        self.interaction_partners.setdefault(i0, set()).add(i1)
        self.interaction_partners.setdefault(i1, set()).add(i0)

    # -- 3. IO -------------------------------------------------------------
    def load_initial_conditions_from_file(self):
        """pyrdsim.pyx:1790-1913."""
        ic = _parse.load_initial_conditions(self.initial_conditions_filename, self.N_dim,
                                            self.N_species, tuple(self.comment_chars))
        self.N_particles = ic['n_particles']
        self.N_molecules = ic['n_molecules']
        self.x = ic['x']
        self.molecule_ids = ic['molecule_ids']
        self.particle_types = ic['particle_types']
        self.first_index_of_molecule = ic['first_index_of_molecule']

    def write_current_state_to_text_file(self, filename='state.txt'):
        """pyrdsim.pyx:1917-1934."""
        with open(filename, 'w') as f:
            f.write(_parse.format_state(self.molecule_ids, self.particle_types, self.x))

    def create_initial_conditions(self, N_particles, particle_types,
                                  reference_configurations=None,
                                  min_distance_threshold=None,
                                  min_distance_threshold_sim=None,
                                  seed=None, relax_steps=None, relax_random_force=False):
        """pyrdsim.pyx:2634-2828, seeded.  N_particles = {molecule_id: count},
        particle_types = {molecule_id: [type of each atom]},
        reference_configurations = {molecule_id: filename} (optional; else the
        geometry minimising the molecule's bonded energy).  Molecules are
        placed rigidly at random with every atom at least
        min_distance_threshold (default rs/4) from the others, then relaxed
        once on the GPU for relax_steps (default N_steps) steps without noise
        (relax_random_force=False); a relaxed state whose smallest distance
        between molecules is below min_distance_threshold_sim (default rs/10)
        is placed again.  Returns the reference's output array (molecule
        number, index in molecule, molecule id, type, positions) and writes
        initial_conditions_filename.  seed=None uses the system's seed."""
        from . import initial_conditions as icg
        if self.N_dim not in (2, 3):
            raise NotImplementedError('the MI355X engine integrates two- and three-dimensional '
                                      'systems')
        rng = np.random.Generator(np.random.PCG64(self.seed if seed is None else seed))
        if min_distance_threshold is None:
            min_distance_threshold = self.rs / 4.
        if min_distance_threshold_sim is None:
            min_distance_threshold_sim = self.rs / 10.
        n_atoms = {i: self.molecule_parameters[i]['n_atoms'] for i in self.molecule_interactions}
        refs = {}
        if not reference_configurations:
            for i in self.molecule_interactions:
                refs[i] = icg.reference_molecule(self.molecule_interactions[i], n_atoms[i],
                                                 self.N_dim, rng)
        else:
            for i, filename in reference_configurations.items():
                refs[i] = np.atleast_2d(np.genfromtxt(filename))
                if len(refs[i]) != n_atoms[i]:
                    raise RuntimeError(("Reference configuration for molecule {0}"
                                        " has {1} atoms, but interaction parameters for that"
                                        " molecule demand {2} atoms").format(
                                            i, len(refs[i]), n_atoms[i]))
        L = np.asarray(self.L, dtype=float)
        steps = self.N_steps if relax_steps is None else int(relax_steps)
        for _ in range(10):
            out = icg.place_molecules(N_particles, particle_types, refs, L, self.periodic,
                                      min_distance_threshold, rng)
            self._set_molecules(out)
            if steps > 0:
                eng = self._engine(relax_random_force)
                try:
                    eng.set_state(self.x, self.particle_types,
                                  self.molecule_ids[:, 0] if eng.has_molecules else None)
                    eng.run(steps)
                    x, _ = eng.state()
                except RuntimeError:
                    continue
                finally:
                    eng.close()
                if np.isnan(x).any():
                    continue
                d = icg.min_intermolecular_distance(x, self.molecule_ids[:, 0], L,
                                                    self.periodic, min_distance_threshold_sim)
                if d < min_distance_threshold_sim:
                    continue
                out[:, 4:] = x
                self.x = x
            self.write_current_state_to_text_file(filename=self.initial_conditions_filename)
            return out
        raise RuntimeError('Could not create an initial condition: relaxed states kept '
                           'molecules closer than {0}'.format(min_distance_threshold_sim))

    def _set_molecules(self, out):
        """State arrays of an output array of create_initial_conditions."""
        self.molecule_ids = np.array(out[:, :3], dtype=np.intc)
        self.particle_types = np.array(out[:, 3], dtype=np.intc)
        self.x = np.array(out[:, 4:], dtype=np.float64)
        self.N_particles = len(out)
        first = np.flatnonzero(self.molecule_ids[:, 1] == 0)
        self.N_molecules = len(first)
        self.first_index_of_molecule = first.astype(np.intc)

    def _bonds(self):
        """Bond and angle lists from the molecule definitions, in the
        reference's definition order (pyrdsim.pyx:1287-1407): bonds
        (i, j, r0, u0) and angles (i1, i2, i3, theta0, k0)."""
        bonds = ([], [], [], [])
        angles = ([], [], [], [], [])
        for i in range(self.N_molecules):
            j = int(self.first_index_of_molecule[i])
            mol_id = int(self.molecule_ids[j, 2])
            for it in self.molecule_interactions.get(mol_id, []):
                p = [q + j for q in it['participants']]
                if it['type'] == 'bond':
                    for col, v in zip(bonds, (p[0], p[1], it['r0'], it['u0'])):
                        col.append(v)
                elif it['type'] == 'angle':
                    for col, v in zip(angles, (p[0], p[1], p[2], it['theta0'], it['u0'])):
                        col.append(v)
        return bonds, angles

    def _trajectory_io(self):
        from . import trajectory
        return trajectory

    # -- 4. simulation -----------------------------------------------------
    def _engine(self, random_force):
        if self.N_dim not in (2, 3):
            raise NotImplementedError('the MI355X engine integrates two- and three-dimensional '
                                      'systems')
        model, rates = build_model(self.interactions, self.reactions, self.N_species)
        rep = make_replica(0, self.N_particles, self.L, self.periodic, self.dt,
                           self.tau_by_tauD, rates, self.seed)
        bonds, angles = self._bonds()
        has_mol = self.N_molecules < self.N_particles
        eng = Engine([model], [rep], precision=self.precision, rng='philox',
                     reaction_prob=self.reaction_probability,
                     bonds=bonds if len(bonds[0]) else None, has_molecules=has_mol,
                     graph_steps=int(self.graph_steps), skin=float(self.skin),
                     angles=angles if len(angles[0]) else None,
                     angle_threshold=self.angle_pi_approximation_threshold,
                     n_dim=self.N_dim)
        eng.set_random_force(random_force)
        return eng

    def simulate(self, verbose=True, random_force=True, append=False):
        """pyrdsim.pyx:2329-2456."""
        traj_io = self._trajectory_io()
        if verbose:
            print("Loading and processing initial conditions..", end='\r')
        if append:
            state = traj_io.load(self.trajectory_filename)
            if state is None:
                self.N_steps_completed = 0
                self.load_initial_conditions_from_file()
            else:
                self.molecule_ids = state['molecule_ids']
                self.first_index_of_molecule = state['first_index_of_molecule']
                self.particle_types = np.array(state['trajectory_types'][-1], dtype=np.intc)
                self.x = np.array(state['trajectory'][-1], dtype=np.float64)
                self.N_steps_completed = int(state['N_steps_completed'])
                self.N_particles = int(state['N_particles'])
                self.N_molecules = int(state['N_molecules'])
        else:
            self.load_initial_conditions_from_file()
            self.N_steps_completed = 0
        N_traj = self.N_traj
        self.trajectory = np.zeros([N_traj, self.N_particles, self.N_dim], dtype=np.float64)
        self.trajectory_types = np.zeros([N_traj, self.N_particles], dtype=np.intc)
        self.a = np.zeros([self.N_particles, self.N_dim], dtype=np.float64)
        self.reaction_rates = np.zeros([self.N_reactions, self.N_particles], dtype=np.float64)
        self.total_reaction_rates = np.zeros(self.N_particles, dtype=np.float64)
        eng = self._engine(random_force)
        eng.set_state(self.x, self.particle_types,
                      self.molecule_ids[:, 0] if eng.has_molecules else None,
                      step0=self.N_steps_completed)
        if verbose:
            print("Loading and processing initial conditions..done", end='\n')
        stride = self.stride
        start = time.time()
        l = self.N_steps_completed
        try:
            while l < self.N_steps:
                # run to the next output step ((l+1) % stride == 0) or the end
                nxt = ((l // stride) + 1) * stride
                n = min(nxt, self.N_steps) - l
                eng.run(n)
                l += n
                if l % stride == 0:
                    x, t = eng.state()
                    m = ((l - 1) // stride) % N_traj
                    self.trajectory[m] = x
                    self.trajectory_types[m] = t
                    if m == N_traj - 1:
                        self.N_steps_completed += N_traj * stride
                        if self.N_steps_completed == N_traj * stride:
                            self.x, self.particle_types = x, t
                            traj_io.save(self.trajectory_filename, self)
                        else:
                            traj_io.append(self.trajectory_filename, self)
                else:
                    eng.sync()
                if verbose:
                    el = time.time() - start
                    print("Running simulation.. Progress: {0}%, elapsed time: {1:.0f} s".format(
                        int(l / self.N_steps * 100.), el), end='\r')
            x, t = eng.state()
        finally:
            eng.close()
        self.x = x
        self.particle_types = t.astype(np.intc)
        return {
            'x': self.x,
            'particle_types': self.particle_types,
            'trajectory': np.array(self.trajectory),
            'trajectory_types': np.array(self.trajectory_types, dtype=int),
            'molecule_ids': self.molecule_ids,
            'first_index_of_molecule': self.first_index_of_molecule,
        }
