"""Readers for PyRDSim's text input formats.

These restate the behaviour of the reference's parsers so that the same input
files load into the same dictionaries:

* ``preprocess_line``            -- pyrdsim/pyrdsim.pyx:590-614
* ``load_parameters_file``       -- pyrdsim/pyrdsim.pyx:616-821
* ``load_bonded_interactions``   -- pyrdsim/pyrdsim.pyx:823-1034
* ``load_initial_conditions``    -- pyrdsim/pyrdsim.pyx:1790-1913
* ``format_state``               -- pyrdsim/pyrdsim.pyx:1917-1934

One deliberate difference: the reference's ``load_parameters_file`` tests a
flag (``process_input``) that is only assigned once a section keyword has
been seen (pyx:726-747 vs pyx:777), so for section-less files (the
``parameters.txt`` format) the flag is read uninitialised.  The behaviour the
author evidently relied on is implemented here: a file without a section
keyword processes every ``key = value`` line, and a sectioned file raises
``RuntimeError("Parameter outside of section.")`` for a ``key = value`` line
before its first section.
"""

import numpy as np

DEFAULT_COMMENT_CHARS = ('#', ';')


def preprocess_line(line, comment_chars=DEFAULT_COMMENT_CHARS):
    """Lower-case, strip comments, tabs and the newline (pyx:590-614).

    Returns ``(line, line_nospace)``.  For every comment character present,
    everything from its first occurrence on is removed.
    """
    line = line.lower()
    for c in comment_chars:
        if c in line:
            line = line.split(c)[0]
    line = line.replace('\t', ' ').replace('\n', '')
    return line, line.replace(' ', '')


def load_parameters_file(filename, datatypes, int_arrays=(), float_arrays=(),
                         section_keyword='', comment_chars=DEFAULT_COMMENT_CHARS):
    """Parse a ``key = value`` file into a list of dicts, one per section.

    Follows pyx:616-821: names are case-insensitive; ``int_arrays`` and
    ``float_arrays`` values are whitespace-separated lists; other keys are
    cast with ``datatypes[key]`` applied to the space-free value text; an
    unknown key raises ``RuntimeError``; a line with more than one ``=``
    raises ``RuntimeError``.  A section that collected nothing is dropped.
    """
    types_lower = {k.lower(): v for k, v in datatypes.items()}
    ints_lower = [e.lower() for e in int_arrays]
    floats_lower = [e.lower() for e in float_arrays]
    out = []
    current = {}
    sectioned = len(section_keyword) > 0
    inside = not sectioned
    with open(filename, 'r') as f:
        for raw in f:
            line, nospace = preprocess_line(raw, comment_chars)
            if sectioned and section_keyword in nospace:
                inside = True
                if len(current) > 0:
                    out.append(current)
                    current = {}
            if '=' not in nospace:
                continue
            parts_ns = nospace.split('=')
            if len(parts_ns) != 2:
                raise RuntimeError(
                    "Current line is not of the form <name of parameters> = "
                    "<value of parameter>:\n{0}".format(line))
            parts = line.split('=')
            if not inside:
                raise RuntimeError("Parameter outside of section.")
            key = parts_ns[0]
            if key in ints_lower:
                value = np.array([int(v) for v in parts[1].split()], dtype=int)
            elif key in floats_lower:
                value = np.array([float(v) for v in parts[1].split()], dtype=float)
            else:
                try:
                    cast = types_lower[key]
                except KeyError:
                    raise RuntimeError(
                        "Unrecognized entry in parameters file: {0}".format(key))
                value = cast(parts_ns[1])
            current[key] = value
    if len(current) > 0:
        out.append(current)
    return out


def load_bonded_interactions(filename, comment_chars=DEFAULT_COMMENT_CHARS):
    """Parse ``bonded_interactions.txt`` (pyx:823-1034).

    Returns ``(molecule_parameters, molecule_interactions)``, both keyed by
    ``molecule_id``.  Bonds are ``{'type': 'bond', 'participants': [i, j],
    'r0': float, 'u0': float}``; angles are ``{'type': 'angle',
    'participants': [i, j, k], 'theta0': radians, 'u0': k*(180/pi)**2}``
    (pyx:982-989).  Lines inside ``[bonds]``/``[angles]`` with a different
    number of fields are ignored; unparsable ones raise ``RuntimeError``.
    """
    mol_param_types = {'molecule_id': int, 'n_atoms': int}
    molecule_parameters = {}
    molecule_interactions = {}
    cur_params = {}
    cur_inter = []
    started = False
    in_bonds = False
    in_angles = False
    with open(filename, 'r') as f:
        for raw in f:
            line, nospace = preprocess_line(raw, comment_chars)
            if '[molecule]' in nospace:
                started = True
                if len(cur_params) > 0:
                    molecule_parameters[cur_params['molecule_id']] = cur_params
                    molecule_interactions[cur_params['molecule_id']] = cur_inter
                cur_params = {}
                cur_inter = []
                in_bonds = False
                in_angles = False
            elif '[bonds]' in nospace:
                in_bonds, in_angles = True, False
            elif '[angles]' in nospace:
                in_bonds, in_angles = False, True
            if not started:
                continue
            if in_bonds:
                fields = line.split()
                if len(fields) == 4:
                    try:
                        cur_inter.append({
                            'type': 'bond',
                            'participants': [int(fields[0]), int(fields[1])],
                            'r0': float(fields[2]),
                            'u0': float(fields[3])})
                    except ValueError:
                        raise RuntimeError("Could not process line: {0}".format(line))
            elif in_angles:
                fields = line.split()
                if len(fields) == 5:
                    try:
                        cur_inter.append({
                            'type': 'angle',
                            'participants': [int(fields[0]), int(fields[1]),
                                             int(fields[2])],
                            'theta0': np.pi * float(fields[3]) / 180.,
                            'u0': float(fields[4]) * (180. / np.pi) ** 2})
                    except ValueError:
                        raise RuntimeError("Could not process line: {0}".format(line))
            else:
                for name, cast in mol_param_types.items():
                    if name in nospace:
                        if name in cur_params:
                            raise RuntimeError(
                                "Repeated molecule parameter {0} in file {1}".format(
                                    name, filename))
                        cur_params[name] = cast(nospace.split('=')[1])
    if len(cur_params) > 0:
        molecule_parameters[cur_params['molecule_id']] = cur_params
        molecule_interactions[cur_params['molecule_id']] = cur_inter
    return molecule_parameters, molecule_interactions


def load_initial_conditions(filename, n_dim, n_species,
                            comment_chars=DEFAULT_COMMENT_CHARS):
    """Parse ``initial_conditions.txt`` (pyx:1790-1913).

    Each line: molecule number, index within molecule, molecule id, particle
    type, then ``n_dim`` coordinates.  The first data line fixes the column
    count (``RuntimeError`` if it disagrees with ``n_dim``); later lines with
    another column count are skipped.  A molecule starts where the index
    within the molecule is 0.

    Returns ``dict(x, molecule_ids, particle_types, first_index_of_molecule,
    n_particles, n_molecules)`` with the reference's dtypes (float64 and
    int32).
    """
    mol_ids, types, xs, firsts = [], [], [], []
    n = 0
    n_mol = 0
    max_type = 0
    ncols = None
    with open(filename, 'r') as f:
        for line in f:
            for c in comment_chars:
                if c in line:
                    line = line.split(c)[0]
            if len(line) == 0:
                continue
            fields = line.replace('\t', ' ').split()
            if n == 0:
                ncols = len(fields)
                if ncols - 4 != n_dim:
                    raise RuntimeError(
                        "Dimension of initial conditions ({0}) inconsistent with "
                        "N_dim = {1}".format(ncols - 4, n_dim))
            elif len(fields) != ncols:
                continue
            mol_ids.append([int(fields[0]), int(fields[1]), int(fields[2])])
            t = int(fields[3])
            types.append(t)
            xs.append([float(fields[4 + k]) for k in range(n_dim)])
            if int(fields[1]) == 0:
                firsts.append(n)
                n_mol += 1
            if t > max_type:
                max_type = t
            n += 1
    if n == 0:
        raise RuntimeError("Initial condition file contains no particles")
    if n_species < max_type + 1:
        raise RuntimeError(
            "Initial conditions contains {0} particle types, but only {1} "
            "friction coefficients have been defined. Please set "
            "particle_friction parameters for all particles.".format(
                max_type + 1, n_species))
    return dict(x=np.array(xs, dtype=np.float64),
                molecule_ids=np.array(mol_ids, dtype=np.intc).reshape(n, 3),
                particle_types=np.array(types, dtype=np.intc),
                first_index_of_molecule=np.array(firsts, dtype=np.intc),
                n_particles=n, n_molecules=n_mol)


def format_state(molecule_ids, particle_types, x):
    """Text of ``write_current_state_to_text_file`` (pyx:1917-1934)."""
    out = []
    n, d = x.shape
    for i in range(n):
        row = ['{0:d}   '.format(int(molecule_ids[i, j])) for j in range(3)]
        row.append('{0:d}   '.format(int(particle_types[i])))
        row.extend('{0:3.5f}   '.format(float(x[i, k])) for k in range(d))
        out.append(''.join(row) + '\n')
    return ''.join(out)


def write_initial_conditions(filename, x, particle_types, molecule_ids=None):
    """Write an ``initial_conditions.txt`` at full precision (``%.17g``).

    The reference's own writer keeps 5 decimals (pyx:1933); synthetic
    workloads are written losslessly so that engine and oracle start from
    bit-identical positions.  With ``molecule_ids`` omitted every particle is
    its own one-atom molecule of id 0.
    """
    n = x.shape[0]
    if molecule_ids is None:
        molecule_ids = np.zeros((n, 3), dtype=np.intc)
        molecule_ids[:, 0] = np.arange(n)
    with open(filename, 'w') as f:
        for i in range(n):
            f.write('{0:d} {1:d} {2:d} {3:d} '.format(
                int(molecule_ids[i, 0]), int(molecule_ids[i, 1]),
                int(molecule_ids[i, 2]), int(particle_types[i])))
            f.write(' '.join('%.17g' % float(v) for v in x[i]))
            f.write('\n')
