"""Trajectory file with the reference's layout (pyrdsim.pyx:1703-1786).

HDF5 through h5py when it is importable; otherwise an ``.npz`` file holding the
same datasets (written next to the requested name with ``.npz`` appended).
Semantics kept from the reference: the first flush writes only the current
frame (pyrdsim.pyx:1733-1736), later flushes append the whole ring buffer of
N_traj frames (pyrdsim.pyx:1775-1786).
"""

import os

import numpy as np

try:
    import h5py
except ImportError:  # not installed in this environment
    h5py = None


def _npz(path):
    return path + '.npz'


def save(path, s):
    """save_trajectory_to_file (pyrdsim.pyx:1725-1766)."""
    data = dict(
        trajectory=np.asarray(s.x, dtype=np.float64).reshape(1, s.N_particles, s.N_dim),
        trajectory_types=np.asarray(s.particle_types).reshape(1, s.N_particles),
        molecule_ids=np.asarray(s.molecule_ids),
        first_index_of_molecule=np.asarray(s.first_index_of_molecule),
        N_steps_completed=np.int64(s.N_steps_completed),
        N_particles=np.int64(s.N_particles),
        N_molecules=np.int64(s.N_molecules),
        dt_trajectory=np.float64(s.dt * s.stride),
        N_species=np.int64(s.N_species),
        stride=np.int64(s.stride),
        L=np.asarray(s.L, dtype=np.float64))
    if h5py is not None:
        with h5py.File(path, 'w') as hf:
            for k in ('trajectory', 'trajectory_types'):
                hf.create_dataset(k, data=data[k], maxshape=(None,) + data[k].shape[1:],
                                  chunks=True)
            for k in ('molecule_ids', 'first_index_of_molecule'):
                hf.create_dataset(k, data=data[k], chunks=True)
            for k in ('N_steps_completed', 'N_particles', 'N_molecules', 'N_species', 'stride'):
                hf.create_dataset(k, data=data[k], dtype=int)
            hf.create_dataset('dt_trajectory', data=data['dt_trajectory'], dtype=float)
            hf.create_dataset('L', data=data['L'])
    else:
        np.savez(_npz(path), **data)


def append(path, s):
    """append_trajectory_to_file (pyrdsim.pyx:1770-1786)."""
    if h5py is not None:
        with h5py.File(path, 'a') as hf:
            for k, v in (('trajectory', s.trajectory), ('trajectory_types', s.trajectory_types)):
                hf[k].resize(hf[k].shape[0] + s.N_traj, axis=0)
                hf[k][-s.N_traj:] = v
            hf['N_steps_completed'][...] = s.N_steps_completed
        return
    f = _npz(path)
    with np.load(f, allow_pickle=False) as z:
        data = {k: z[k] for k in z.files}
    data['trajectory'] = np.concatenate([data['trajectory'], s.trajectory], axis=0)
    data['trajectory_types'] = np.concatenate(
        [data['trajectory_types'], np.asarray(s.trajectory_types)], axis=0)
    data['N_steps_completed'] = np.int64(s.N_steps_completed)
    np.savez(f, **data)


def load(path):
    """load_trajectory_from_file (pyrdsim.pyx:1703-1723); None if absent."""
    if h5py is not None:
        if not os.path.exists(path):
            return None
        with h5py.File(path, 'r') as hf:
            return dict(molecule_ids=hf['molecule_ids'][()],
                        first_index_of_molecule=hf['first_index_of_molecule'][()],
                        trajectory_types=hf['trajectory_types'][-1:],
                        trajectory=hf['trajectory'][-1:],
                        N_steps_completed=hf['N_steps_completed'][()],
                        N_particles=hf['N_particles'][()],
                        N_molecules=hf['N_molecules'][()])
    f = _npz(path)
    if not os.path.exists(f):
        return None
    with np.load(f, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}
