"""ctypes binding of ``librds_hip.so`` (declared in include/rds.h).

The library is the only compute path: importing this module fails loudly when
the library has not been built.
"""

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'librds_hip.so')

MAX_TYPES = 8
MAX_TERMS = 32
MAX_WALLS = 8
MAX_REACTIONS = 8
STATIC_LABELS_MAX = 8388608   # RDS_STATIC_LABELS_MAX
ABI_VERSION = 4

OK, E_ARG, E_HIP, E_BOUNDS, E_STATE, E_STREAM, E_ANGLE = 0, -1, -2, -3, -4, -5, -6
FP64, MIXED, COMPENSATED = 0, 1, 2
RNG_PHILOX, RNG_INJECTED = 0, 1
PROB_LINEAR, PROB_EXP = 0, 1
LJ, TANH, YUKAWA, HARMONIC, PIECEWISE_LINEAR = 1, 2, 3, 4, 5
RX_LINEAR, RX_CATALYZED, RX_INHIBITED = 1, 2, 3

PAIR_KINDS = {'lj': LJ, 'tanh': TANH, 'yukawa': YUKAWA, 'harmonic': HARMONIC,
              'piecewise_linear': PIECEWISE_LINEAR}
REACTION_KINDS = {'linear': RX_LINEAR, 'locally_catalyzed': RX_CATALYZED,
                  'locally_inhibited': RX_INHIBITED}

# names of the timed kernels of rds_time_kernels
KERNELS = ('force', 'scan', 'scatter', 'react', 'bonds', 'list')


class PairTerm(C.Structure):
    _fields_ = [('kind', C.c_int32), ('type_a', C.c_int32), ('type_b', C.c_int32),
                ('n', C.c_int32), ('u0', C.c_double), ('r0', C.c_double),
                ('rc', C.c_double), ('a', C.c_double), ('dr', C.c_double)]


class WallTerm(C.Structure):
    _fields_ = [('particle_type', C.c_int32), ('direction', C.c_int32),
                ('u0', C.c_double), ('r0', C.c_double), ('a', C.c_double)]


class ReactionDef(C.Structure):
    _fields_ = [('kind', C.c_int32), ('reactant', C.c_int32), ('product', C.c_int32),
                ('catalyst', C.c_int32), ('inhibitor', C.c_int32),
                ('inhibitor2', C.c_int32), ('rate', C.c_double), ('r0', C.c_double),
                ('rate_decrease', C.c_double)]


class Model(C.Structure):
    _fields_ = [('n_types', C.c_int32), ('n_terms', C.c_int32), ('n_walls', C.c_int32),
                ('n_reactions', C.c_int32), ('terms', PairTerm * MAX_TERMS),
                ('walls', WallTerm * MAX_WALLS),
                ('reactions', ReactionDef * MAX_REACTIONS)]


class Replica(C.Structure):
    _fields_ = [('model', C.c_int32), ('n_particles', C.c_int32), ('L', C.c_double * 3),
                ('periodic', C.c_int32 * 3),
                ('force_prefac', C.c_double * MAX_TYPES),
                ('noise_prefac', C.c_double * MAX_TYPES),
                ('rate', C.c_double * MAX_REACTIONS), ('dt', C.c_double),
                ('seed', C.c_uint64)]


class Config(C.Structure):
    _fields_ = [('n_models', C.c_int32), ('n_replicas', C.c_int32),
                ('models', C.POINTER(Model)), ('replicas', C.POINTER(Replica)),
                ('precision', C.c_int32), ('rng', C.c_int32),
                ('reaction_prob', C.c_int32), ('has_molecules', C.c_int32),
                ('tile_particles', C.c_int32), ('cells_per_cutoff', C.c_int32),
                ('graph_steps', C.c_int32), ('tile_rows', C.c_int32),
                ('n_bonds', C.c_int64), ('bond_i', C.POINTER(C.c_int32)),
                ('bond_j', C.POINTER(C.c_int32)), ('bond_r0', C.POINTER(C.c_double)),
                ('bond_u0', C.POINTER(C.c_double)), ('skin', C.c_double),
                ('list_capacity', C.c_int32), ('n_dim', C.c_int32),
                ('n_angles', C.c_int64), ('angle_i', C.POINTER(C.c_int32)),
                ('angle_j', C.POINTER(C.c_int32)), ('angle_k', C.POINTER(C.c_int32)),
                ('angle_theta0', C.POINTER(C.c_double)),
                ('angle_k0', C.POINTER(C.c_double)), ('angle_threshold', C.c_double)]


class Status(C.Structure):
    _fields_ = [('steps_done', C.c_int64), ('bounds_error', C.c_int32),
                ('stream_error', C.c_int32), ('error_step', C.c_int64),
                ('error_particle', C.c_int32), ('error_axis', C.c_int32),
                ('error_value', C.c_double), ('reaction_events', C.c_int64),
                ('integrity', C.c_int32), ('angle_error', C.c_int32),
                ('rebuilds', C.c_int64), ('list_overflow', C.c_int64),
                ('list_entries', C.c_int64)]


# every entry point of include/rds.h: name -> (restype, argtypes)
_VP = C.c_void_p
SIGNATURES = {
    'rds_abi_version': (C.c_int, []),
    'rds_last_error': (C.c_char_p, []),
    'rds_ctx_create': (C.c_int, [C.POINTER(Config), C.POINTER(_VP)]),
    'rds_ctx_destroy': (C.c_int, [_VP]),
    'rds_set_state': (C.c_int, [_VP, _VP, _VP, _VP, C.c_int64]),
    'rds_get_state': (C.c_int, [_VP, _VP, _VP]),
    'rds_set_injected': (C.c_int, [_VP, _VP, _VP, _VP, C.c_int64, C.c_int64]),
    'rds_set_random_force': (C.c_int, [_VP, C.c_int32]),
    'rds_run': (C.c_int, [_VP, C.c_int64]),
    'rds_sync': (C.c_int, [_VP, C.POINTER(Status)]),
    'rds_compute_forces': (C.c_int, [_VP, _VP]),
    'rds_cell_ids': (C.c_int, [_VP, _VP]),
    'rds_grid': (C.c_int, [_VP, _VP, _VP, _VP]),
    'rds_grid3': (C.c_int, [_VP, _VP, _VP, _VP]),
    'rds_neighbor_counts': (C.c_int, [_VP, C.c_double, _VP]),
    'rds_pair_histogram': (C.c_int, [_VP, C.c_double, C.c_double, C.c_int32, C.c_uint32, _VP]),
    'rds_type_counts': (C.c_int, [_VP, _VP]),
    'rds_cluster_labels': (C.c_int, [_VP, C.c_double, C.c_uint32, _VP]),
    'rds_cluster_labels_static': (C.c_int, [_VP, C.c_double, C.c_uint32, _VP]),
    'rds_time_kernels': (C.c_int, [_VP, C.c_int64, _VP, _VP]),
    'rds_debug_fill_lds': (C.c_int, [_VP, C.c_uint32]),
    'rds_set_tiling': (C.c_int, [_VP, C.c_int32, C.c_int32]),
    'rds_get_tiling': (C.c_int, [_VP, _VP, _VP, _VP]),
    'rds_sizes': (C.c_int, [_VP, _VP, _VP, _VP]),
    'rds_stream': (C.c_int, [_VP, C.POINTER(_VP)]),
}

_lib = None


def load(path=LIB_PATH):
    """Load the engine library once; raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            'librds_hip.so is not built (%s); run `python -m pyrdsim_amd.build`' % path)
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.rds_abi_version() != ABI_VERSION:
        raise RuntimeError('librds_hip.so ABI %d, expected %d'
                           % (lib.rds_abi_version(), ABI_VERSION))
    _lib = lib
    return lib


class EngineError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(code):
    """Raise EngineError for a negative return code."""
    if code != OK:
        msg = _lib.rds_last_error().decode(errors='replace')
        raise EngineError(code, msg)
    return code
