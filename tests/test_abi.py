"""The C ABI library (CPU only: no GPU call is made).

Checks that librds_hip.so loads, reports its ABI version, exports every entry
point include/rds.h declares, and that the ctypes layouts match the header's
struct definitions field by field.
"""

import ctypes as C
import os
import re

from pyrdsim_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'rds.h')


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r'^(?:int|const char\*)\s+(rds_\w+)\s*\(', text, re.M)))


def test_library_loads_and_reports_version():
    lib = _lib.load()
    assert lib.rds_abi_version() == _lib.ABI_VERSION


def test_every_declared_symbol_is_exported_and_bound():
    lib = _lib.load()
    names = declared_functions()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name
    assert set(_lib.SIGNATURES) == set(names)


def _header_struct_fields(name):
    text = open(HEADER).read()
    end = re.search(r'\}\s*' + name + ';', text).start()
    start = text.rindex('typedef struct {', 0, end) + len('typedef struct {')
    body = re.sub(r'/\*.*?\*/', '', text[start:end], flags=re.S)
    fields = []
    for decl in body.split(';'):
        decl = decl.strip()
        if not decl:
            continue
        parts = [w for w in decl.replace('*', ' ').split() if w != 'const']
        for f in ' '.join(parts[1:]).split(','):
            f = f.strip()
            if f in ('int32_t', 'double', 'uint64_t', 'int64_t'):
                continue
            fields.append(re.sub(r'\[.*\]', '', f).strip())
    return [f for f in fields if f and f not in ('const', 'rds_model', 'rds_replica',
                                                  'rds_pair_term', 'rds_wall_term',
                                                  'rds_reaction_def')]


def test_struct_layouts_follow_the_header():
    pairs = [('rds_pair_term', _lib.PairTerm), ('rds_wall_term', _lib.WallTerm),
             ('rds_reaction_def', _lib.ReactionDef), ('rds_model', _lib.Model),
             ('rds_replica', _lib.Replica), ('rds_config', _lib.Config),
             ('rds_status', _lib.Status)]
    for cname, py in pairs:
        want = _header_struct_fields(cname)
        got = [f[0] for f in py._fields_]
        assert got == want, (cname, got, want)
    assert C.sizeof(_lib.PairTerm) == 56 and C.sizeof(_lib.WallTerm) == 32
    assert C.sizeof(_lib.ReactionDef) == 48
    assert C.sizeof(_lib.Model) == 16 + 32 * 56 + 8 * 32 + 8 * 48


def test_errors_without_context():
    lib = _lib.load()
    assert lib.rds_run(None, 1) == _lib.E_ARG
    assert b'null' in lib.rds_last_error()
    assert lib.rds_sync(None, None) == _lib.E_ARG


def _config(precision, skin):
    model = _lib.Model()
    model.n_types = 1
    rep = _lib.Replica()
    rep.n_particles = 16
    rep.L[0] = rep.L[1] = 10.0
    rep.periodic[0] = rep.periodic[1] = 1
    rep.dt = 1e-4
    cfg = _lib.Config()
    cfg.n_models = cfg.n_replicas = 1
    keep = (model, rep)
    cfg.models = C.pointer(model)
    cfg.replicas = C.pointer(rep)
    cfg.precision = precision
    cfg.skin = skin
    return cfg, keep


def test_config_errors_before_any_device_call():
    """Rejected configurations return RDS_E_ARG before touching a GPU."""
    lib = _lib.load()
    ctx = C.c_void_p()
    for prec, skin, text in ((7, 0.0, b'unknown precision'),
                             (_lib.COMPENSATED, 0.5, b'compensated'),
                             (_lib.MIXED, -1.0, b'skin')):
        cfg, keep = _config(prec, skin)
        assert lib.rds_ctx_create(C.byref(cfg), C.byref(ctx)) == _lib.E_ARG
        assert text in lib.rds_last_error(), lib.rds_last_error()
        assert not ctx.value


def test_topology_errors_before_any_device_call():
    """Bonds and angles that name a particle twice, a missing particle or two
    replicas are refused before touching a GPU."""
    import numpy as np
    lib = _lib.load()
    ctx = C.c_void_p()
    i32 = lambda v: np.ascontiguousarray(v, dtype=np.int32)
    f64 = lambda v: np.ascontiguousarray(v, dtype=np.float64)
    P32, PD = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    cases = (((3, 4, 3), b'invalid particles'), ((3, 4, 16), b'invalid particles'),
             ((3, -1, 5), b'invalid particles'))
    for (i, j, k), text in cases:
        cfg, keep = _config(_lib.FP64, 0.0)
        arrs = (i32([i]), i32([j]), i32([k]), f64([2.0]), f64([10.0]))
        cfg.n_angles = 1
        cfg.angle_i, cfg.angle_j, cfg.angle_k = (a.ctypes.data_as(P32) for a in arrs[:3])
        cfg.angle_theta0, cfg.angle_k0 = (a.ctypes.data_as(PD) for a in arrs[3:])
        assert lib.rds_ctx_create(C.byref(cfg), C.byref(ctx)) == _lib.E_ARG
        assert text in lib.rds_last_error(), lib.rds_last_error()
        assert not ctx.value
    cfg, keep = _config(_lib.FP64, 0.0)
    cfg.n_angles = 2
    assert lib.rds_ctx_create(C.byref(cfg), C.byref(ctx)) == _lib.E_ARG
    assert b'angle arrays missing' in lib.rds_last_error()
    # two replicas of 16 particles: a bond (15, 16) spans them
    model = _lib.Model()
    model.n_types = 1
    reps = (_lib.Replica * 2)()
    for r in reps:
        r.n_particles = 16
        r.L[0] = r.L[1] = 10.0
        r.periodic[0] = r.periodic[1] = 1
        r.dt = 1e-4
    cfg = _lib.Config()
    cfg.n_models, cfg.n_replicas = 1, 2
    cfg.models = C.pointer(model)
    cfg.replicas = C.cast(reps, C.POINTER(_lib.Replica))
    bi, bj, br, bu = i32([15]), i32([16]), f64([1.0]), f64([1.0])
    cfg.n_bonds = 1
    cfg.bond_i, cfg.bond_j = bi.ctypes.data_as(P32), bj.ctypes.data_as(P32)
    cfg.bond_r0, cfg.bond_u0 = br.ctypes.data_as(PD), bu.ctypes.data_as(PD)
    assert lib.rds_ctx_create(C.byref(cfg), C.byref(ctx)) == _lib.E_ARG
    assert b'spans replicas' in lib.rds_last_error(), lib.rds_last_error()
