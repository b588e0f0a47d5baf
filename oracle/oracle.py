"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this
module, as the checker; the engine never calls it.  The C++ restatement of the
reference's step loop lives in rds_oracle.cpp (see its header for the
pyrdsim.pyx lines each part follows).
"""

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle.so')

K_LJ, K_TANH, K_YUKAWA, K_HARMONIC, K_PWLINEAR, K_WALL = 1, 2, 3, 4, 5, 6
RX_LINEAR, RX_CATALYZED, RX_INHIBITED = 1, 2, 3
RNG_MT, RNG_INJECTED, RNG_PHILOX = 0, 1, 2
NL_FAITHFUL, NL_EXACT = 0, 1

KINDS = {'lj': K_LJ, 'tanh': K_TANH, 'yukawa': K_YUKAWA, 'harmonic': K_HARMONIC,
         'piecewise_linear': K_PWLINEAR, 'wall': K_WALL}
RX_KINDS = {'linear': RX_LINEAR, 'locally_catalyzed': RX_CATALYZED,
            'locally_inhibited': RX_INHIBITED}


class Interaction(C.Structure):
    _fields_ = [('kind', C.c_int32), ('p0', C.c_int32), ('p1', C.c_int32), ('n', C.c_int32),
                ('u0', C.c_double), ('r0', C.c_double), ('rc', C.c_double),
                ('a', C.c_double), ('dr', C.c_double)]


class Reaction(C.Structure):
    _fields_ = [('kind', C.c_int32), ('reactant', C.c_int32), ('product', C.c_int32),
                ('catalyst', C.c_int32), ('inhibitor', C.c_int32), ('inhibitor2', C.c_int32),
                ('rate', C.c_double), ('r0', C.c_double), ('rate_decrease', C.c_double)]


class Bond(C.Structure):
    _fields_ = [('i1', C.c_int32), ('i2', C.c_int32), ('r0', C.c_double), ('u0', C.c_double)]


class Spec(C.Structure):
    _fields_ = [('n', C.c_int32), ('n_dim', C.c_int32), ('n_species', C.c_int32),
                ('n_inter', C.c_int32), ('n_rx', C.c_int32), ('n_bonds', C.c_int32),
                ('L', C.c_double * 3), ('periodic', C.c_double * 3), ('dt', C.c_double),
                ('include_random_force', C.c_double),
                ('force_prefac', C.POINTER(C.c_double)), ('noise_prefac', C.POINTER(C.c_double)),
                ('inter', C.POINTER(Interaction)), ('rx', C.POINTER(Reaction)),
                ('bonds', C.POINTER(Bond)), ('molecule', C.POINTER(C.c_int32)),
                ('partners', C.POINTER(C.c_int32)), ('rs', C.c_double),
                ('ratio', C.c_int32), ('nl_freq', C.c_int32), ('nl_mode', C.c_int32),
                ('rng_mode', C.c_int32), ('reaction_prob', C.c_int32), ('pad', C.c_int32),
                ('seed', C.c_uint64)]


_lib = None


def build():
    subprocess.run(['make', '-s', '-C', HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        'orc_create': (vp, [C.POINTER(Spec), vp, vp]),
        'orc_destroy': (None, [vp]),
        'orc_set_injected': (C.c_int, [vp, vp, vp, vp, C.c_int64, C.c_int64]),
        'orc_init': (C.c_int, [vp]),
        'orc_set_angles': (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp, C.c_double]),
        'orc_run': (C.c_int, [vp, C.c_int64, C.c_int64]),
        'orc_get_state': (C.c_int, [vp, vp, vp, vp]),
        'orc_set_state': (C.c_int, [vp, vp, vp, vp]),
        'orc_reaction_info': (C.c_int64, [vp, vp, vp]),
        'orc_nl_size': (C.c_int64, [vp]),
        'orc_nl_pairs': (C.c_int64, [vp, vp, C.c_int64]),
        'orc_mt_sequence': (C.c_int, [C.c_uint64, C.c_int64, vp, vp]),
        'orc_mt_raw': (C.c_int, [C.c_uint64, C.c_int64, vp]),
        'orc_philox_streams': (C.c_int, [C.c_uint64, C.c_int64, C.c_int32, vp, vp]),
        'orc_philox4x32_10': (None, [vp, vp, vp]),
        'orc_force_law': (C.c_double, [C.POINTER(Interaction), C.c_double]),
        'orc_last_error': (C.c_char_p, []),
    }
    for k, (res, args) in sig.items():
        f = getattr(lib, k)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _lower(d):
    return {str(k).lower(): v for k, v in d.items()}


def make_interaction(d):
    d = _lower(d)
    it = Interaction()
    kind = str(d['type']).lower()
    it.kind = KINDS[kind]
    if kind == 'wall':
        it.p0 = int(d['particle_type'])
        it.p1 = int(d['direction'])
    else:
        pair = [int(v) for v in np.asarray(d['pair']).ravel()]
        it.p0, it.p1 = pair
    it.n = int(d.get('n', 0))
    it.u0 = float(d['u0'])
    it.r0 = float(d.get('r0', 0.0))
    it.rc = float(d['r0'] if kind == 'piecewise_linear' else d.get('rc', 0.0))
    it.a = float(d.get('a', 1.0))
    it.dr = float(d.get('dr', 1.0))
    return it


def force_law(d, r):
    """f(r) = dU/dr of one interaction dict (or the wall law at coordinate r)."""
    it = make_interaction(d)
    return load().orc_force_law(C.byref(it), float(r))


class Oracle:
    """One system stepped by the CPU restatement of the reference."""

    def __init__(self, x0, types0, L, periodic, dt, tau_by_tauD, interactions,
                 reactions=(), molecules=None, bonds=(), seed=0, rng='philox',
                 nl='exact', rs=None, ratio=2, nl_freq=10, include_random_force=1.0,
                 reaction_prob='linear', angles=(), angle_threshold=0.1):
        lib = load()
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        n, d = x0.shape
        types0 = np.ascontiguousarray(types0, dtype=np.int32)
        tau = np.asarray(tau_by_tauD, dtype=np.float64)
        ns = len(tau)
        mask = tau < 0
        self.noise = np.ones_like(tau)
        self.force = np.ones_like(tau)
        self.noise[mask] = 0.
        self.noise[~mask] = np.sqrt(2. * dt * tau[~mask])
        self.force[mask] = 0.
        self.force[~mask] = dt * tau[~mask]
        its = [make_interaction(i) for i in interactions]
        self.inter = (Interaction * max(1, len(its)))(*its)
        rxs = []
        partners = np.zeros((ns, ns), dtype=np.int32)
        for i in interactions:
            i = _lower(i)
            if str(i['type']).lower() != 'wall':
                a, b = [int(v) for v in np.asarray(i['pair']).ravel()]
                if a < ns and b < ns:
                    partners[a, b] = partners[b, a] = 1
        max_rc = max([float(_lower(i).get('rc', _lower(i).get('r0', 0.0)))
                      if str(_lower(i)['type']).lower() != 'piecewise_linear'
                      else float(_lower(i)['r0']) for i in interactions] + [0.0])
        for r in reactions:
            r = _lower(r)
            q = Reaction()
            q.kind = RX_KINDS[str(r['type']).lower()]
            q.reactant = int(r['reactant'])
            q.product = int(r['product'])
            q.catalyst = int(r.get('catalyst', -1))
            q.inhibitor = int(r.get('inhibitor', -1))
            q.inhibitor2 = int(r.get('inhibitor2', -1))
            q.rate = float(r['rate'])
            q.r0 = float(r.get('r0', 0.0))
            q.rate_decrease = float(r.get('rate_decrease', 0.0))
            if q.kind == RX_CATALYZED:
                partners[q.reactant, q.catalyst] = partners[q.catalyst, q.reactant] = 1
            if q.kind == RX_INHIBITED:
                partners[q.reactant, q.inhibitor] = partners[q.inhibitor, q.reactant] = 1
                if q.inhibitor2 >= 0:
                    partners[q.reactant, q.inhibitor2] = partners[q.inhibitor2, q.reactant] = 1
            rxs.append(q)
        self.rx = (Reaction * max(1, len(rxs)))(*rxs)
        bl = [Bond(int(b[0]), int(b[1]), float(b[2]), float(b[3])) for b in bonds]
        self.bonds = (Bond * max(1, len(bl)))(*bl)
        if molecules is None:
            molecules = np.arange(n)
        self.molecule = np.ascontiguousarray(molecules, dtype=np.int32)
        self.partners = np.ascontiguousarray(partners)
        s = Spec()
        s.n, s.n_dim, s.n_species = n, d, ns
        s.n_inter, s.n_rx, s.n_bonds = len(its), len(rxs), len(bl)
        for k in range(d):
            s.L[k] = float(L[k])
            s.periodic[k] = float(periodic[k])
        s.dt = float(dt)
        s.include_random_force = float(include_random_force)
        s.force_prefac = self.force.ctypes.data_as(C.POINTER(C.c_double))
        s.noise_prefac = self.noise.ctypes.data_as(C.POINTER(C.c_double))
        s.inter = C.cast(self.inter, C.POINTER(Interaction))
        s.rx = C.cast(self.rx, C.POINTER(Reaction))
        s.bonds = C.cast(self.bonds, C.POINTER(Bond))
        s.molecule = self.molecule.ctypes.data_as(C.POINTER(C.c_int32))
        s.partners = self.partners.ctypes.data_as(C.POINTER(C.c_int32))
        s.rs = float(rs) if rs is not None else 2.0 * max_rc
        s.ratio = int(ratio)
        s.nl_freq = int(nl_freq)
        s.nl_mode = NL_EXACT if nl == 'exact' else NL_FAITHFUL
        s.rng_mode = {'mt19937': RNG_MT, 'injected': RNG_INJECTED, 'philox': RNG_PHILOX}[rng]
        s.reaction_prob = 0 if reaction_prob == 'linear' else 1
        s.seed = int(seed) & 0xffffffffffffffff
        self.spec = s
        self.n, self.d = n, d
        self.n_rx = len(rxs)
        self.h = lib.orc_create(C.byref(s), x0.ctypes.data, types0.ctypes.data)
        if not self.h:
            raise RuntimeError(lib.orc_last_error().decode())
        # angles: (i1, i2, i3, theta0 [rad], k0 [per rad^2]) in definition order
        if len(angles):
            ang = list(angles)
            self._ang = (np.ascontiguousarray([a[0] for a in ang], dtype=np.int32),
                         np.ascontiguousarray([a[1] for a in ang], dtype=np.int32),
                         np.ascontiguousarray([a[2] for a in ang], dtype=np.int32),
                         np.ascontiguousarray([a[3] for a in ang], dtype=np.float64),
                         np.ascontiguousarray([a[4] for a in ang], dtype=np.float64))
            lib.orc_set_angles(self.h, len(ang), *[v.ctypes.data for v in self._ang],
                               float(angle_threshold))
        self._inj = None
        self.step = 0

    def __del__(self):
        try:
            if self.h:
                load().orc_destroy(self.h)
        except Exception:
            pass

    def set_injected(self, normals, u1, u2, first_step=1):
        n = np.ascontiguousarray(normals, dtype=np.float64)
        a = np.ascontiguousarray(u1, dtype=np.float64)
        b = np.ascontiguousarray(u2, dtype=np.float64)
        self._inj = (n, a, b)
        load().orc_set_injected(self.h, n.ctypes.data, a.ctypes.data, b.ctypes.data,
                                n.shape[0], int(first_step))

    def init(self):
        """Initial neighbour list and drift (pyrdsim.pyx:2398-2401)."""
        load().orc_init(self.h)

    def run(self, first_step, n_steps):
        rc = load().orc_run(self.h, int(first_step), int(n_steps))
        if rc == -3:
            raise RuntimeError(load().orc_last_error().decode())
        if rc != 0:
            raise RuntimeError(load().orc_last_error().decode())
        self.step = first_step + n_steps - 1

    def state(self):
        x = np.zeros((self.n, self.d))
        t = np.zeros(self.n, dtype=np.int32)
        a = np.zeros((self.n, self.d))
        load().orc_get_state(self.h, x.ctypes.data, t.ctypes.data, a.ctypes.data)
        return x, t, a

    def set_state(self, x, types, drift=None):
        """Replace positions and types (and the drift of the last step); the
        neighbour list stays as it is until its next refresh."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        t = np.ascontiguousarray(types, dtype=np.int32)
        a = None if drift is None else np.ascontiguousarray(drift, dtype=np.float64)
        load().orc_set_state(self.h, x.ctypes.data, t.ctypes.data,
                             None if a is None else a.ctypes.data)

    def reaction_info(self):
        rates = np.zeros((max(1, self.n_rx), self.n))
        total = np.zeros(self.n)
        ev = load().orc_reaction_info(self.h, rates.ctypes.data, total.ctypes.data)
        return rates[:self.n_rx], total, ev

    def nl_pairs(self):
        lib = load()
        m = lib.orc_nl_size(self.h)
        out = np.zeros((max(1, m), 2), dtype=np.int32)
        lib.orc_nl_pairs(self.h, out.ctypes.data, m)
        return out[:m]


def mt_raw(seed, n):
    out = np.zeros(n, dtype=np.uint32)
    load().orc_mt_raw(int(seed), int(n), out.ctypes.data)
    return out


def mt_sequence(seed, kinds):
    """Draws of the reference stream: kinds 0 = normal, 1 = uniform."""
    k = np.ascontiguousarray(kinds, dtype=np.int8)
    out = np.zeros(len(k))
    load().orc_mt_sequence(int(seed), len(k), k.ctypes.data, out.ctypes.data)
    return out


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    load().orc_philox4x32_10(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def philox_streams(seed, step, n):
    """The engine's noise (normals [n][2]) and reaction uniforms ([n][2])."""
    nm = np.zeros((n, 2))
    u = np.zeros((n, 2))
    load().orc_philox_streams(int(seed), int(step), int(n), nm.ctypes.data, u.ctypes.data)
    return nm, u
