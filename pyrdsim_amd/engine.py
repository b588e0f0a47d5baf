"""Host side of the engine: model tables, replicas and device buffers.

``Engine`` owns one ``rds_ctx`` of ``librds_hip.so``.  PyTorch-ROCm tensors
serve only as device buffers for the arrays that cross the C ABI; all compute
runs in the library's HIP kernels.
"""

import ctypes as C

import numpy as np
import torch

from . import _lib as L

PRECISIONS = {'fp64': L.FP64, 'mixed': L.MIXED, 'compensated': L.COMPENSATED}
RNGS = {'philox': L.RNG_PHILOX, 'injected': L.RNG_INJECTED}
PROBS = {'linear': L.PROB_LINEAR, 'exp': L.PROB_EXP}


def _lower(d):
    return {str(k).lower(): v for k, v in d.items()}


def build_model(interactions, reactions, n_types):
    """Model table from the reference's interaction and reaction dicts.

    ``interactions`` are the sections of nonbonded_interactions.txt (or the
    ``interactions`` entry of the constructor dict), ``reactions`` those of
    reactions.txt; both follow pyrdsim.pyx:1037-1154 and 2070-2222.  Returns
    ``(model, rates)`` with the per-reaction rates in file order.
    """
    m = L.Model()
    if not 1 <= n_types <= L.MAX_TYPES:
        raise RuntimeError('the engine supports 1..%d particle types, got %d'
                           % (L.MAX_TYPES, n_types))
    m.n_types = n_types
    nt = nw = 0
    for raw in interactions:
        it = _lower(raw)
        kind = str(it['type']).lower()
        if kind == 'wall':
            if nw >= L.MAX_WALLS:
                raise RuntimeError('too many wall interactions')
            w = m.walls[nw]
            w.particle_type = int(it['particle_type'])
            w.direction = int(it['direction'])
            w.u0 = float(it['u0'])
            w.r0 = float(it['r0'])
            w.a = float(it['a'])
            if w.direction not in (0, 1):
                raise RuntimeError('wall direction must be 0 or 1 in two dimensions')
            nw += 1
            continue
        if kind not in L.PAIR_KINDS:
            raise RuntimeError("Interaction {0} not recognized".format(kind))
        if nt >= L.MAX_TERMS:
            raise RuntimeError('too many pair interactions')
        t = m.terms[nt]
        t.kind = L.PAIR_KINDS[kind]
        pair = [int(v) for v in np.asarray(it['pair']).ravel()]
        t.type_a, t.type_b = pair[0], pair[1]
        t.u0 = float(it['u0'])
        t.r0 = float(it.get('r0', 0.0))
        t.rc = float(it['r0'] if kind == 'piecewise_linear' else it['rc'])
        t.n = int(it.get('n', 0))
        t.a = float(it.get('a', 1.0))
        t.dr = float(it.get('dr', 1.0))
        nt += 1
    m.n_terms = nt
    m.n_walls = nw
    rates = []
    for q, raw in enumerate(reactions):
        rx = _lower(raw)
        if q >= L.MAX_REACTIONS:
            raise RuntimeError('too many reactions')
        kind = str(rx['type']).lower()
        if kind not in L.REACTION_KINDS:
            raise RuntimeError('reaction type {0} not recognized'.format(kind))
        r = m.reactions[q]
        r.kind = L.REACTION_KINDS[kind]
        r.reactant = int(rx['reactant'])
        r.product = int(rx['product'])
        r.catalyst = int(rx.get('catalyst', -1))
        r.inhibitor = int(rx.get('inhibitor', -1))
        r.inhibitor2 = int(rx.get('inhibitor2', -1))
        r.rate = float(rx['rate'])
        r.r0 = float(rx.get('r0', 0.0))
        r.rate_decrease = float(rx.get('rate_decrease', 0.0))
        rates.append(r.rate)
    m.n_reactions = len(rates)
    return m, rates


def prefactors(dt, tau_by_tauD):
    """force_prefacs and noise_prefacs exactly as pyrdsim.pyx:488-496."""
    tau = np.array(tau_by_tauD, dtype=np.float64)
    mask = tau < 0
    noise = np.ones_like(tau)
    force = np.ones_like(tau)
    noise[mask] = 0.
    noise[~mask] = np.sqrt(2. * dt * tau[~mask])
    force[mask] = 0.
    force[~mask] = dt * tau[~mask]
    return force, noise


def make_replica(model_index, n_particles, L_box, periodic, dt, tau_by_tauD, rates, seed):
    """One replica; L_box and periodic have one entry per axis (2 or 3)."""
    r = L.Replica()
    r.model = model_index
    r.n_particles = int(n_particles)
    nd = len(L_box)
    if nd not in (2, 3) or len(periodic) < nd:
        raise RuntimeError('L_box needs 2 or 3 axes and periodic one entry per axis')
    for k in range(nd):
        r.L[k] = float(L_box[k])
        pk = float(periodic[k])
        if pk == 0.0 or pk != pk:
            # the reference wraps positions unless periodic < 0
            # (pyrdsim.pyx:2274) but takes minimum images only where
            # periodic > 0 (pyrdsim.pyx:1629): 0 mixes the two
            raise RuntimeError(
                'periodic[%d] = %r: use a positive value (periodic) or a negative one '
                '(reflecting); 0 wraps positions but not distances in the reference' % (k, pk))
        r.periodic[k] = 1 if pk > 0 else 0
    fp, npf = prefactors(dt, tau_by_tauD)
    if len(fp) > L.MAX_TYPES:
        raise RuntimeError('at most %d particle species' % L.MAX_TYPES)
    for t in range(len(fp)):
        r.force_prefac[t] = fp[t]
        r.noise_prefac[t] = npf[t]
    for q, v in enumerate(rates):
        r.rate[q] = float(v)
    r.dt = float(dt)
    r.seed = int(seed) & 0xffffffffffffffff
    return r


class Engine:
    """One engine context over one or more replicas on one GPU.

    skin = 0 re-sorts by cell and scans the 3x3 stencil every step; skin > 0
    keeps Verlet lists within (pair range + skin), rebuilt with the sort only
    after some particle moved more than skin/2 (list_capacity entries per
    particle, 0 = 64; a particle with more falls back to the stencil scan).

    bonds = (i, j, r0, u0) and angles = (i1, i2, i3, theta0 [rad], k0 [per
    rad^2]) are arrays over global particle indices, in definition order
    (pyrdsim.pyx:1287-1407); i2 is an angle's vertex.

    n_dim = 3 integrates three-dimensional replicas (cell search, 'fp64' or
    'mixed'); positions, forces and injected normals then carry three
    components.
    """

    def __init__(self, models, replicas, precision='mixed', rng='philox',
                 reaction_prob='linear', bonds=None, has_molecules=False,
                 tile_particles=0, graph_steps=0, tile_rows=1, device='cuda',
                 skin=0.0, list_capacity=0, angles=None, angle_threshold=0.1, n_dim=2):
        self._init = (list(models), list(replicas),
                      dict(precision=precision, rng=rng, reaction_prob=reaction_prob,
                           bonds=bonds, has_molecules=has_molecules,
                           tile_particles=tile_particles, graph_steps=graph_steps,
                           tile_rows=tile_rows, device=device, skin=skin,
                           list_capacity=list_capacity, angles=angles,
                           angle_threshold=angle_threshold, n_dim=n_dim))
        self.lib = L.load()
        if not torch.cuda.is_available():
            raise RuntimeError('the engine needs a ROCm GPU; none is visible')
        self.device = torch.device(device)
        if self.device.index is None:
            self.device = torch.device('cuda', torch.cuda.current_device())
        torch.cuda.set_device(self.device)
        self.models = (L.Model * len(models))(*models)
        self.replicas = (L.Replica * len(replicas))(*replicas)
        self.n_total = int(sum(r.n_particles for r in replicas))
        self.has_molecules = bool(has_molecules)
        cfg = L.Config()
        cfg.n_models = len(models)
        cfg.n_replicas = len(replicas)
        cfg.models = C.cast(self.models, C.POINTER(L.Model))
        cfg.replicas = C.cast(self.replicas, C.POINTER(L.Replica))
        cfg.precision = PRECISIONS[precision]
        cfg.rng = RNGS[rng]
        cfg.reaction_prob = PROBS[reaction_prob]
        cfg.has_molecules = 1 if has_molecules else 0
        cfg.tile_particles = int(tile_particles)
        cfg.cells_per_cutoff = 1
        cfg.graph_steps = int(graph_steps)
        cfg.tile_rows = int(tile_rows)
        cfg.skin = float(skin)
        cfg.list_capacity = int(list_capacity)
        if int(n_dim) not in (2, 3):
            raise RuntimeError('n_dim must be 2 or 3')
        cfg.n_dim = int(n_dim)
        self.n_dim = int(n_dim)
        self.skin = float(skin)
        self._bond_arrays = None
        if bonds is not None and len(bonds[0]) > 0:
            bi = np.ascontiguousarray(bonds[0], dtype=np.int32)
            bj = np.ascontiguousarray(bonds[1], dtype=np.int32)
            br = np.ascontiguousarray(bonds[2], dtype=np.float64)
            bu = np.ascontiguousarray(bonds[3], dtype=np.float64)
            self._bond_arrays = (bi, bj, br, bu)
            cfg.n_bonds = len(bi)
            cfg.bond_i = bi.ctypes.data_as(C.POINTER(C.c_int32))
            cfg.bond_j = bj.ctypes.data_as(C.POINTER(C.c_int32))
            cfg.bond_r0 = br.ctypes.data_as(C.POINTER(C.c_double))
            cfg.bond_u0 = bu.ctypes.data_as(C.POINTER(C.c_double))
        self._angle_arrays = None
        cfg.angle_threshold = float(angle_threshold)
        if angles is not None and len(angles[0]) > 0:
            ai, aj, ak = (np.ascontiguousarray(a, dtype=np.int32) for a in angles[:3])
            at = np.ascontiguousarray(angles[3], dtype=np.float64)
            ak0 = np.ascontiguousarray(angles[4], dtype=np.float64)
            if not (len(ai) == len(aj) == len(ak) == len(at) == len(ak0)):
                raise ValueError('angle arrays differ in length')
            self._angle_arrays = (ai, aj, ak, at, ak0)
            cfg.n_angles = len(ai)
            cfg.angle_i = ai.ctypes.data_as(C.POINTER(C.c_int32))
            cfg.angle_j = aj.ctypes.data_as(C.POINTER(C.c_int32))
            cfg.angle_k = ak.ctypes.data_as(C.POINTER(C.c_int32))
            cfg.angle_theta0 = at.ctypes.data_as(C.POINTER(C.c_double))
            cfg.angle_k0 = ak0.ctypes.data_as(C.POINTER(C.c_double))
        self.precision = precision
        self.rng = rng
        ctx = C.c_void_p()
        L.check(self.lib.rds_ctx_create(C.byref(cfg), C.byref(ctx)))
        self.ctx = ctx
        self._inj = None
        self.step = 0

    def clone(self, rng='philox'):
        """A new context with this one's configuration, tiling and current
        state (positions, types, molecules, step counter), for work that must
        not touch this one -- the autotuner times kernels on a clone.  The
        clone draws Philox numbers by default: injected streams stay here."""
        models, reps, kw = self._init
        kw = dict(kw, rng=rng)
        rows, per, _ = self.tiling()
        kw.update(tile_rows=rows, tile_particles=per)
        e = Engine(models, reps, **kw)
        x, t = self.state()
        e.set_state(x, t, getattr(self, '_molecules', None), step0=self.step)
        return e

    def close(self):
        if getattr(self, 'ctx', None):
            self.lib.rds_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- buffers ----------------------------------------------------------
    def _dev(self, arr, dtype):
        t = torch.as_tensor(np.ascontiguousarray(arr), dtype=dtype)
        return t.to(self.device, non_blocking=False).contiguous()

    def set_state(self, x, types, molecules=None, step0=0):
        """Positions [N][n_dim] f64, types [N], molecule numbers [N] (caller order)."""
        x = np.asarray(x, dtype=np.float64).reshape(self.n_total, self.n_dim)
        types = np.asarray(types).reshape(-1)
        lo = 0
        for r in self.replicas:
            nt = self.models[r.model].n_types
            seg = types[lo:lo + r.n_particles]
            if seg.size and (seg.min() < 0 or seg.max() >= nt):
                raise RuntimeError('particle types must lie in [0, %d) for this model' % nt)
            lo += r.n_particles
        xd = self._dev(x, torch.float64)
        td = self._dev(np.asarray(types, dtype=np.int32).reshape(-1), torch.int32)
        md = None
        if self.has_molecules:
            self._molecules = np.asarray(molecules, dtype=np.int32).reshape(-1).copy()
            md = self._dev(self._molecules, torch.int32)
        torch.cuda.synchronize(self.device)
        L.check(self.lib.rds_set_state(self.ctx, xd.data_ptr(), td.data_ptr(),
                                       md.data_ptr() if md is not None else None,
                                       int(step0)))
        self.step = int(step0)

    def set_injected(self, normals, u1, u2=None, first_step=1):
        """Injected streams: normals [S][N][n_dim], u1/u2 [S][N] for steps
        first_step .. first_step+S-1 (the reference's step numbers)."""
        n = self._dev(np.asarray(normals, dtype=np.float64), torch.float64)
        a = self._dev(np.asarray(u1, dtype=np.float64), torch.float64)
        b = self._dev(np.asarray(u2, dtype=np.float64), torch.float64) if u2 is not None else None
        S = n.shape[0]
        if n.numel() != S * self.n_total * self.n_dim or a.numel() != S * self.n_total or \
                (b is not None and b.numel() != S * self.n_total):
            raise ValueError('injected stream shapes do not match the particle count')
        torch.cuda.synchronize(self.device)
        self._inj = (n, a, b)  # keep alive
        L.check(self.lib.rds_set_injected(self.ctx, n.data_ptr(), a.data_ptr(),
                                          b.data_ptr() if b is not None else None,
                                          int(first_step), int(S)))

    def set_random_force(self, include):
        L.check(self.lib.rds_set_random_force(self.ctx, 1 if include else 0))

    # -- stepping ---------------------------------------------------------
    def run(self, n_steps):
        L.check(self.lib.rds_run(self.ctx, int(n_steps)))
        self.step += int(n_steps)

    def sync(self):
        """Wait for the GPU; returns the status, raises on a kernel-side error."""
        st = L.Status()
        rc = self.lib.rds_sync(self.ctx, C.byref(st))
        if rc == L.E_BOUNDS:
            raise RuntimeError(self.lib.rds_last_error().decode())
        L.check(rc)
        return st

    # -- outputs ----------------------------------------------------------
    def state(self):
        x = torch.empty((self.n_total, self.n_dim), dtype=torch.float64, device=self.device)
        t = torch.empty((self.n_total,), dtype=torch.int32, device=self.device)
        L.check(self.lib.rds_get_state(self.ctx, x.data_ptr(), t.data_ptr()))
        self.sync()
        return x.cpu().numpy(), t.cpu().numpy()

    def state_device(self, x, t):
        """Write the state into caller-provided device tensors (no sync)."""
        L.check(self.lib.rds_get_state(self.ctx, x.data_ptr(), t.data_ptr()))

    def forces(self):
        f = torch.empty((self.n_total, self.n_dim), dtype=torch.float64, device=self.device)
        L.check(self.lib.rds_compute_forces(self.ctx, f.data_ptr()))
        self.sync()
        return f.cpu().numpy()

    def cell_ids(self):
        c = torch.empty((self.n_total,), dtype=torch.int32, device=self.device)
        L.check(self.lib.rds_cell_ids(self.ctx, c.data_ptr()))
        self.sync()
        return c.cpu().numpy()

    def grid(self):
        """Cells per axis [R][n_dim], cell lengths [R][n_dim], first global cell [R]."""
        n = len(self.replicas)
        lxyz = np.zeros((n, 3), dtype=np.int32)
        cl = np.zeros((n, 3), dtype=np.float64)
        base = np.zeros((n,), dtype=np.int32)
        L.check(self.lib.rds_grid3(self.ctx, lxyz.ctypes.data, cl.ctypes.data, base.ctypes.data))
        return lxyz[:, :self.n_dim].copy(), cl[:, :self.n_dim].copy(), base

    def neighbor_counts(self, r):
        c = torch.empty((self.n_total,), dtype=torch.int32, device=self.device)
        L.check(self.lib.rds_neighbor_counts(self.ctx, float(r), c.data_ptr()))
        self.sync()
        return c.cpu().numpy()

    # -- analysis on the device -------------------------------------------
    @staticmethod
    def _mask(types):
        if types is None:
            return 0
        m = 0
        for t in types:
            m |= 1 << int(t)
        return m

    def pair_histogram(self, r_max=5.0, dr=0.02, types=None):
        """Per replica [n_replicas][n_bins]: unordered pairs of the given types
        (None = all) at distance r <= r_max, bin floor(r/dr); n_bins =
        round(r_max/dr)."""
        nb = int(round(r_max / dr))
        h = torch.empty((len(self.replicas), nb), dtype=torch.int64, device=self.device)
        L.check(self.lib.rds_pair_histogram(self.ctx, float(r_max), float(dr), nb,
                                            self._mask(types), h.data_ptr()))
        self.sync()
        return h.cpu().numpy()

    def type_counts(self):
        """Particles of each type per replica, [n_replicas][MAX_TYPES]."""
        cnt = torch.empty((len(self.replicas), L.MAX_TYPES), dtype=torch.int64,
                          device=self.device)
        L.check(self.lib.rds_type_counts(self.ctx, cnt.data_ptr()))
        self.sync()
        return cnt.cpu().numpy()

    def cluster_labels(self, r_link=1.5, types=(1,)):
        """Per particle (caller order): the smallest index of its cluster of
        particles of the given types linked at r <= r_link, or -1."""
        lab = torch.empty((self.n_total,), dtype=torch.int32, device=self.device)
        L.check(self.lib.rds_cluster_labels(self.ctx, float(r_link), self._mask(types),
                                            lab.data_ptr()))
        self.sync()
        return lab.cpu().numpy()

    def time_kernels(self, n_steps):
        """Average HIP-event duration (ms) of each kernel over n_steps eager steps."""
        avg = np.zeros(len(L.KERNELS), dtype=np.float64)
        cnt = np.zeros(len(L.KERNELS), dtype=np.int64)
        L.check(self.lib.rds_time_kernels(self.ctx, int(n_steps), avg.ctypes.data,
                                          cnt.ctypes.data))
        self.step += int(n_steps)
        return {k: (float(avg[i]), int(cnt[i])) for i, k in enumerate(L.KERNELS)}

    def set_tiling(self, tile_rows, tile_particles):
        L.check(self.lib.rds_set_tiling(self.ctx, int(tile_rows), int(tile_particles)))

    def tiling(self):
        r, t, c = C.c_int32(), C.c_int32(), C.c_int32()
        L.check(self.lib.rds_get_tiling(self.ctx, C.byref(r), C.byref(t), C.byref(c)))
        return r.value, t.value, c.value

    def debug_fill_lds(self, pattern=0x7fc00001):
        """Test hook: fill LDS with a (NaN) pattern before the next launch."""
        L.check(self.lib.rds_debug_fill_lds(self.ctx, int(pattern) & 0xffffffff))

    def sizes(self):
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        L.check(self.lib.rds_sizes(self.ctx, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value
