"""Synthetic workloads: the reference-format models and seeded initial states.

Five configurations (all 2-D, kT = 1):

* ``lj1``      -- N=1024 one-species 12-6 LJ (sigma = eps = 1, rc = 2.5) at
  density 0.5 in a periodic square box, dt = 1e-4, FP64.
* ``kalj``     -- N=262,144 two-species LJ mixture (50/50), density 0.6,
  eps_AA=1, eps_BB=0.5, eps_AB=1.5, sigma_AA=1, sigma_BB=0.88, sigma_AB=0.8,
  rc = 2.5 sigma_AA, dt = 5e-5.
* ``wall``     -- N=65,536 at density 0.4, periodic in x and reflecting in y;
  inactive particles (type 0) repel (WCA), active ones (type 1) attract
  (LJ eps=2, rc=2.5) and feel a tanh wall well at y = 0 of depth 3 kT and
  width 0.5 sigma; activation 0 -> 1 at rate k_a = 0.1, deactivation 1 -> 0
  at rate k_d = 1.0 per active neighbour within 1.5 sigma; dt = 1e-4.
* ``patchy``   -- 16,384 molecules, each a core bead (type 0) with 3 patch beads
  at radius 0.5 and 120 degrees, harmonic springs (k = 1000) core-patch and
  patch-patch; cores repel by Yukawa (kappa = 5, A = 10), active patches
  (type 2) attract by a tanh well (depth 5, range 0.3); activation 1 -> 2 at
  0.05, catalysed deactivation 2 -> 1 at 0.5 per active patch within 0.3;
  packing fraction 0.3; dt = 1e-5.
* ``ensemble`` -- 512 replicas, half of the kalj model and half of the wall
  model, N per replica log-uniform over 2,048..32,768, rates log-uniform over
  0.01..1 for the wall replicas, one Philox seed per replica.

Initial positions are uniform random with overlap rejection at 0.9 sigma
(random sequential addition), from a NumPy PCG64 seed.

Three-dimensional variants (``lj3``, ``wall3``, ``patchy3``) take the same
models in a cube, with positions on a randomly occupied simple-cubic lattice
jittered by up to 10% of its spacing.
"""

import math

import numpy as np

LJ_R0 = 2.0 ** (1.0 / 6.0)   # minimum of the 12-6 LJ: r0 = 2^(1/6) sigma


def rsa_positions(n, L, dmin, rng, max_rounds=10000):
    """n points uniform in the box [0,Lx) x [0,Ly) with all pair distances
    >= dmin (minimum image), by batched random sequential addition."""
    Lx, Ly = float(L[0]), float(L[1])
    h = dmin / math.sqrt(2.0)          # at most one point per grid cell
    gx, gy = max(1, int(Lx / h)), max(1, int(Ly / h))
    hx, hy = Lx / gx, Ly / gy
    grid = -np.ones((gx, gy), dtype=np.int64)
    pts = np.zeros((n, 2))
    have = 0
    reach_x = int(math.ceil(dmin / hx))
    reach_y = int(math.ceil(dmin / hy))
    offs = [(a, b) for a in range(-reach_x, reach_x + 1) for b in range(-reach_y, reach_y + 1)]
    d2min = dmin * dmin
    for _ in range(max_rounds):
        if have >= n:
            break
        m = min(max(4 * (n - have), 1024), 4 * n + 1024)
        cand = rng.random((m, 2)) * np.array([Lx, Ly])
        cx = np.minimum((cand[:, 0] / hx).astype(np.int64), gx - 1)
        cy = np.minimum((cand[:, 1] / hy).astype(np.int64), gy - 1)
        ok = grid[cx, cy] < 0
        # unique cell among the candidates (keep the first)
        key = cx * gy + cy
        _, first = np.unique(key, return_index=True)
        uniq = np.zeros(m, dtype=bool)
        uniq[first] = True
        ok &= uniq
        idx = np.nonzero(ok)[0]
        cand, cx, cy = cand[idx], cx[idx], cy[idx]
        # conflicts with accepted points
        bad = np.zeros(len(idx), dtype=bool)
        for a, b in offs:
            nx = (cx + a) % gx
            ny = (cy + b) % gy
            j = grid[nx, ny]
            has = j >= 0
            if not has.any():
                continue
            q = pts[np.where(has, j, 0)]
            dx = q[:, 0] - cand[:, 0]
            dy = q[:, 1] - cand[:, 1]
            dx -= Lx * np.round(dx / Lx)
            dy -= Ly * np.round(dy / Ly)
            bad |= has & (dx * dx + dy * dy < d2min)
        keep = ~bad
        cand, cx, cy = cand[keep], cx[keep], cy[keep]
        # conflicts among the candidates: provisional grid, later index loses
        tmp = -np.ones((gx, gy), dtype=np.int64)
        tmp[cx, cy] = np.arange(len(cand))
        lose = np.zeros(len(cand), dtype=bool)
        me = np.arange(len(cand))
        for a, b in offs:
            if a == 0 and b == 0:
                continue
            j = tmp[(cx + a) % gx, (cy + b) % gy]
            has = (j >= 0) & (j < me)
            if not has.any():
                continue
            q = cand[np.where(has, j, 0)]
            dx = q[:, 0] - cand[:, 0]
            dy = q[:, 1] - cand[:, 1]
            dx -= Lx * np.round(dx / Lx)
            dy -= Ly * np.round(dy / Ly)
            lose |= has & (dx * dx + dy * dy < d2min)
        cand, cx, cy = cand[~lose], cx[~lose], cy[~lose]
        take = min(len(cand), n - have)
        cand, cx, cy = cand[:take], cx[:take], cy[:take]
        pts[have:have + take] = cand
        grid[cx, cy] = np.arange(have, have + take)
        have += take
    if have < n:
        raise RuntimeError('random sequential addition did not place %d points' % n)
    return pts


def square_box(n, density):
    s = math.sqrt(n / density)
    return np.array([s, s])


# ---------------------------------------------------------------------------
# models in the reference's dict form (nonbonded_interactions.txt sections)
# ---------------------------------------------------------------------------

def lj1_model():
    inter = [dict(type='lj', pair=[0, 0], u0=1.0, r0=LJ_R0, n=6, rc=2.5)]
    return dict(interactions=inter, reactions=[], tau_by_tauD=[1.0])


def kalj_model():
    inter = [
        dict(type='lj', pair=[0, 0], u0=1.0, r0=LJ_R0 * 1.0, n=6, rc=2.5),
        dict(type='lj', pair=[0, 1], u0=1.5, r0=LJ_R0 * 0.8, n=6, rc=2.5),
        dict(type='lj', pair=[1, 1], u0=0.5, r0=LJ_R0 * 0.88, n=6, rc=2.5),
    ]
    return dict(interactions=inter, reactions=[], tau_by_tauD=[1.0, 1.0])


def wall_model(k_a=0.1, k_d=1.0):
    inter = [
        dict(type='lj', pair=[0, 0], u0=1.0, r0=LJ_R0, n=6, rc=LJ_R0),
        dict(type='lj', pair=[0, 1], u0=1.0, r0=LJ_R0, n=6, rc=LJ_R0),
        dict(type='lj', pair=[1, 1], u0=2.0, r0=LJ_R0, n=6, rc=2.5),
        dict(type='wall', particle_type=1, direction=1, u0=3.0, r0=1.0, a=0.5, rc=3.0),
    ]
    rx = [
        dict(type='linear', rate=k_a, reactant=0, product=1),
        dict(type='locally_catalyzed', rate=k_d, r0=1.5, reactant=1, catalyst=1, product=0),
    ]
    return dict(interactions=inter, reactions=rx, tau_by_tauD=[1.0, 1.0])


def patchy_model():
    inter = [
        dict(type='yukawa', pair=[0, 0], u0=10.0, r0=0.2, rc=2.0),
        dict(type='tanh', pair=[2, 2], u0=5.0, r0=0.3, a=0.05, rc=0.6),
    ]
    rx = [
        dict(type='linear', rate=0.05, reactant=1, product=2),
        dict(type='locally_catalyzed', rate=0.5, r0=0.3, reactant=2, catalyst=2, product=1),
    ]
    return dict(interactions=inter, reactions=rx, tau_by_tauD=[1.0, 1.0, 1.0])


# ---------------------------------------------------------------------------
# systems: dict(x, types, molecule_ids, L, periodic, dt, model, bonds, ...)
# ---------------------------------------------------------------------------

def _atoms(x, types):
    n = len(x)
    mol = np.zeros((n, 3), dtype=np.intc)
    mol[:, 0] = np.arange(n)
    return mol


def lj1(n=1024, density=0.5, seed=1001):
    L = square_box(n, density)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rsa_positions(n, L, 0.9, rng)
    types = np.zeros(n, dtype=np.intc)
    return dict(name='lj1', x=x, types=types, molecule_ids=_atoms(x, types), L=L,
                periodic=np.array([1.0, 1.0]), dt=1e-4, n_steps=2000, model=lj1_model(),
                bonds=None)


def kalj(n=262144, density=0.6, seed=1002, dt=5e-5):
    L = square_box(n, density)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rsa_positions(n, L, 0.9, rng)
    types = np.zeros(n, dtype=np.intc)
    types[rng.permutation(n)[: n // 2]] = 1
    return dict(name='kalj', x=x, types=types, molecule_ids=_atoms(x, types), L=L,
                periodic=np.array([1.0, 1.0]), dt=dt, n_steps=10000, model=kalj_model(),
                bonds=None)


def wall(n=65536, density=0.4, seed=1003, k_a=0.1, k_d=1.0, dt=1e-4):
    L = square_box(n, density)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rsa_positions(n, L, 0.9, rng)
    # the box is reflecting in y: keep particles off the walls
    x[:, 1] = np.clip(x[:, 1], 1e-6, L[1] - 1e-6)
    types = np.zeros(n, dtype=np.intc)
    return dict(name='wall', x=x, types=types, molecule_ids=_atoms(x, types), L=L,
                periodic=np.array([1.0, -1.0]), dt=dt, n_steps=20000,
                model=wall_model(k_a, k_d), bonds=None)


def patchy(n_mol=16384, packing=0.3, seed=1004, dt=1e-5):
    s = math.sqrt(n_mol * math.pi / 4.0 / packing)
    L = np.array([s, s])
    rng = np.random.Generator(np.random.PCG64(seed))
    cores = rsa_positions(n_mol, L, 1.05, rng)
    th = rng.random(n_mol) * 2 * math.pi
    x = np.zeros((4 * n_mol, 2))
    types = np.zeros(4 * n_mol, dtype=np.intc)
    mol = np.zeros((4 * n_mol, 3), dtype=np.intc)
    for k in range(4):
        sl = slice(k, 4 * n_mol, 4)
        if k == 0:
            x[sl] = cores
        else:
            a = th + 2 * math.pi * (k - 1) / 3.0
            x[sl] = cores + 0.5 * np.stack([np.cos(a), np.sin(a)], axis=1)
            types[sl] = 1
        mol[sl, 0] = np.arange(n_mol)
        mol[sl, 1] = k
    x %= L
    # bonds: core-patch r0 = 0.5, patch-patch r0 = 0.5*sqrt(3); U0 = 1000
    per = [(0, 1, 0.5), (0, 2, 0.5), (0, 3, 0.5),
           (1, 2, 0.5 * math.sqrt(3)), (2, 3, 0.5 * math.sqrt(3)), (3, 1, 0.5 * math.sqrt(3))]
    bi, bj, br, bu = [], [], [], []
    base = 4 * np.arange(n_mol)
    for a, b, r0 in per:
        bi.append(base + a)
        bj.append(base + b)
        br.append(np.full(n_mol, r0))
        bu.append(np.full(n_mol, 1000.0))
    bonds = (np.concatenate(bi), np.concatenate(bj), np.concatenate(br), np.concatenate(bu))
    molecule_bonds = [dict(type='bond', participants=[a, b], r0=r0, u0=1000.0)
                      for a, b, r0 in per]
    return dict(name='patchy', x=x, types=types, molecule_ids=mol, L=L,
                periodic=np.array([1.0, 1.0]), dt=dt, n_steps=10000, model=patchy_model(),
                bonds=bonds, molecule_bonds=molecule_bonds)


def ensemble_plan(n_rep=512, seed=1005, n_min=2048, n_max=32768):
    """Per replica: (kind, n, k_a, k_d, seed); half kalj, half wall."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ns = np.round(np.exp(rng.uniform(math.log(n_min), math.log(n_max), n_rep))).astype(int)
    ka = np.exp(rng.uniform(math.log(0.01), math.log(1.0), n_rep))
    kd = np.exp(rng.uniform(math.log(0.01), math.log(1.0), n_rep))
    plan = []
    for r in range(n_rep):
        kind = 'kalj' if r < n_rep // 2 else 'wall'
        plan.append(dict(kind=kind, n=int(ns[r]), k_a=float(ka[r]), k_d=float(kd[r]),
                         seed=(seed << 20) + r))
    return plan


def ensemble_systems(plan, seed=1005):
    out = []
    for r, p in enumerate(plan):
        if p['kind'] == 'kalj':
            s = kalj(n=p['n'], seed=seed * 1000 + r)
        else:
            s = wall(n=p['n'], seed=seed * 1000 + r, k_a=p['k_a'], k_d=p['k_d'])
        s['philox_seed'] = p['seed']
        out.append(s)
    return out


BY_NAME = {'lj1': lj1, 'kalj': kalj, 'wall': wall, 'patchy': patchy}


# ---------------------------------------------------------------------------
# three-dimensional systems (the reference's N_dim = 3)
# ---------------------------------------------------------------------------

def lattice_positions3(n, L, rng, jitter=0.1):
    """n points in the cube [0, L)^3: n random sites of the smallest simple
    cubic lattice with >= n sites, each moved by U(-jitter, jitter) spacings
    per axis; pair distances are >= (1 - 2 jitter) spacings."""
    m = int(math.ceil(n ** (1.0 / 3.0) - 1e-9))
    while m ** 3 < n:
        m += 1
    a = float(L) / m
    site = rng.permutation(m ** 3)[:n]
    idx = np.stack([site % m, (site // m) % m, site // (m * m)], axis=1).astype(float)
    return (idx + 0.5 + rng.uniform(-jitter, jitter, (n, 3))) * a


def lj3(n=4096, density=0.6, seed=2001, dt=5e-5):
    """The two-species LJ mixture of ``kalj`` in a periodic cube."""
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = lattice_positions3(n, L, rng)
    types = np.zeros(n, dtype=np.intc)
    types[rng.permutation(n)[: n // 2]] = 1
    return dict(name='lj3', x=x, types=types, molecule_ids=_atoms(x, types),
                L=np.array([L, L, L]), periodic=np.array([1.0, 1.0, 1.0]), dt=dt,
                n_steps=10000, model=kalj_model(), bonds=None)


def wall3(n=4096, density=0.4, seed=2002, k_a=0.1, k_d=1.0, dt=1e-4):
    """The ``wall`` model in a cube periodic in x and z and reflecting in y
    (the wall well at y = 0)."""
    L = (n / density) ** (1.0 / 3.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    x = lattice_positions3(n, L, rng)
    types = np.zeros(n, dtype=np.intc)
    return dict(name='wall3', x=x, types=types, molecule_ids=_atoms(x, types),
                L=np.array([L, L, L]), periodic=np.array([1.0, -1.0, 1.0]), dt=dt,
                n_steps=20000, model=wall_model(k_a, k_d), bonds=None)


def patchy3(n_mol=1024, density=0.15, seed=2003, dt=1e-5, angle_k0=50.0):
    """The ``patchy`` molecules in a periodic cube: three patches at radius 0.5
    and 120 degrees in a random plane through the core, the six springs of
    ``patchy``, and three patch-core-patch angles of 120 degrees (k0 per
    rad^2) in definition order."""
    L = (n_mol / density) ** (1.0 / 3.0)
    rng = np.random.Generator(np.random.PCG64(seed))
    cores = lattice_positions3(n_mol, L, rng)
    # an orthonormal pair (u, v) per molecule: the patch plane
    g = rng.standard_normal((n_mol, 3))
    u = g / np.linalg.norm(g, axis=1, keepdims=True)
    h = rng.standard_normal((n_mol, 3))
    h -= (h * u).sum(axis=1, keepdims=True) * u
    v = h / np.linalg.norm(h, axis=1, keepdims=True)
    x = np.zeros((4 * n_mol, 3))
    types = np.zeros(4 * n_mol, dtype=np.intc)
    mol = np.zeros((4 * n_mol, 3), dtype=np.intc)
    for k in range(4):
        sl = slice(k, 4 * n_mol, 4)
        if k == 0:
            x[sl] = cores
        else:
            a = 2 * math.pi * (k - 1) / 3.0
            x[sl] = cores + 0.5 * (math.cos(a) * u + math.sin(a) * v)
            types[sl] = 1
        mol[sl, 0] = np.arange(n_mol)
        mol[sl, 1] = k
    x %= L
    per = [(0, 1, 0.5), (0, 2, 0.5), (0, 3, 0.5),
           (1, 2, 0.5 * math.sqrt(3)), (2, 3, 0.5 * math.sqrt(3)), (3, 1, 0.5 * math.sqrt(3))]
    base = 4 * np.arange(n_mol)
    bonds = tuple(np.concatenate(c) for c in zip(*[
        (base + a, base + b, np.full(n_mol, r0), np.full(n_mol, 1000.0)) for a, b, r0 in per]))
    tri = [(1, 0, 2), (2, 0, 3), (3, 0, 1)]
    angles = tuple(np.concatenate(c) for c in zip(*[
        (base + i, base + j, base + k, np.full(n_mol, 2 * math.pi / 3),
         np.full(n_mol, float(angle_k0))) for i, j, k in tri]))
    return dict(name='patchy3', x=x, types=types, molecule_ids=mol,
                L=np.array([L, L, L]), periodic=np.array([1.0, 1.0, 1.0]), dt=dt,
                n_steps=10000, model=patchy_model(), bonds=bonds, angles=angles)
