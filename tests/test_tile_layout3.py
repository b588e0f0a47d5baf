"""The index arithmetic of the three-dimensional force tiles, restated on the
CPU and checked against brute force (no GPU).

`k3_tile` (pyrdsim_amd/csrc/rds_engine.hip) works on tiles of T3Y x T3Z home
cell rows (y, z) by W cells in x, built by `build_tiles3`; `tile_setup3`
turns the cell starts into per-row offsets, image segments and a home-number
table, `tile_stage3` maps each staged index to a slot and a periodic shift,
and each lane walks nine row windows of three cells.  This file restates
those steps in Python, line for line in their arithmetic, and checks on
random grids (periodic and reflecting axes, partial tiles, several LDS
rounds) that every index stays inside its buffer, that every particle is a
home particle of exactly one tile, and that each home particle's windows
hold exactly the slots of its 27-cell stencil with the right images.
"""

import itertools

import numpy as np

T3Y = T3Z = 2
T3R = (T3Y + 2) * (T3Z + 2)
MAXW3 = 16


def _tiles(lx, ly, lz, occ, target):
    """build_tiles3: (ny, nz, y0, z0, x0, W) per tile."""
    W = int(np.floor(target / max(occ, 1e-9)))
    W = max(1, min(W, min(MAXW3, lx)))
    nstrip = (lx + W - 1) // W
    Wq, Wr = lx // nstrip, lx % nstrip
    out = []
    for z in range(0, lz, T3Z):
        for y in range(0, ly, T3Y):
            x0 = 0
            for k in range(nstrip):
                wk = Wq + (1 if k < Wr else 0)
                out.append((min(T3Y, ly - y), min(T3Z, lz - z), y, z, x0, wk))
                x0 += wk
    return out


def _check_grid(lx, ly, lz, per, target, cap, rng):
    sx, sy, sz = [int(p and n >= 3) for p, n in zip(per, (lx, ly, lz))]
    L = np.array([float(lx), float(ly), float(lz)])
    ncell = lx * ly * lz
    cnt = rng.poisson(rng.uniform(0.5, 12), ncell)
    cstart = np.concatenate([[0], np.cumsum(cnt)]).astype(int)
    N = int(cstart[-1])
    cell_of_slot = np.repeat(np.arange(ncell), cnt)
    homes_seen = np.zeros(N, int)
    for (ny, nz, y0, z0, x0, W) in _tiles(lx, ly, lz, N / ncell, target):
        # tile_setup3: cell starts and counts of the staged cells
        cs = np.zeros((T3R, W + 2), int)
        off = np.zeros((T3R, W + 3), int)
        for r in range(T3R):
            qy, qz = r % (T3Y + 2), r // (T3Y + 2)
            for k in range(W + 2):
                ok = qy <= ny + 1 and qz <= nz + 1
                y, z, x = y0 + qy - 1, z0 + qz - 1, x0 - 1 + k
                if y < 0:
                    ok, y = ok and sy, y + ly
                elif y >= ly:
                    ok, y = ok and sy, y - ly
                if z < 0:
                    ok, z = ok and sz, z + lz
                elif z >= lz:
                    ok, z = ok and sz, z - lz
                if x < 0:
                    ok, x = ok and sx, x + lx
                elif x >= lx:
                    ok, x = ok and sx, x - lx
                if ok:
                    c = (z * ly + y) * lx + x
                    assert 0 <= c < ncell
                    cs[r, k] = cstart[c]
                    off[r, k + 1] = cstart[c + 1] - cstart[c]
        off[:, 1:] = np.cumsum(off[:, 1:], axis=1)
        # segments (left image, in-range run, right image) and row shifts
        seg, row_sh = [], []
        for r in range(T3R):
            qy, qz = r % (T3Y + 2), r // (T3Y + 2)
            y, z = y0 + qy - 1, z0 + qz - 1
            row_sh.append((-L[1] if y < 0 else (L[1] if y >= ly else 0.0),
                           -L[2] if z < 0 else (L[2] if z >= lz else 0.0)))
            sg = [[0, 0, 0, 0.0] for _ in range(3)]  # first slot, n, staged offset, x shift
            if x0 == 0 and sx:
                sg[0] = [cs[r, 0], off[r, 1], 0, -L[0]]
            ka = 1 if x0 == 0 else 0
            kb = W if x0 + W >= lx else W + 1
            sg[1] = [cs[r, ka], off[r, kb + 1] - off[r, ka], off[r, ka], 0.0]
            if x0 + W >= lx and sx:
                sg[2] = [cs[r, W + 1], off[r, W + 2] - off[r, W + 1], off[r, W + 1], L[0]]
            seg.append(sg)
        row_base = np.concatenate([[0], np.cumsum(off[:, W + 2])])
        total = int(row_base[-1])
        nc = ny * nz * W
        assert nc <= 64   # one lane per home cell in the home-cell scan
        hc = np.zeros(nc, int)
        for i in range(nc):
            hr, k = divmod(i, W)
            r = (hr // ny + 1) * (T3Y + 2) + (hr % ny + 1)
            hc[i] = off[r, k + 2] - off[r, k + 1]
        chp = np.concatenate([[0], np.cumsum(hc)])
        # tile_stage3: staged index -> (slot, shifts); row by binary search
        staged = []
        for p in range(total):
            r = 0
            for step in (8, 4, 2, 1):
                if row_base[r + step] <= p:
                    r += step
            loc = p - row_base[r]
            g = 0
            if seg[r][1][1] > 0 and loc >= seg[r][1][2]:
                g = 1
            if seg[r][2][1] > 0 and loc >= seg[r][2][2]:
                g = 2
            src = seg[r][g][0] + loc - seg[r][g][2]
            assert 0 <= src < N
            staged.append((int(src), seg[r][g][3], row_sh[r][0], row_sh[r][1]))
        # each home particle: its cell, slot and nine windows
        for g in range(int(chp[-1])):
            lo_i, hi_i = 0, nc
            while hi_i - lo_i > 1:
                mid = (lo_i + hi_i) // 2
                if chp[mid] <= g:
                    lo_i = mid
                else:
                    hi_i = mid
            hr, k = divmod(lo_i, W)
            rh = (hr // ny + 1) * (T3Y + 2) + (hr % ny + 1)
            s = int(cs[rh, k + 1] + g - chp[hr * W + k])
            assert 0 <= s < N
            homes_seen[s] += 1
            got = []
            for d in range(9):
                rr = rh + (d // 3 - 1) * (T3Y + 2) + (d % 3 - 1)
                assert 0 <= rr < T3R
                lo = row_base[rr] + off[rr, k]
                hi = row_base[rr] + off[rr, k + 3]
                assert 0 <= lo <= hi <= total
                for p0 in range(0, max(total, 1), cap):
                    a, b = max(lo, p0) - p0, min(hi, p0 + cap) - p0
                    # the pair loop reads one past a window's end: the sentinel
                    # slot min(total - p0, cap) or a staged particle
                    assert b <= min(total - p0, cap)
                    got += [staged[p0 + j] for j in range(a, b)]
            # brute force: the 27 cells around the home cell, with images
            hz, rem = divmod(int(cell_of_slot[s]), lx * ly)
            hy, hx = divmod(rem, lx)
            want = []
            for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
                v = [hx + dx, hy + dy, hz + dz]
                sh, ok = [0.0, 0.0, 0.0], True
                for ax, (n, flag) in enumerate(((lx, sx), (ly, sy), (lz, sz))):
                    if v[ax] < 0 or v[ax] >= n:
                        ok = ok and bool(flag)
                        sh[ax] = -L[ax] if v[ax] < 0 else L[ax]
                if not ok:
                    continue
                c = ((v[2] % lz) * ly + (v[1] % ly)) * lx + (v[0] % lx)
                want += [(j, sh[0], sh[1], sh[2]) for j in range(cstart[c], cstart[c + 1])]
            assert sorted(got) == sorted(want)
    assert (homes_seen == 1).all()


def test_tile_windows_cover_the_27_cell_stencil():
    rng = np.random.Generator(np.random.PCG64(1))
    for _ in range(60):
        per = [bool(rng.integers(0, 2)) for _ in range(3)]
        dims = [int(rng.integers(1, 12)) for _ in range(3)]
        # a periodic axis with a stencil has at least three cells
        dims = [max(d, 3) if p else d for d, p in zip(dims, per)]
        _check_grid(*dims, per, int(rng.choice([8, 48, 200])), int(rng.choice([64, 256, 100000])),
                    rng)
