"""Autotuner for the force kernel's work decomposition.

The tile shape -- cell rows per tile (one wavefront per row up to four rows,
four wavefronts for eight; the tile's home particles dealt to the wavefronts
64 at a time) and particles per row strip --
trades LDS staging overhead (bigger tiles stage fewer halo cells per particle)
against parallelism, LDS occupancy and lane utilisation.  The
best shape depends on the workload (density, replica sizes, GPU fill), so it is
measured: each candidate runs a few real steps on a clone of the engine (same
configuration, same current state) and the force kernel's HIP-event time
decides.  The engine itself is not stepped: its state, step counter and random
streams are as before, only its tiling changes.  Results are cached by a
workload signature in a small JSON file.
"""

import json
import os

CANDIDATES = ((1, 56), (1, 64), (2, 56), (2, 64), (4, 48), (4, 56), (4, 64), (4, 72), (4, 80),
              (4, 96), (4, 112), (4, 128), (8, 48), (8, 56), (8, 64))


def signature(engine):
    """Workload key: precision, replica count, total particles, cell count."""
    n, c, _ = engine.sizes()
    return '%s/%d/%d/%d' % (engine.precision, len(engine.replicas), n, c)


def autotune(engine, candidates=CANDIDATES, steps=4, cache=None):
    """Pick (tile_rows, tile_particles) for `engine` by timing steps of a
    clone of it; leaves the engine set to the best shape and otherwise
    unchanged.  Returns (best, {shape: ms})."""
    key = signature(engine)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table:
            best = tuple(table[key]['best'])
            engine.set_tiling(*best)
            return best, {tuple(map(int, k.split('x'))): v for k, v in table[key]['ms'].items()}
    times = {}
    scratch = engine.clone()
    try:
        for rows, per in candidates:
            scratch.set_tiling(rows, per)
            scratch.time_kernels(1)                 # first launch of this variant
            times[(rows, per)] = scratch.time_kernels(steps)['force'][0]
    finally:
        scratch.close()
    best = min(times, key=times.get)
    engine.set_tiling(*best)
    if cache:
        table[key] = dict(best=list(best), ms={'%dx%d' % k: v for k, v in times.items()})
        with open(cache, 'w') as f:
            json.dump(table, f, indent=1)
    return best, times
