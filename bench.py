"""Benchmark of the step loop on one MI355X.

Headline (`value`): particle-steps/s of the batched ensemble -- 512 independent
replicas (256 of the two-species LJ mixture, 256 of the wall + reaction model,
N per replica log-uniform over 2,048..32,768, sum N = 5,789,872) stepped
together with FP32 pair forces and FP64 positions, hipGraph-captured.  Beside it:

* `plain`: the single-trajectory FP64 path, one replica at a time, on a
  stratified subset of the same replicas, and the ratio batched / plain;
* `roofline`: the fused force + integrate + bin kernel's average duration from
  HIP events on the engine's stream during an eager pass over the same state,
  against its algorithmic HBM bytes (48 B per particle + 4 B per cell);
* `cpu_baseline`: the CPU oracle in reference-faithful mode (Verlet list on the
  reference's grid, FP64, one core) on a bounded sample of the same replicas;
* `analysis`: the RDF pair histogram, type counts and cluster labels of the
  whole ensemble on the device, beside the host RDF of one replica.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pyrdsim_amd import workloads  # noqa: E402
from pyrdsim_amd.autotune import autotune  # noqa: E402

HBM_PEAK_GBS = 8000.0
FORCE_BYTES_PER_PARTICLE = 48   # read pos 16 + type 4 + pid 4; write pos 16 + cell 4 + rank 4
FORCE_BYTES_PER_CELL = 4        # cell starts
# list mode: read pos 16 + type 4 + pid 4 + reference pos 16 + sorted cell 4 + list length 2,
# write pos 16; plus 2 bytes per list entry
LIST_BYTES_PER_PARTICLE = 62
LIST_BYTES_PER_ENTRY = 2


def _make_replica_state(args):
    kind, n, k_a, k_d, seed, ic_seed = args
    if kind == 'kalj':
        s = workloads.kalj(n=n, seed=ic_seed)
    else:
        s = workloads.wall(n=n, seed=ic_seed, k_a=k_a, k_d=k_d)
    return s['x'], s['types']


def build_ensemble(n_rep, workers, cache=None):
    """Replica plan and initial states; `cache` (an .npz path) is read when it
    exists and written otherwise, so that profiled runs start without forking
    worker processes."""
    plan = workloads.ensemble_plan(n_rep=n_rep)
    if cache and os.path.exists(cache):
        with np.load(cache, allow_pickle=False) as z:
            x, t, n = z['x'], z['t'], z['n']
        if len(n) == len(plan) and all(int(a) == p['n'] for a, p in zip(n, plan)):
            off = np.concatenate([[0], np.cumsum(n)])
            return plan, [(x[off[r]:off[r + 1]], t[off[r]:off[r + 1]]) for r in range(len(plan))]
    jobs = [(p['kind'], p['n'], p['k_a'], p['k_d'], p['seed'], 1005 * 1000 + r)
            for r, p in enumerate(plan)]
    if workers > 1:
        ctx = mp.get_context('fork')
        with ctx.Pool(workers) as pool:
            states = pool.map(_make_replica_state, jobs, chunksize=4)
    else:
        states = [_make_replica_state(j) for j in jobs]
    if cache:
        np.savez(cache, x=np.concatenate([s[0] for s in states]),
                 t=np.concatenate([s[1] for s in states]),
                 n=np.array([p['n'] for p in plan]))
    return plan, states


def engine_pieces(plan):
    from pyrdsim_amd.engine import build_model, make_replica
    mk = workloads.kalj_model()
    mw = workloads.wall_model()
    kalj_model, _ = build_model(mk['interactions'], mk['reactions'], 2)
    wall_model, _ = build_model(mw['interactions'], mw['reactions'], 2)
    L_k = lambda n: workloads.square_box(n, 0.6)
    L_w = lambda n: workloads.square_box(n, 0.4)
    reps = []
    for p in plan:
        if p['kind'] == 'kalj':
            reps.append(make_replica(0, p['n'], L_k(p['n']), [1, 1], 5e-5, [1.0, 1.0], [],
                                     p['seed']))
        else:
            reps.append(make_replica(1, p['n'], L_w(p['n']), [1, -1], 1e-4, [1.0, 1.0],
                                     [p['k_a'], p['k_d']], p['seed']))
    return [kalj_model, wall_model], reps


def time_engine(eng, steps, warmup):
    import torch
    eng.run(warmup)
    eng.sync()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.run(steps)
    eng.sync()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def accuracy_leg(plan, states, models, reps, x0, t0s, graph_steps, tiling, steps, n_check,
                 chunk=1000):
    """Statistical accuracy of the batched FP32-force runs (mixed: FP64
    positions; compensated: float-pair positions) against the FP64 path over
    `steps` steps with the same Philox streams from the same states: a
    stratified subset of replicas is re-run alone in FP64.  The RDF and the
    activated fraction are averaged over snapshots every `chunk` steps in the
    second half of the run; the largest position deviation is taken after the
    first chunk (later the trajectories decorrelate, as chaotic dynamics do)."""
    from pyrdsim_amd import analysis
    from pyrdsim_amd.engine import Engine
    off = np.concatenate([[0], np.cumsum([p['n'] for p in plan])])
    picks = []
    for kind in ('kalj', 'wall'):
        idx = sorted((r for r, p in enumerate(plan) if p['kind'] == kind),
                     key=lambda r: plan[r]['n'])
        m = max(1, n_check // 2)
        picks += [idx[int((i + 0.5) * len(idx) / m)] for i in range(m)]
    n_chunks = max(2, steps // chunk)
    second = range(n_chunks // 2, n_chunks)
    geo = {r: ([reps[r].L[0], reps[r].L[1]], [reps[r].periodic[0], reps[r].periodic[1]])
           for r in picks}

    def observe(acc, r, x, t, k):
        L, per = geo[r]
        if k == 0:
            acc.setdefault('x1', {})[r] = x.copy()
        if k in second:
            _, g = analysis.rdf(x, L, per)
            acc.setdefault('g', {}).setdefault(r, []).append(g)
            acc.setdefault('a', {}).setdefault(r, []).append(analysis.activated_fraction(t))

    ref = {}
    for r in picks:
        e = Engine([models[reps[r].model]], [_with_model0(reps[r])], precision='fp64',
                   graph_steps=graph_steps)
        e.set_state(states[r][0], states[r][1])
        for k in range(n_chunks):
            e.run(chunk)
            x, t = e.state()
            observe(ref, r, x, t, k)
        e.close()
    out = dict(replicas=len(picks), steps=int(n_chunks * chunk), snapshots=len(second),
               reference='fp64 single replica, same streams')
    for prec in ('mixed', 'compensated'):
        e = Engine(models, reps, precision=prec, graph_steps=graph_steps, tile_rows=tiling[0],
                   tile_particles=tiling[1])
        e.set_state(x0, t0s)
        acc = {}
        for k in range(n_chunks):
            e.run(chunk)
            xa, ta = e.state()
            for r in picks:
                observe(acc, r, xa[off[r]:off[r + 1]], ta[off[r]:off[r + 1]], k)
        e.close()
        rdf_err, act_err, dev = [], [], []
        for r in picks:
            L, per = geo[r]
            rdf_err.append(analysis.rdf_error(np.mean(acc['g'][r], axis=0),
                                              np.mean(ref['g'][r], axis=0)))
            act_err.append(abs(np.mean(acc['a'][r]) - np.mean(ref['a'][r])))
            d = acc['x1'][r] - ref['x1'][r]
            for q in range(2):
                if per[q] > 0:
                    d[:, q] -= L[q] * np.round(d[:, q] / L[q])
            dev.append(float(np.abs(d).max()))
        out[prec] = {'rdf_l1_rel_max': float(max(rdf_err)),
                     'activated_fraction_abs_max': float(max(act_err)),
                     'position_dev_max_after_%d_steps' % chunk: float(max(dev))}
    out['rdf_l1_rel_max'] = out['mixed']['rdf_l1_rel_max']
    out['activated_fraction_abs_max'] = out['mixed']['activated_fraction_abs_max']
    return out


def analysis_leg(eng, plan, reps):
    """RDF pair histogram (r <= 5, bins of 0.02), type counts and cluster
    labels (active particles, links <= 1.5) of the whole ensemble on the
    device, wall time including the copy of the results to the host; the
    same RDF on the host (cKDTree) for the median-size replica, per particle."""
    from pyrdsim_amd import analysis
    eng.pair_histogram(5.0, 0.02)
    t = time.perf_counter()
    hist = eng.pair_histogram(5.0, 0.02)
    t_rdf = time.perf_counter() - t
    t = time.perf_counter()
    eng.type_counts()
    t_cnt = time.perf_counter() - t
    t = time.perf_counter()
    eng.cluster_labels(1.5, types=(1,))
    t_cl = time.perf_counter() - t
    x, _ = eng.state()
    n = np.array([p['n'] for p in plan])
    off = np.concatenate([[0], np.cumsum(n)])
    r = int(np.argsort(n, kind='stable')[len(n) // 2])
    rep = reps[r]
    L = np.array(rep.L[:])
    per = [1 if rep.periodic[0] else -1, 1 if rep.periodic[1] else -1]
    t = time.perf_counter()
    analysis.rdf(x[off[r]:off[r + 1]], L, per, 5.0, 0.02)
    t_host = time.perf_counter() - t
    n_tot = int(n.sum())
    return dict(particles=n_tot, pairs=int(hist.sum()), rdf_device_ms=round(1e3 * t_rdf, 3),
                type_counts_device_ms=round(1e3 * t_cnt, 3),
                clusters_device_ms=round(1e3 * t_cl, 3),
                rdf_host_ms_per_1e6_particles=round(1e3 * t_host / n[r] * 1e6, 1),
                host_sample='analysis.rdf on replica %d (N=%d), one core' % (r, int(n[r])))


def measure_traffic(cache, rows, per, replicas, timeout=300):
    """HBM bytes per force-kernel launch from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over a few eager steps of this workload in a
    child process, each scaled by what the same counter reads for an
    elementwise kernel that reads and writes 256 MiB (--copy-calibration).
    Run before this process touches the GPU.  None when rocprofv3 is missing
    or a pass fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which('rocprofv3')
    if not prof:
        return None
    tmp = tempfile.mkdtemp(prefix='rds_traffic_')
    child = [sys.executable, os.path.abspath(__file__), '--steps', '6', '--warmup', '2',
             '--graph-steps', '0', '--replicas', str(replicas), '--no-plain', '--no-cpu',
             '--no-accuracy', '--no-autotune', '--no-traffic', '--tile-rows', str(rows),
             '--tile-particles', str(per), '--copy-calibration', '--ic-cache', cache]
    out = {}
    try:
        for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
            d = os.path.join(tmp, counter)
            cmd = [prof, '--pmc', counter, '--output-format', 'csv', '-d', d, '-o', 'c', '--'] + child
            subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=timeout)
            files = glob.glob(os.path.join(d, '**', 'c_counter_collection.csv'), recursive=True)
            if not files:
                return None
            with open(files[0]) as f:
                recs = [r for r in csv.DictReader(f) if r['Counter_Name'] == counter]
            force = [float(r['Counter_Value']) for r in recs if 'k_force' in r['Kernel_Name']]
            cal = [float(r['Counter_Value']) for r in recs if 'AUnaryFunctor' in r['Kernel_Name']]
            if not force or not cal or cal[0] <= 0:
                return None
            factor = cal[0] / (268435456 / 1024)       # KiB counted per KiB moved
            out[counter] = float(np.mean(force)) * 1024 / factor
    except (subprocess.SubprocessError, OSError, KeyError, ValueError):
        return None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


def cpu_baseline(plan, states, budget_s):
    """The oracle, reference-faithful (Verlet list, rs = 2 rc, ratio 2, refresh
    every 10 steps, FP64, one core), on the smallest replica of each model."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    from oracle import Oracle
    picks = []
    for kind in ('kalj', 'wall'):
        idx = [r for r, p in enumerate(plan) if p['kind'] == kind]
        picks.append(min(idx, key=lambda r: plan[r]['n']))
    done = 0
    t_total = 0.0
    desc = []
    for r in picks:
        p = plan[r]
        x, t = states[r]
        if p['kind'] == 'kalj':
            m = workloads.kalj_model()
            L, per, dt = workloads.square_box(p['n'], 0.6), [1, 1], 5e-5
        else:
            m = workloads.wall_model(p['k_a'], p['k_d'])
            L, per, dt = workloads.square_box(p['n'], 0.4), [1, -1], 1e-4
        o = Oracle(x, t, L, per, dt, m['tau_by_tauD'], m['interactions'], m['reactions'],
                   seed=p['seed'], rng='mt19937', nl='faithful', ratio=2, nl_freq=10)
        o.init()
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s / len(picks):
            o.run(steps + 1, 10)
            steps += 10
        el = time.perf_counter() - t0
        done += p['n'] * steps
        t_total += el
        desc.append('%s N=%d x %d steps' % (p['kind'], p['n'], steps))
    return done / t_total, '; '.join(desc) + ' (%.1f s)' % t_total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--replicas', type=int, default=512)
    ap.add_argument('--plain-subset', type=int, default=32)
    ap.add_argument('--plain-steps', type=int, default=500)
    ap.add_argument('--cpu-seconds', type=float, default=12.0)
    ap.add_argument('--graph-steps', type=int, default=50)
    ap.add_argument('--tile-particles', type=int, default=0)
    ap.add_argument('--tile-rows', type=int, default=1)
    ap.add_argument('--skin', type=float, default=0.0,
                    help='Verlet skin of the batched engine (0: cell search every step)')
    ap.add_argument('--ic-cache', default=None, help='.npz file of initial states (read or written)')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-autotune', dest='autotune', action='store_false')
    ap.add_argument('--no-accuracy', action='store_true')
    ap.add_argument('--accuracy-replicas', type=int, default=8)
    ap.add_argument('--accuracy-steps', type=int, default=10000)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run one elementwise kernel of known traffic (256 MiB read + 256 MiB '
                         'written) first, to scale HBM counters in a profile of this command')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-analysis', action='store_true')
    ap.add_argument('--no-traffic', dest='traffic', action='store_false',
                    help='skip the rocprofv3 counter passes for roofline.traffic')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    cache = args.ic_cache
    own_cache = None
    if args.traffic and not cache:
        import tempfile
        fd, own_cache = tempfile.mkstemp(prefix='rds_ic_', suffix='.npz')
        os.close(fd)
        os.unlink(own_cache)
        cache = own_cache
    plan, states = build_ensemble(args.replicas, min(16, os.cpu_count() or 1), cache)
    print('# built %d replica states in %.1f s' % (len(plan), time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    traffic = None
    # the counter passes use the tile shape the autotuner picks for this
    # workload (4 rows x 112 particles), unless one is given
    traffic_tiling = (args.tile_rows if args.tile_rows > 1 else 4,
                      args.tile_particles if args.tile_particles > 0 else 112)
    if args.traffic:
        traffic = measure_traffic(cache, traffic_tiling[0], traffic_tiling[1], args.replicas)
        print('# traffic: %s' % traffic, file=sys.stderr, flush=True)
    if own_cache and os.path.exists(own_cache):
        os.unlink(own_cache)

    import torch
    from pyrdsim_amd.engine import Engine
    if args.copy_calibration:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device='cuda')
        torch.cuda.synchronize()
        b = a * 2.0          # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del a, b
        print('# copy calibration: 268435456 B read + 268435456 B written', file=sys.stderr)
    models, reps = engine_pieces(plan)
    n_tot = sum(p['n'] for p in plan)
    x0 = np.concatenate([s[0] for s in states])
    t0s = np.concatenate([s[1] for s in states])

    eng = Engine(models, reps, precision='mixed', graph_steps=args.graph_steps,
                 tile_particles=args.tile_particles, tile_rows=args.tile_rows, skin=args.skin)
    eng.set_state(x0, t0s)
    tuned = None
    if args.autotune:
        best, times = autotune(eng)
        tuned = dict(best=list(best), force_ms={'%dx%d' % k: round(v, 5) for k, v in times.items()})
        print('# autotune: rows x particles = %s' % (best,), file=sys.stderr, flush=True)
    st0 = eng.sync()
    secs = time_engine(eng, args.steps, args.warmup)
    st1 = eng.sync()
    value = n_tot * args.steps / secs
    print('# batched: %.3e particle-steps/s (%.3f ms/step)' % (value, 1e3 * secs / args.steps),
          file=sys.stderr, flush=True)

    kt = eng.time_kernels(20)
    eng.sync()
    st2 = eng.sync()
    _, n_cells, n_tiles = eng.sizes()
    force_ms = kt['force'][0]
    lists = None
    if args.skin > 0:
        mean_len = st2.list_entries / max(1, st2.rebuilds) / n_tot
        force_bytes = (LIST_BYTES_PER_PARTICLE + LIST_BYTES_PER_ENTRY * mean_len) * n_tot + \
            FORCE_BYTES_PER_CELL * n_cells
        lists = dict(skin=args.skin, mean_list_length=round(mean_len, 2),
                     steps_per_rebuild=round((args.steps + args.warmup) /
                                             max(1, st1.rebuilds - st0.rebuilds), 2),
                     overflow=int(st2.list_overflow))
    else:
        force_bytes = FORCE_BYTES_PER_PARTICLE * n_tot + FORCE_BYTES_PER_CELL * n_cells
    achieved = force_bytes / (force_ms * 1e-3) / 1e9
    traffic_bytes = int(traffic['FETCH_SIZE'] + traffic['WRITE_SIZE']) if traffic else None
    roofline = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_bytes,
                    traffic_read=int(traffic['FETCH_SIZE']) if traffic else None,
                    traffic_write=int(traffic['WRITE_SIZE']) if traffic else None,
                    traffic_tiling=list(traffic_tiling) if traffic else None,
                    kernel='k_force (pair forces + Euler-Maruyama + binning)',
                    algorithmic_bytes=int(force_bytes),
                    avg_ms={k: round(v[0], 5) for k, v in kt.items() if v[1]})
    tiling = eng.tiling()
    ana = None
    if not args.no_analysis:
        ana = analysis_leg(eng, plan, reps)
        print('# analysis: %s' % ana, file=sys.stderr, flush=True)
    del eng
    torch.cuda.synchronize()

    accuracy = None
    compensated = None
    batched_fp64 = None
    if not args.no_accuracy:
        # throughput of the same batched run with compensated FP32 positions
        ec = Engine(models, reps, precision='compensated', graph_steps=args.graph_steps,
                    tile_rows=tiling[0], tile_particles=tiling[1])
        ec.set_state(x0, t0s)
        secs_c = time_engine(ec, args.steps, args.warmup)
        compensated = dict(value=n_tot * args.steps / secs_c, unit='particle-steps/s',
                           ms_per_step=1e3 * secs_c / args.steps)
        del ec
        torch.cuda.synchronize()
        # the same batched run in FP64: separates the gain of batching from
        # that of FP32 pair arithmetic
        e64 = Engine(models, reps, precision='fp64', graph_steps=args.graph_steps,
                     tile_rows=tiling[0], tile_particles=tiling[1])
        e64.set_state(x0, t0s)
        secs_64 = time_engine(e64, args.steps, args.warmup)
        batched_fp64 = dict(value=n_tot * args.steps / secs_64, unit='particle-steps/s',
                            ms_per_step=1e3 * secs_64 / args.steps)
        del e64
        torch.cuda.synchronize()
        accuracy = accuracy_leg(plan, states, models, reps, x0, t0s, args.graph_steps, tiling,
                                args.accuracy_steps, args.accuracy_replicas)
        print('# accuracy: %s' % accuracy, file=sys.stderr, flush=True)

    plain = None
    if not args.no_plain:
        order = np.argsort([p['n'] for p in plan], kind='stable')
        k = min(args.plain_subset, len(plan))
        # stratified over size, alternating the two models
        pick = [int(order[int((i + 0.5) * len(order) / k)]) for i in range(k)]
        work = 0.0
        tsum = 0.0
        plain_tiling = (args.tile_rows, args.tile_particles)
        if args.autotune:
            # tuned once, on the median-size replica of the subset
            r = pick[len(pick) // 2]
            e1 = Engine([models[reps[r].model]], [_with_model0(reps[r])], precision='fp64')
            e1.set_state(states[r][0], states[r][1])
            plain_tiling, _ = autotune(e1, steps=32)
            e1.close()
        for r in pick:
            e1 = Engine([models[reps[r].model]], [_with_model0(reps[r])], precision='fp64',
                        graph_steps=args.graph_steps, tile_rows=plain_tiling[0],
                        tile_particles=plain_tiling[1])
            e1.set_state(states[r][0], states[r][1])
            dt = time_engine(e1, args.plain_steps, 20)
            work += plan[r]['n'] * args.plain_steps
            tsum += dt
            e1.close()
        plain = dict(value=work / tsum, unit='particle-steps/s', replicas=k,
                     steps=args.plain_steps, precision='fp64', graph_steps=args.graph_steps,
                     tiling=list(plain_tiling))
        print('# plain: %.3e particle-steps/s' % plain['value'], file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        v, desc = cpu_baseline(plan, states, args.cpu_seconds)
        cpu = dict(value=v, unit='particle-steps/s', cores=1, kind='port',
                   sample='oracle, reference-faithful FP64: ' + desc)

    out = {
        'metric': 'particle-steps/s (batched vs plain, 1 MI355X) at RDF error <=1%; % HBM roofline',
        'value': value,
        'unit': 'particle-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * secs / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32 forces + f64 positions',
        'data': 'synthetic (seeded random initial states)',
        'config': {
            'workload': 'ensemble: %d replicas (half LJ mixture rho=0.6 dt=5e-5, half '
                        'wall+reactions rho=0.4 dt=1e-4), N log-uniform 2048..32768, '
                        'sum N=%d' % (len(plan), n_tot),
            'precision': 'mixed', 'graph_steps': args.graph_steps,
            'tiling': list(tiling[:2]), 'lds_cap': tiling[2], 'autotune': tuned,
            'cells': int(n_cells), 'tiles': int(n_tiles), 'lists': lists},
        'roofline': roofline,
        'plain': plain,
        'batched_vs_plain': (value / plain['value']) if plain else None,
        'accuracy': accuracy,
        'compensated_positions': compensated,
        'batched_fp64': batched_fp64,
        'analysis': ana,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


def _with_model0(rep):
    import ctypes
    from pyrdsim_amd import _lib as L
    r = L.Replica()
    ctypes.pointer(r)[0] = rep
    r.model = 0
    return r


if __name__ == '__main__':
    main()
