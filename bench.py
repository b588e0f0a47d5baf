"""Benchmark of the step loop on one MI355X.

Headline (`value`): particle-steps/s of the batched ensemble -- 512 independent
replicas (256 of the two-species LJ mixture, 256 of the wall + reaction model,
N per replica log-uniform over 2,048..32,768, sum N = 5,789,872) stepped
together with FP32 pair forces and FP64 positions, hipGraph-captured.  The
ensemble is first advanced --pre-advance steps (default 2,000, untimed) from
its all-inactive start into its reacting regime, the force kernel's tile
shape is then tuned on that state (on clones of the engine), and the timed
window replays graphs of at most --steps steps.  Beside it:

* `plain`: the single-trajectory FP64 path, one replica at a time, on a
  stratified subset of the same replicas, and the ratio batched / plain;
* `roofline`: the fused force + integrate + bin kernel's average duration from
  HIP events on the engine's stream during an eager pass over the same state,
  against its algorithmic HBM bytes (44 B per particle + 4 B per cell);
* `cpu_baseline`: the CPU oracle in reference-faithful mode (Verlet list on the
  reference's grid, FP64, one core) on a bounded sample of the same replicas;
  `c1`: the one-species LJ system (N = 1,024, 2,000 steps, FP64) on the GPU
  and on the CPU oracle (median of 3);
* `analysis`: the RDF pair histogram, type counts and cluster labels of the
  whole ensemble on the device, beside the host RDF of one replica;
* `steady_state`: 300 further graphed steps timed right after the window, as
  a check that the window is representative.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from pyrdsim_amd import ensemble, workloads  # noqa: E402
from pyrdsim_amd.autotune import autotune  # noqa: E402

HBM_PEAK_GBS = 8000.0
FORCE_BYTES_PER_PARTICLE = 44   # read pos 16 + type 4 + pid 4; write pos 16 + rank 4
FORCE_BYTES_PER_CELL = 4        # cell starts
# list mode: read pos 16 + type 4 + pid 4 + reference pos 16 + sorted cell 4 + list length 2,
# write pos 16; plus 2 bytes per list entry
LIST_BYTES_PER_PARTICLE = 62
LIST_BYTES_PER_ENTRY = 2
# the force-only figure of the pair-force step (read pos 16 + type 4, write the
# force 8 per particle; 8 per cell), for comparison with the fused count above
FORCE_ONLY_BYTES_PER_PARTICLE = 28
FORCE_ONLY_BYTES_PER_CELL = 8


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
        e = Engine([models[reps[r].model]], [ensemble.with_model0(reps[r])], precision='fp64',
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


def measure_traffic(cache, state, rows, per, replicas, timeout=300):
    """HBM bytes per force-kernel launch from two rocprofv3 counter passes
    (FETCH_SIZE, then WRITE_SIZE) over this workload in a child process that
    starts from `state` (the main run's advanced state, saved by
    save_state): the mean over its last eager k_force launches, each pass
    scaled by what the same counter reads for an elementwise kernel that
    reads and writes 256 MiB (--copy-calibration).  None when rocprofv3 is
    missing or a pass fails."""
    import csv
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which('rocprofv3')
    if not prof:
        return None
    tmp = tempfile.mkdtemp(prefix='rds_traffic_')
    child = [sys.executable, os.path.abspath(__file__), '--steps', '6', '--warmup', '0',
             '--graph-steps', '0', '--pre-advance', '0', '--replicas', str(replicas),
             '--no-plain', '--no-cpu', '--no-accuracy', '--no-autotune', '--no-traffic',
             '--no-analysis', '--no-c1', '--no-3d', '--no-steady-state', '--tile-rows', str(rows),
             '--tile-particles', str(per), '--copy-calibration', '--ic-cache', cache,
             '--state', state]
    out = {}
    try:
        for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
            print('# counter pass %s' % counter, file=sys.stderr, flush=True)
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
            force = force[-16:]
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
    every 10 steps, FP64, one core), on the smallest replica of each model of
    the ensemble, about budget_s seconds in all."""
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
        sp = ensemble.replica_spec(p)
        m = sp['model']
        o = Oracle(x, t, sp['L'], sp['periodic'], sp['dt'], m['tau_by_tauD'], m['interactions'],
                   m['reactions'], seed=p['seed'], rng='mt19937', nl='faithful', ratio=2,
                   nl_freq=10)
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


def c1_leg(runs=3, steps=2000):
    """The one-species LJ system (N = 1,024, density 0.5, rc = 2.5, dt = 1e-4,
    2,000 steps, FP64): the plain GPU path (graphs of 50 steps, median of
    `runs`) and the oracle in reference-faithful mode on one core (median of
    `runs`), in particle-steps/s."""
    import torch
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    from oracle import Oracle
    s = workloads.lj1()
    m = s['model']
    n = len(s['x'])
    model, rates = build_model(m['interactions'], m['reactions'], 1)
    gpu = []
    for k in range(runs):
        rep = make_replica(0, n, s['L'], s['periodic'], s['dt'], m['tau_by_tauD'], rates, 2001)
        e = Engine([model], [rep], precision='fp64', graph_steps=50)
        e.set_state(s['x'], s['types'])
        gpu.append(n * steps / time_engine(e, steps, 50))
        e.close()
    torch.cuda.synchronize()
    cpu = []
    for k in range(runs):
        o = Oracle(s['x'], s['types'], s['L'], s['periodic'], s['dt'], m['tau_by_tauD'],
                   m['interactions'], m['reactions'], seed=2001, rng='mt19937',
                   nl='faithful', ratio=2, nl_freq=10)
        o.init()
        t0 = time.perf_counter()
        o.run(1, steps)
        cpu.append(n * steps / (time.perf_counter() - t0))
    return dict(workload='one-species LJ, N=1024, rho=0.5, rc=2.5, dt=1e-4, %d steps, FP64'
                % steps, gpu_plain=float(np.median(gpu)),
                cpu_oracle_faithful_1core=float(np.median(cpu)), runs=runs,
                unit='particle-steps/s')


def three_d_leg(n=262144, steps=200, warmup=50, cpu_n=2048, cpu_seconds=8.0, cpu=True):
    """The three-dimensional engine (n_dim = 3, 27-cell stencil): the
    two-species LJ mixture of the ensemble in a periodic cube at density 0.6,
    N = 262,144, graphs of 50 steps, in mixed precision and in FP64, in
    particle-steps/s; beside it the oracle in reference-faithful mode (Verlet
    list, rs = 2 rc, ratio 2, refresh every 10 steps, mt19937 stream, FP64)
    on one core, on the same model at N = cpu_n for about cpu_seconds."""
    import torch
    from pyrdsim_amd.engine import Engine, build_model, make_replica
    s = workloads.lj3(n=n)
    m = s['model']
    out = dict(workload='two-species LJ mixture, 3-D periodic cube, density 0.6, N=%d, '
                        '%d steps after %d' % (n, steps, warmup), unit='particle-steps/s')
    for prec in ('mixed', 'fp64'):
        model, rates = build_model(m['interactions'], m['reactions'], len(m['tau_by_tauD']))
        rep = make_replica(0, n, s['L'], s['periodic'], s['dt'], m['tau_by_tauD'], rates, 2001)
        e = Engine([model], [rep], precision=prec, graph_steps=50, n_dim=3)
        e.set_state(s['x'], s['types'])
        out[prec] = n * steps / time_engine(e, steps, warmup)
        e.close()
    torch.cuda.synchronize()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        from oracle import Oracle
        sc = workloads.lj3(n=cpu_n)
        o = Oracle(sc['x'], sc['types'], sc['L'], sc['periodic'], sc['dt'], m['tau_by_tauD'],
                   m['interactions'], m['reactions'], seed=2001, rng='mt19937', nl='faithful',
                   ratio=2, nl_freq=10)
        o.init()
        done = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < cpu_seconds:
            o.run(done + 1, 10)
            done += 10
        el = time.perf_counter() - t0
        out['cpu_oracle_faithful_1core'] = cpu_n * done / el
        out['cpu_sample'] = 'oracle, reference-faithful FP64, N=%d x %d steps (%.1f s)' % (
            cpu_n, done, el)
    return out


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Streaming copy of nbytes between two device buffers (torch's copy
    kernel), timed with HIP events: (read + write bytes) / time in GB/s, the
    best of `reps`."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device='cuda').fill_(1.0)
    b = torch.empty_like(a)
    best = 0.0
    for _ in range(reps + 1):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        b.copy_(a)
        e1.record()
        e1.synchronize()
        best = max(best, 2 * nbytes / (e0.elapsed_time(e1) * 1e-3) / 1e9)
    del a, b
    torch.cuda.synchronize()
    return best


def save_state(eng, path):
    """Positions, types and step counter of an engine, for a child process."""
    x, t = eng.state()
    np.savez(path, x=x, t=t, step=np.int64(eng.step))


def advance(eng, n):
    """n untimed steps (graph replays where the engine has graphs)."""
    if n > 0:
        eng.run(n)
        eng.sync()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=500)
    ap.add_argument('--warmup', type=int, default=50)
    ap.add_argument('--pre-advance', type=int, default=2000,
                    help='untimed steps from the all-inactive start into the reacting regime '
                         'before any timed window')
    ap.add_argument('--replicas', type=int, default=512)
    ap.add_argument('--plain-subset', type=int, default=32)
    ap.add_argument('--plain-steps', type=int, default=500)
    ap.add_argument('--cpu-seconds', type=float, default=12.0)
    ap.add_argument('--graph-steps', type=int, default=50,
                    help='steps per captured graph (at most --steps, so that the timed '
                         'window replays graphs)')
    ap.add_argument('--tile-particles', type=int, default=0)
    ap.add_argument('--tile-rows', type=int, default=1)
    ap.add_argument('--skin', type=float, default=0.0,
                    help='Verlet skin of the batched engine (0: cell search every step)')
    ap.add_argument('--ic-cache', default=None, help='.npz file of initial states (read or written)')
    ap.add_argument('--state', default=None,
                    help='.npz file (x, t, step) to start the batched run from instead of the '
                         'initial states (the counter passes use it)')
    ap.add_argument('--save-state', default=None,
                    help='.npz file to write the batched state to after the timed window, so '
                         'that a profiled run can start from it with --state')
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-autotune', dest='autotune', action='store_false')
    ap.add_argument('--no-accuracy', action='store_true')
    ap.add_argument('--accuracy-replicas', type=int, default=8)
    ap.add_argument('--accuracy-steps', type=int, default=10000)
    ap.add_argument('--copy-calibration', action='store_true',
                    help='run one elementwise kernel of known traffic (256 MiB read + 256 MiB '
                         'written) first, to scale HBM counters in a profile of this command')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-c1', action='store_true')
    ap.add_argument('--no-3d', dest='three_d', action='store_false')
    ap.add_argument('--no-analysis', action='store_true')
    ap.add_argument('--no-steady-state', action='store_true')
    ap.add_argument('--no-traffic', dest='traffic', action='store_false',
                    help='skip the rocprofv3 counter passes for roofline.traffic')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py runs on one GPU only: this build has no multi-GPU path '
              '(--gpus %d requested)' % args.gpus, file=sys.stderr)
        sys.exit(2)
    graph_steps = max(0, min(args.graph_steps, args.steps))

    t0 = time.perf_counter()
    cache = args.ic_cache
    own_cache = None
    if args.traffic and not cache:
        import tempfile
        fd, own_cache = tempfile.mkstemp(prefix='rds_ic_', suffix='.npz')
        os.close(fd)
        os.unlink(own_cache)
        cache = own_cache
    plan = workloads.ensemble_plan(n_rep=args.replicas)
    states = ensemble.states(plan, min(16, os.cpu_count() or 1), cache)
    print('# built %d replica states in %.1f s' % (len(plan), time.perf_counter() - t0),
          file=sys.stderr, flush=True)

    import torch
    from pyrdsim_amd.engine import Engine
    if args.copy_calibration:
        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device='cuda')
        torch.cuda.synchronize()
        b = a * 2.0          # reads 256 MiB, writes 256 MiB
        torch.cuda.synchronize()
        del a, b
        print('# copy calibration: 268435456 B read + 268435456 B written', file=sys.stderr)
    copy_gbs = copy_bandwidth()
    print('# streaming copy: %.0f GB/s' % copy_gbs, file=sys.stderr, flush=True)
    models, reps = ensemble.engine_inputs(plan)
    n_tot = sum(p['n'] for p in plan)
    x0 = np.concatenate([s[0] for s in states])
    t0s = np.concatenate([s[1] for s in states])

    eng = Engine(models, reps, precision='mixed', graph_steps=graph_steps,
                 tile_particles=args.tile_particles, tile_rows=args.tile_rows, skin=args.skin)
    step0 = 0
    if args.state:
        with np.load(args.state, allow_pickle=False) as z:
            x0, t0s, step0 = z['x'], z['t'], int(z['step'])
    eng.set_state(x0, t0s, step0=step0)
    t = time.perf_counter()
    advance(eng, args.pre_advance)
    print('# pre-advanced %d steps in %.1f s' % (args.pre_advance, time.perf_counter() - t),
          file=sys.stderr, flush=True)
    # the tile shape is tuned on the state that is then timed (the autotuner
    # times clones and leaves the run itself where it is)
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
    steady = None
    if not args.no_steady_state:
        ss_steps = 300
        ss = time_engine(eng, ss_steps, 0)
        steady = dict(steps=ss_steps, ms_per_step=round(1e3 * ss / ss_steps, 5),
                      first_step=int(args.pre_advance + args.warmup + args.steps + 1),
                      graph_steps=graph_steps)
        print('# steady state: %s' % steady, file=sys.stderr, flush=True)

    if args.save_state:
        save_state(eng, args.save_state)
    kt = eng.time_kernels(20)
    eng.sync()
    st2 = eng.sync()
    counts = eng.type_counts()
    state_file = None
    if args.traffic:
        import tempfile
        fd, state_file = tempfile.mkstemp(prefix='rds_state_', suffix='.npz')
        os.close(fd)
        save_state(eng, state_file)
    act = [counts[r, 1] / p['n'] for r, p in enumerate(plan) if p['kind'] == 'wall']
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
    tiling = eng.tiling()
    ana = None
    if not args.no_analysis:
        ana = analysis_leg(eng, plan, reps)
        print('# analysis: %s' % ana, file=sys.stderr, flush=True)
    del eng
    torch.cuda.synchronize()
    traffic = None
    # the counter passes use this run's tile shape
    traffic_tiling = tuple(tiling[:2])
    if args.traffic:
        t = time.perf_counter()
        traffic = measure_traffic(cache, state_file, traffic_tiling[0], traffic_tiling[1],
                                  args.replicas)
        print('# traffic: %s (%.1f s)' % (traffic, time.perf_counter() - t), file=sys.stderr,
              flush=True)
        os.unlink(state_file)
    if own_cache and os.path.exists(own_cache):
        os.unlink(own_cache)
    traffic_bytes = int(traffic['FETCH_SIZE'] + traffic['WRITE_SIZE']) if traffic else None
    roofline = dict(bound='hbm', achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic_bytes,
                    traffic_read=int(traffic['FETCH_SIZE']) if traffic else None,
                    traffic_write=int(traffic['WRITE_SIZE']) if traffic else None,
                    traffic_tiling=list(traffic_tiling) if traffic else None,
                    kernel='k_force (pair forces + Euler-Maruyama + binning)',
                    algorithmic_bytes=int(force_bytes),
                    copy_measured=round(copy_gbs, 1),
                    achieved_force_only_bytes=round(
                        (FORCE_ONLY_BYTES_PER_PARTICLE * n_tot +
                         FORCE_ONLY_BYTES_PER_CELL * n_cells) / (force_ms * 1e-3) / 1e9, 1),
                    frac_of_measured_copy=round(achieved / copy_gbs, 4),
                    avg_ms={k: round(v[0], 5) for k, v in kt.items() if v[1]},
                    at_step=int(args.pre_advance + args.warmup + args.steps +
                                (0 if steady is None else steady['steps'])))

    accuracy = None
    compensated = None
    batched_fp64 = None
    if not args.no_accuracy:
        # throughput of the same batched run with compensated FP32 positions,
        # and in FP64 (separates the gain of batching from that of FP32 pair
        # arithmetic), each pre-advanced like the headline run
        legs = {}
        for prec in ('compensated', 'fp64'):
            e = Engine(models, reps, precision=prec, graph_steps=graph_steps,
                       tile_rows=tiling[0], tile_particles=tiling[1])
            e.set_state(x0, t0s)
            advance(e, args.pre_advance)
            sec = time_engine(e, args.steps, args.warmup)
            legs[prec] = dict(value=n_tot * args.steps / sec, unit='particle-steps/s',
                              ms_per_step=1e3 * sec / args.steps)
            del e
            torch.cuda.synchronize()
        compensated, batched_fp64 = legs['compensated'], legs['fp64']
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
            # tuned once, on the median-size replica of the subset, after the
            # same pre-advance as the timed runs
            r = pick[len(pick) // 2]
            e1 = Engine([models[reps[r].model]], [ensemble.with_model0(reps[r])], precision='fp64')
            e1.set_state(states[r][0], states[r][1])
            advance(e1, args.pre_advance)
            plain_tiling, _ = autotune(e1, steps=32)
            e1.close()
        plain_graph = min(args.graph_steps, args.plain_steps)
        for r in pick:
            e1 = Engine([models[reps[r].model]], [ensemble.with_model0(reps[r])], precision='fp64',
                        graph_steps=plain_graph, tile_rows=plain_tiling[0],
                        tile_particles=plain_tiling[1])
            e1.set_state(states[r][0], states[r][1])
            advance(e1, args.pre_advance)
            dt = time_engine(e1, args.plain_steps, 20)
            work += plan[r]['n'] * args.plain_steps
            tsum += dt
            e1.close()
        plain = dict(value=work / tsum, unit='particle-steps/s', replicas=k,
                     steps=args.plain_steps, precision='fp64', graph_steps=plain_graph,
                     tiling=list(plain_tiling), pre_advance=args.pre_advance)
        print('# plain: %.3e particle-steps/s' % plain['value'], file=sys.stderr, flush=True)

    c1 = None
    if not args.no_c1:
        c1 = c1_leg()
        print('# c1: %s' % c1, file=sys.stderr, flush=True)
    three_d = None
    if args.three_d:
        three_d = three_d_leg(cpu=not args.no_cpu)
        print('# 3-D: %s' % three_d, file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu:
        v, desc = cpu_baseline(plan, states, args.cpu_seconds)
        cpu = dict(value=v, unit='particle-steps/s', cores=1, kind='port',
                   sample='oracle, reference-faithful FP64, mt19937 stream: ' + desc,
                   c1=c1)

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
            'precision': 'mixed', 'graph_steps': graph_steps,
            'pre_advance': args.pre_advance,
            'timed_steps': [int(args.pre_advance + args.warmup + 1),
                            int(args.pre_advance + args.warmup + args.steps)],
            'activated_fraction_wall_mean': round(float(np.mean(act)), 5) if act else None,
            'tiling': list(tiling[:2]), 'lds_cap': tiling[2], 'autotune': tuned,
            'cells': int(n_cells), 'tiles': int(n_tiles), 'lists': lists},
        'roofline': roofline,
        'steady_state': steady,
        'plain': plain,
        'batched_vs_plain': (value / plain['value']) if plain else None,
        'accuracy': accuracy,
        'compensated_positions': compensated,
        'batched_fp64': batched_fp64,
        'analysis': ana,
        'three_d': three_d,
        'cpu_baseline': cpu,
    }
    print(json.dumps(out))


if __name__ == '__main__':
    main()
