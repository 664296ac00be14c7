"""Ros2 column-steps/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-mixed] [--no-steady]

A step is one accepted Ros2 step of the column (one iteration of VULCAN's
integration loop, op.py:813-923), rejected attempts included in the time.
The timed workload is the full-size plain method: one column, nz = 150,
ni = 93 (600 reversible pairs, 15 % termolecular with Troe falloff, plus 20
photolysis channels), all FP64, started from dt = 1e-10 s.  Inputs are
resident on the device before timing starts.

The JSON line also carries
* ``roofline`` for the dominant kernel, the block factorisation of the Ros2
  lhs (for one column the tuner picks block cyclic reduction: a sequence of
  launches, timed as one), with the average duration measured by HIP events
  on the launch stream (eager attempts right after the timed region; the
  timed region itself replays the captured per-attempt HIP graph) and the
  HBM traffic per factorisation measured with rocprofv3 byte counters
  (read from the committed counter summary, TRAFFIC_FILE below), and the
  FP64 work that runs on the matrix cores by rocprofv3's MFMA counters
  (``mfma_counted``, from MFMA_FILE);
* ``cpu_baseline``: the CPU oracle (reference algorithm: dense lhs, band
  repack, two banded LAPACK solves per attempt) timed on the host in a
  single-threaded subprocess on a bounded sample, and (``aggregate``) in P
  concurrent single-threaded processes on P columns of the batch;
* ``kernels``: HBM rates of the rate, right-hand-side and lhs-assembly
  kernels (algorithmic bytes over HIP-event time);
* ``batched``: the same network on 256 columns with nz uniform on [80, 200]
  (one captured launch sequence per attempt for all columns);
* ``mixed``: 512 columns over three networks (ni = 40, 93, 150), one batch
  per network on its own stream, block-inversion kernel tuned per ni;
* ``steady_state``: the other half of the metric, time to steady state from
  a warm and from a cold start: held-out seeded columns of the same network
  (nz = 150, random T profiles, Kzz and element abundances) run to VULCAN's
  own ``conv()`` criteria (defaults) in one batch each, once from chemical
  equilibrium and once from the committed warm-start network's prediction
  (WARMSTART_MODEL below, projected on each column's element inventory);
  medians over the columns that converge in both runs within the attempt cap.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # B/s
VARIANTS = ['register Gauss-Jordan chain', 'blocked Gauss-Jordan chain (MFMA Schur updates)',
            'blocked Gauss-Jordan chain, two workgroups per CU', 'two-wave Gauss-Jordan chain',
            'block cyclic reduction (MFMA Schur updates)']
FP64_PEAK = 78.6e12        # FLOP/s (vector = matrix on MI355X)
# HBM bytes per factorisation come from the committed byte-counter summary
# (FETCH_SIZE + WRITE_SIZE of the factorisation kernels of eager attempts,
# python3 -m vulcan_amd.profile_kernels under rocprofv3 --pmc, one counter
# per run, each scaled by its reading for a 512 MiB device copy); the file
# gives the command that made it.
TRAFFIC_FILE = os.path.join('profiles', 'hbm_traffic_ros2_attempt.md')
MFMA_FILE = os.path.join('profiles', 'mfma_counters_factorisation.md')
# the warm-start network (307 -> 4 x 1024 SiLU -> 93 x 150), trained on the
# device on steady states the batched solver generated (python -m
# vulcan_amd.warmstart_study generate / train); plain arrays, BF16 weights
WARMSTART_MODEL = os.path.join('vulcan_amd', 'weights', 'warmstart_ni93_nz150.npz')


def factor_traffic(section):
    """Factorisation bytes per attempt of a workload section ('one column',
    '256 columns') of the counter summary, or None."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    try:
        text = open(path).read()
    except OSError:
        return None
    part = text.split('## ' + section, 1)
    if len(part) < 2:
        return None
    for line in part[1].split('\n## ')[0].splitlines():
        if line.startswith('Factorisation traffic'):
            return float(line.split()[2]) * 1e6
    return None


def factor_mfma(section):
    """Counted FP64 MFMA work of the factorisation in a workload section of
    the MFMA counter summary: {'gflop', 'tflops', 'busy_frac'} per attempt
    (SQ_INSTS_VALU_MFMA_MOPS_F64 x 512 and SQ_VALU_MFMA_BUSY_CYCLES over the
    dispatch time), or None."""
    try:
        text = open(os.path.join(ROOT, MFMA_FILE)).read()
    except OSError:
        return None
    part = text.split('## ' + section, 1)
    if len(part) < 2:
        return None
    for line in part[1].split('\n## ')[0].splitlines():
        if line.startswith('| factorisation |'):
            f = [c.strip() for c in line.strip('|').split('|')]
            return {'gflop': float(f[1]), 'tflops': float(f[5]), 'busy_frac': float(f[6]) / 100,
                    'source': MFMA_FILE}
    return None


def _sync():
    import torch
    torch.cuda.synchronize()


def run_steps(batch, nsteps, col=0):
    """Attempts until column `col` has taken `nsteps` more accepted steps.
    An attempt accepts at most one step, so each call launches as many
    attempts as steps remain (the column never takes more than `nsteps`) and
    the host reads the step count once per call, not after every attempt."""
    start = int(batch.host_state()['count'][col])
    attempts, done = 0, 0
    while done < nsteps:
        n = nsteps - done
        batch.integrate(n, use_conv=False)
        attempts += n
        done = int(batch.host_state()['count'][col]) - start
    return attempts


def factor_timing(b, attempts):
    """Average factor-kernel launch time over `attempts` eager attempts,
    from HIP events recorded on the launch stream."""
    b.set_graph(False)
    b.set_timing(True)
    ms0, n0 = b.factor_stats()
    for _ in range(attempts):
        b.integrate(1, use_conv=False)
    _sync()
    ms1, n1 = b.factor_stats()
    b.set_timing(False)
    return (ms1 - ms0) / max(n1 - n0, 1)


def bench_single(steps, warmup, graph=True):
    import numpy as np
    from vulcan_amd import workloads as W
    from vulcan_amd.device import ColumnBatch, DevicePlan
    c = W.full_case()
    plan = DevicePlan(c.net)
    b = ColumnBatch(plan, [c.col], [c.y0], c.cfg, graph=graph)
    _sync()
    run_steps(b, warmup)
    _sync()
    t0 = time.perf_counter()
    attempts = run_steps(b, steps)
    _sync()
    dt = time.perf_counter() - t0
    st = b.host_state()[0]
    factor_ms = factor_timing(b, 20)
    nz, ni = c.col.nz, c.net.ni
    variant = b._lib.vk_plan_factor_variant(plan.handle)
    kr = kernel_rates(b)
    return dict(time=dt, steps=steps, attempts=attempts, nz=nz, ni=ni, npairs=c.net.npairs,
                factor_ms=factor_ms, t_model=float(st['t']), dt_model=float(st['dt']), variant=variant,
                tuner_ms=[round(v, 3) for v in b.factor_ms], kernels=kr)


def kernel_rates(b, reps=20):
    """Average time and HBM rate of the rate, right-hand-side and lhs-assembly
    entry points on batch b's buffers (HIP events on the launch stream), with
    the algorithmic bytes of each:
      rates: write k (L nr 8) and read T, M, tau;
      rhs (chemdf + diffdf): read k, y, the atmosphere, write the rhs;
      lhs assembly (Jacobian + diffusion, c0 I - J): read k, y, write the
      ni x ni blocks and the two coupling vectors."""
    import torch
    from vulcan_amd import _lib
    from vulcan_amd.device import dptr, stream_handle
    import ctypes as C
    lib = b._lib
    L, ni, nr = b.layout.L, b.ni, b.plan.net.nr
    ncol = b.layout.ncols
    out = torch.empty_like(b.y)
    D = torch.empty((L, ni, ni), dtype=torch.float64, device='cuda')
    Ls, Us = torch.empty_like(b.y), torch.empty_like(b.y)
    dt = torch.full((ncol,), 1e-6, dtype=torch.float64, device='cuda')
    ktmp = torch.empty_like(b.k)
    atm_b = 8 * L * (ni + 8)            # Dzz, Kzz, dzi, Ti, Hpi, M, Tco, ...
    calls = {
        'rates': (lambda: lib.vk_rates(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(ktmp),
                                       stream_handle()), 8 * L * (nr + 3)),
        'rhs (chemdf + diffdf)': (lambda: lib.vk_rhs(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y),
                                                     dptr(b.k), 1, None, None, dptr(out), stream_handle()),
                                  8 * L * (nr + 2 * ni) + atm_b),
        'lhs assembly': (lambda: lib.vk_lhs(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), dptr(b.k), 1,
                                            dptr(dt), dptr(D), dptr(Ls), dptr(Us), stream_handle()),
                         8 * L * (ni * ni + nr + 3 * ni) + 8 * (L - ncol) * 2 * ni + atm_b),
    }
    res = {}
    for name, (fn, nbytes) in calls.items():
        _lib.check(fn(), name)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {'avg_ms': ms, 'algorithmic_bytes': nbytes, 'achieved_GBps': nbytes / (ms * 1e-3) / 1e9,
                     'frac_of_hbm_peak': nbytes / (ms * 1e-3) / HBM_PEAK}
    return res


def bench_batched(attempts, warmup, ncols=256, graph=True):
    import numpy as np
    from vulcan_amd import workloads as W
    from vulcan_amd.device import ColumnBatch, DevicePlan
    cases = W.batched_cases(ncols)
    plan = DevicePlan(cases[0].net)
    b = ColumnBatch(plan, [c.col for c in cases], [c.y0 for c in cases], cases[0].cfg, graph=graph)
    _sync()
    b.integrate(warmup, use_conv=False)
    _sync()
    c0 = b.host_state()['count'].astype(np.int64).sum()
    t0 = time.perf_counter()
    b.integrate(attempts, use_conv=False)
    _sync()
    dt = time.perf_counter() - t0
    c1 = b.host_state()['count'].astype(np.int64).sum()
    factor_ms = factor_timing(b, 3)
    nzs = np.array([c.col.nz for c in cases])
    kr = kernel_rates(b)
    return dict(time=dt, col_steps=int(c1 - c0), attempts=attempts, ncols=ncols, sum_nz=int(nzs.sum()),
                ni=cases[0].net.ni, factor_ms=factor_ms, kernels=kr)


def bench_mixed(attempts, warmup, ncols=512):
    """512 columns over the ni = 40 / 93 / 150 networks, nz uniform on [100, 150]."""
    import numpy as np
    from vulcan_amd import workloads as W
    from vulcan_amd.device import MixedBatch
    cases = W.mixed_cases(ncols)
    mb = MixedBatch(cases, graph=True)
    mb.integrate(warmup, use_conv=False)
    c0 = mb.host_state()['count'].astype(np.int64).sum()
    t0 = time.perf_counter()
    mb.integrate(attempts, use_conv=False)
    mb.synchronize()
    dt = time.perf_counter() - t0
    c1 = mb.host_state()['count'].astype(np.int64).sum()
    groups = [dict(ni=int(b.ni), ncols=int(len(idx)), factor_variant_ms=[round(v, 3) for v in b.factor_ms])
              for idx, b, _ in mb.groups]
    mb.close()
    return dict(time=dt, col_steps=int(c1 - c0), attempts=attempts, ncols=ncols, groups=groups)


def bench_steady(columns, max_attempts, poll=32):
    """Median time to steady state, cold vs warm, on `columns` held-out columns."""
    from vulcan_amd.warmstart_study import evaluate
    path = os.path.join(ROOT, WARMSTART_MODEL)
    if not os.path.exists(path):
        return {'skipped': 'no trained network at %s' % WARMSTART_MODEL}
    t0 = time.perf_counter()
    rep = evaluate(path, columns, 6006, max_attempts, poll)
    rep['leg_seconds'] = time.perf_counter() - t0
    rep['model'] = WARMSTART_MODEL
    return rep


CPU_SNIPPET = r'''
import json, sys, time
sys.path.insert(0, %(root)r)
import numpy as np
from oracle import ros2_oracle as O
from vulcan_amd import workloads as W
c = W.full_case()
k = O.rates(c.net, c.col.Tco, c.col.M, c.col.tau)
st = O.new_state(c.net, c.y0, k, c.cfg)
tb = O.nw.build_tables(c.net)
budget = %(budget)f
t0 = time.perf_counter(); steps = 0; att0 = st.attempts
while time.perf_counter() - t0 < budget:
    st.y_prev = st.y.copy(); st.atom_loss_prev = st.atom_loss.copy()
    O.one_step(c.net, c.col, st, c.cfg, tb, dense_lhs=True)
    st.y = np.vstack(c.col.M) * st.ymix
    st.t += st.dt; st.count += 1
    O.step_size(st, c.cfg)
    steps += 1
dt = time.perf_counter() - t0
print(json.dumps(dict(steps=steps, attempts=st.attempts - att0, time=dt)))
'''


CPU_COLUMN_SNIPPET = r'''
import json, sys, time
sys.path.insert(0, %(root)r)
import numpy as np
from oracle import ros2_oracle as O
from vulcan_amd import workloads as W
c = W.batched_case(%(index)d, %(ncols)d)
k = O.rates(c.net, c.col.Tco, c.col.M, c.col.tau)
st = O.new_state(c.net, c.y0, k, c.cfg)
tb = O.nw.build_tables(c.net)
budget = %(budget)f
print('ready', flush=True)
sys.stdin.readline()                   # every process starts timing together
t0 = time.perf_counter(); steps = 0
while time.perf_counter() - t0 < budget:
    st.y_prev = st.y.copy(); st.atom_loss_prev = st.atom_loss.copy()
    O.one_step(c.net, c.col, st, c.cfg, tb, dense_lhs=True)
    st.y = np.vstack(c.col.M) * st.ymix
    st.t += st.dt; st.count += 1
    O.step_size(st, c.cfg)
    steps += 1
print(json.dumps(dict(steps=steps, time=time.perf_counter() - t0, nz=int(c.col.nz))))
'''


def cpu_model():
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith('Model name:'):
                return line.split(':', 1)[1].strip()
    except Exception:
        pass
    try:
        for line in open('/proc/cpuinfo'):
            if line.startswith('model name'):
                return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def cpu_aggregate(budget=20.0, pmax=16):
    """P single-threaded oracle processes (OMP_NUM_THREADS=1 each, as
    vulcan.py:52 sets), each on its own column of the 256-column batch,
    concurrently for `budget` seconds: aggregate accepted column-steps/s.
    P = the host cores this process may use, capped at `pmax` (the host share
    of one GPU) and by memory (about 2 GB per process for the dense lhs of a
    nz <= 200 column)."""
    avail = len(os.sched_getaffinity(0))
    try:
        mem = os.sysconf('SC_PAGE_SIZE') * os.sysconf('SC_AVPHYS_PAGES')
    except (ValueError, OSError):
        mem = 64 << 30
    P = max(1, min(avail, pmax, int(mem // (3 << 30))))
    env = dict(os.environ, OMP_NUM_THREADS='1', OPENBLAS_NUM_THREADS='1', MKL_NUM_THREADS='1')
    procs = [subprocess.Popen([sys.executable, '-c', CPU_COLUMN_SNIPPET % dict(root=ROOT, ncols=256, index=i,
                                                                                budget=budget)],
                              stdin=subprocess.PIPE, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                              env=env)
             for i in range(P)]
    for pr in procs:                     # set-up done everywhere, then all start
        if pr.stdout.readline().strip() != 'ready':
            sys.stderr.write(pr.stderr.read()[-2000:])
            return None
    for pr in procs:
        pr.stdin.write('go\n')
        pr.stdin.flush()
    steps, nzs, times = 0, [], []
    for pr in procs:
        out, err = pr.communicate(timeout=600)
        if pr.returncode != 0:
            sys.stderr.write(err[-2000:])
            return None
        d = json.loads(out.strip().splitlines()[-1])
        steps += d['steps']
        nzs.append(d['nz'])
        times.append(d['time'])
    wall = max(times)
    return dict(value=steps / wall, unit='column-steps/s', processes=P, cores_available=avail,
                cpu_model=cpu_model(),
                sample='%d concurrent single-threaded processes, one column of the 256-column batch each (nz %d..%d, '
                       'ni=93), started together after their set-up: %d accepted steps in %.1f s'
                       % (P, min(nzs), max(nzs), steps, wall))


def cpu_baseline(budget=20.0):
    env = dict(os.environ, OMP_NUM_THREADS='1', OPENBLAS_NUM_THREADS='1', MKL_NUM_THREADS='1')
    code = CPU_SNIPPET % dict(root=ROOT, budget=budget)
    r = subprocess.run([sys.executable, '-c', code], capture_output=True, text=True, env=env, timeout=600)
    if r.returncode != 0:
        sys.stderr.write(r.stderr[-2000:])
        return None
    d = json.loads(r.stdout.strip().splitlines()[-1])
    return dict(value=d['steps'] / d['time'], unit='column-steps/s', cores=1, kind='port',
                sample='%d accepted steps (%d attempts) of the nz=150, ni=93 column from dt=1e-10 s in %.1f s; '
                       'dense lhs + band repack + 2 scipy solve_banded per attempt, OMP_NUM_THREADS=1'
                       % (d['steps'], d['attempts'], d['time']))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=200)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--batched-attempts', type=int, default=40)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--mixed-attempts', type=int, default=10)
    ap.add_argument('--no-mixed', action='store_true')
    ap.add_argument('--cpu-budget', type=float, default=20.0)
    ap.add_argument('--no-cpu-aggregate', action='store_true')
    ap.add_argument('--no-graph', action='store_true', help='eager launches instead of the captured HIP graph')
    ap.add_argument('--steady-columns', type=int, default=128)
    ap.add_argument('--steady-attempts', type=int, default=4000, help='attempt cap of the steady-state runs')
    ap.add_argument('--no-steady', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: every box in this setup has one MI355X and the columns are '
              'independent (zero per-step communication), so --gpus must be 1', file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print('no HIP device visible', file=sys.stderr)
        sys.exit(1)
    print('[bench] single column nz=150 ni=93: %d warmup + %d timed steps' % (args.warmup, args.steps),
          file=sys.stderr, flush=True)
    s = bench_single(args.steps, args.warmup, graph=not args.no_graph)
    value = s['steps'] / s['time']
    nz, ni = s['nz'], s['ni']
    flops_launch = nz * (8.0 / 3.0) * ni ** 3
    achieved = flops_launch / (s['factor_ms'] * 1e-3)
    out = {
        'metric': 'Ros2 column-steps/s (ni=93,nz=150,FP64)',
        'value': value,
        'unit': 'column-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * s['time'] / s['steps'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': 'plain method, one column: nz=150, ni=93, 600 pairs (90 Troe) + 20 photolysis, '
                               'dt0=1e-10 s', 'attempts': s['attempts'], 'model_time_s': s['t_model'],
                   'launch': 'eager' if args.no_graph else 'hip-graph per attempt'},
        'roofline': {'kernel': 'block factorisation of the Ros2 lhs: ' + VARIANTS[s['variant']]
                               + ' (chosen by the tuner; per-variant ms for factor + 2 solves: %s)' % s['tuner_ms'],
                     'bound': 'mfma',
                     'achieved': achieved / 1e12, 'peak': FP64_PEAK / 1e12, 'unit': 'TFLOP/s',
                     'frac': achieved / FP64_PEAK, 'traffic': factor_traffic('one column'),
                     'traffic_unit': 'bytes per factorisation (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)',
                     'traffic_source': TRAFFIC_FILE,
                     'algorithmic_bytes_per_factorisation': nz * 2 * ni * ni * 8,
                     'flops_per_factorisation': flops_launch, 'avg_factorisation_ms': s['factor_ms'],
                     'mfma_counted': factor_mfma('one column')},
        'kernels': {'workload': 'the same column', 'unit': 'GB/s (algorithmic bytes / HIP-event time)',
                    'hbm_peak_GBps': HBM_PEAK / 1e9, **s['kernels']},
    }
    if not args.no_batched:
        print('[bench] batched 256 columns: %d attempts' % args.batched_attempts, file=sys.stderr, flush=True)
        bt = bench_batched(args.batched_attempts, 5, graph=not args.no_graph)
        fl = bt['sum_nz'] * (8.0 / 3.0) * bt['ni'] ** 3
        out['batched'] = {'workload': '256 columns, nz uniform [80,200], ni=93', 'value': bt['col_steps'] / bt['time'],
                          'unit': 'column-steps/s', 'attempts': bt['attempts'], 'col_steps': bt['col_steps'],
                          'time_s': bt['time'],
                          'factor_avg_launch_ms': bt['factor_ms'],
                          'factor_tflops': fl / (bt['factor_ms'] * 1e-3) / 1e12,
                          'factor_frac': fl / (bt['factor_ms'] * 1e-3) / FP64_PEAK,
                          'factor_traffic': factor_traffic('256 columns'),
                          'factor_mfma_counted': factor_mfma('256 columns'),
                          'vs_single': (bt['col_steps'] / bt['time']) / value,
                          'kernels': bt['kernels']}
    if not args.no_mixed:
        print('[bench] mixed networks, 512 columns: %d attempts' % args.mixed_attempts, file=sys.stderr, flush=True)
        mx = bench_mixed(args.mixed_attempts, 3)
        out['mixed'] = {'workload': '512 columns over ni=40/93/150 networks (200/600/1100 pairs), nz uniform [100,150]',
                        'value': mx['col_steps'] / mx['time'], 'unit': 'column-steps/s', 'attempts': mx['attempts'],
                        'col_steps': mx['col_steps'], 'time_s': mx['time'], 'groups': mx['groups']}
    if not args.no_steady:
        print('[bench] time to steady state, cold vs warm: %d columns, cap %d attempts'
              % (args.steady_columns, args.steady_attempts), file=sys.stderr, flush=True)
        st = bench_steady(args.steady_columns, args.steady_attempts)
        st['workload'] = ('%d held-out columns (seed 6006) of the ni=93 network, nz=150, random T profile, Kzz and '
                          'element abundances; cold = chemical equilibrium, warm = warm-start network prediction '
                          'projected on the element inventory; VULCAN default conv() criteria; one batch per run'
                          % args.steady_columns)
        if st.get('speedup_median_wall'):
            st['metric'] = 'median time-to-steady-state, cold / warm (wall s)'
        out['steady_state'] = st
    if not args.no_cpu:
        print('[bench] cpu baseline (%.0f s budget)' % args.cpu_budget, file=sys.stderr, flush=True)
        out['cpu_baseline'] = cpu_baseline(args.cpu_budget)
        if not args.no_cpu_aggregate:
            print('[bench] cpu aggregate (%.0f s budget)' % args.cpu_budget, file=sys.stderr, flush=True)
            agg = cpu_aggregate(args.cpu_budget)
            if out['cpu_baseline'] is not None and agg is not None:
                out['cpu_baseline']['aggregate'] = agg
    print(json.dumps(out))


if __name__ == '__main__':
    main()
