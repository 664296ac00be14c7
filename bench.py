"""Ros2 column-steps/s on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched]

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
  (FACTOR_TRAFFIC below);
* ``cpu_baseline``: the CPU oracle (reference algorithm: dense lhs, band
  repack, two banded LAPACK solves per attempt) timed on the host in a
  single-threaded subprocess on a bounded sample;
* ``batched``: the same network on 256 columns with nz uniform on [80, 200]
  (one captured launch sequence per attempt for all columns);
* ``mixed``: 512 columns over three networks (ni = 40, 93, 150), one batch
  per network on its own stream, block-inversion kernel tuned per ni.
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
# HBM bytes per factorisation of the nz = 150, ni = 93 column, by factor
# variant: FETCH_SIZE + WRITE_SIZE summed over the factorisation kernels of
# one eager attempt (python3 -m vulcan_amd.profile_kernels under
# rocprofv3 --pmc, one counter per run), each counter scaled by its reading
# for a 512 MiB device copy.  Algorithmic: 2 ni^2 x 8 B per layer = 20.8 MB.
FACTOR_TRAFFIC = {4: 343401327}   # block cyclic reduction: 343.4 MB (16.5x; slab products re-read their operands)


def _sync():
    import torch
    torch.cuda.synchronize()


def run_steps(batch, nsteps, col=0):
    """Attempts until column `col` has taken `nsteps` more accepted steps."""
    start = int(batch.host_state()['count'][col])
    attempts = 0
    while int(batch.host_state()['count'][col]) - start < nsteps:
        batch.integrate(1, use_conv=False)
        attempts += 1
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
    return dict(time=dt, steps=steps, attempts=attempts, nz=nz, ni=ni, npairs=c.net.npairs,
                factor_ms=factor_ms, t_model=float(st['t']), dt_model=float(st['dt']), variant=variant,
                tuner_ms=[round(v, 3) for v in b.factor_ms])


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
    return dict(time=dt, col_steps=int(c1 - c0), attempts=attempts, ncols=ncols, sum_nz=int(nzs.sum()),
                ni=cases[0].net.ni, factor_ms=factor_ms)


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
    ap.add_argument('--no-graph', action='store_true', help='eager launches instead of the captured HIP graph')
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
                     'frac': achieved / FP64_PEAK, 'traffic': FACTOR_TRAFFIC.get(s['variant']),
                     'traffic_unit': 'bytes per factorisation (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)',
                     'algorithmic_bytes_per_factorisation': nz * 2 * ni * ni * 8,
                     'flops_per_factorisation': flops_launch, 'avg_factorisation_ms': s['factor_ms']},
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
                          'vs_single': (bt['col_steps'] / bt['time']) / value}
    if not args.no_mixed:
        print('[bench] mixed networks, 512 columns: %d attempts' % args.mixed_attempts, file=sys.stderr, flush=True)
        mx = bench_mixed(args.mixed_attempts, 3)
        out['mixed'] = {'workload': '512 columns over ni=40/93/150 networks (200/600/1100 pairs), nz uniform [100,150]',
                        'value': mx['col_steps'] / mx['time'], 'unit': 'column-steps/s', 'attempts': mx['attempts'],
                        'col_steps': mx['col_steps'], 'time_s': mx['time'], 'groups': mx['groups']}
    if not args.no_cpu:
        print('[bench] cpu baseline (%.0f s budget)' % args.cpu_budget, file=sys.stderr, flush=True)
        out['cpu_baseline'] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out))


if __name__ == '__main__':
    main()
