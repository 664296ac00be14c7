"""Run a few eager Ros2 attempts of the bench workloads, for counter collection.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o pmc -- python3 -m vulcan_amd.profile_kernels
    rocprofv3 --pmc WRITE_SIZE -d <dir> -o pmc -- python3 -m vulcan_amd.profile_kernels

Dispatches, in order: a plain device copy of COPY_MIB MiB (COPY_MIB MiB read
and written: the calibration of the byte counters), then for the one-column
workload (nz = 150, ni = 93) and for the 256-column batch (nz uniform on
[80, 200]) `attempts` eager attempts each (every kernel of an attempt is a
separate dispatch; no graph replay), with the factor variants bench.py's
tuner picks for them (set with --variants).
"""
from __future__ import annotations

import argparse

import torch

from . import _lib
from . import workloads as W
from .device import ColumnBatch, DevicePlan

COPY_MIB = 512


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--attempts', type=int, default=3)
    ap.add_argument('--no-batched', action='store_true')
    ap.add_argument('--variants', type=int, nargs=2, default=[4, 2], metavar=('SINGLE', 'BATCHED'),
                    help='factor variants of the two workloads (the ones the tuner picks in bench.py: cyclic '
                         'reduction for one column, the two-workgroups-per-CU chain for 256); counter collection '
                         'distorts the tuner\'s timings, so the variant is set rather than tuned here')
    a = ap.parse_args(argv)
    lib = _lib.load()
    src = torch.ones(COPY_MIB * 2 ** 20 // 8, dtype=torch.float64, device='cuda')
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print('copy: %d MiB read, %d MiB written' % (COPY_MIB, COPY_MIB), flush=True)
    del src, dst
    cases = [('single', [W.full_case()])]
    if not a.no_batched:
        cases.append(('batched', W.batched_cases(256)))
    for (name, cs), variant in zip(cases, a.variants):
        plan = DevicePlan(cs[0].net)
        b = ColumnBatch(plan, [c.col for c in cs], [c.y0 for c in cs], cs[0].cfg, tune=False)
        _lib.check(lib.vk_plan_set_factor_variant(plan.handle, variant), 'vk_plan_set_factor_variant')
        b.set_graph(False)
        torch.cuda.synchronize()
        print('%s: L = %d layers, ni = %d, factor variant %d; %d eager attempts follow'
              % (name, b.layout.L, plan.net.ni, lib.vk_plan_factor_variant(plan.handle), a.attempts), flush=True)
        b.integrate(a.attempts, use_conv=False)
        torch.cuda.synchronize()
        b.close()


if __name__ == '__main__':
    main()
