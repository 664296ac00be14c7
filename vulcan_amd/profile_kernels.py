"""Launch the dominant kernels a few times, for counter collection.

    rocprofv3 --pmc FETCH_SIZE -d <dir> -o pmc -- python3 -m vulcan_amd.profile_kernels
    rocprofv3 --pmc WRITE_SIZE -d <dir> -o pmc -- python3 -m vulcan_amd.profile_kernels

Dispatches, in order: a plain device copy of COPY_MIB MiB (torch's copy
kernel: COPY_MIB MiB read and written, the calibration of the byte
counters), then for the one-column workload (nz = 150, ni = 93) the lhs
assembly and `reps` block factorisations with the tuned kernel, then the
same for the 256-column batch.  The factor variant used is printed.
"""
from __future__ import annotations

import argparse
import ctypes as C

import torch

from . import _lib
from . import workloads as W
from .device import ColumnBatch, DevicePlan, dptr, stream_handle

COPY_MIB = 512


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--reps', type=int, default=3)
    ap.add_argument('--no-batched', action='store_true')
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
    for name, cs in cases:
        plan = DevicePlan(cs[0].net)
        b = ColumnBatch(plan, [c.col for c in cs], [c.y0 for c in cs], cs[0].cfg)
        L, ni = b.layout.L, plan.net.ni
        D = torch.empty((L, ni, ni), dtype=torch.float64, device='cuda')
        Di = torch.empty_like(D)
        Ls = torch.empty((L, ni), dtype=torch.float64, device='cuda')
        Us = torch.empty_like(Ls)
        dt = torch.full((len(cs),), 1e-6, dtype=torch.float64, device='cuda')
        s = stream_handle()
        _lib.check(lib.vk_lhs(plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), dptr(b.k), 1, dptr(dt),
                              dptr(D), dptr(Ls), dptr(Us), s), 'vk_lhs')
        for _ in range(a.reps):
            _lib.check(lib.vk_block_factor(plan.handle, b.layout.handle, dptr(D), dptr(Ls), dptr(Us), dptr(Di), None,
                                           s), 'vk_block_factor')
        torch.cuda.synchronize()
        print('%s: L = %d layers, ni = %d, factor variant %d, %d factor launches'
              % (name, L, ni, lib.vk_plan_factor_variant(plan.handle), a.reps), flush=True)
        b.close()


if __name__ == '__main__':
    main()
