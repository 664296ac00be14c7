"""The bounds-checked build of the chemistry kernels (libvulcan_amd_checked.so).

In that build every table index the rate, right-hand-side and assembly
kernels gather through is checked against the length of its array (an
index out of range is recorded by site and replaced by 0, so it is never
used).  The GPU test runs the same integration steps through the normal and
the checked library: the full-size synthetic column, one column of each
network of the mixed batch (ni = 40, 93, 150) and the two real VULCAN
networks (their Lindemann, k0-only and CH3OH rate rows); the states and
rates must agree bit for bit and no index may be out of range.  Each
library is loaded in a child process of its own.
"""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from vulcan_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
gpu = pytest.mark.gpu


def test_checked_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.CHECKED_LIB_PATH):
        pytest.fail('%s is missing: build both libraries with __graft_entry__.build()' % _lib.CHECKED_LIB_PATH)
    hdr = open(os.path.join(ROOT, 'include', 'vulcan_amd.h')).read()
    declared = set(re.findall(r'^int\s+(vk_\w+)\s*\(', hdr, flags=re.M))
    lib = C.CDLL(_lib.CHECKED_LIB_PATH)      # loading needs no device
    for name in declared:
        assert hasattr(lib, name), name


RUN = r'''
import json, os, sys
sys.path.insert(0, %(root)r)
sys.path.insert(0, os.path.join(%(root)r, 'tests'))
import ctypes as C
import numpy as np
import torch
from vulcan_amd import _lib
_lib.LIB_PATH = sys.argv[1]
from vulcan_amd import workloads as W
from vulcan_amd.device import ColumnBatch, DevicePlan
from test_real_networks import real_case, NETWORKS
cases = [('full', W.full_case())] + [('mixed%%d' %% i, c) for i, c in enumerate(W.mixed_cases(3))]
cases += [(n, real_case(n)) for n in NETWORKS]
lib = _lib.load()
site = C.c_int(-1)
rc0 = lib.vk_debug_checks(C.byref(site), 1)
out = {}
for name, c in cases:
    b = ColumnBatch(DevicePlan(c.net), [c.col], [c.y0], c.cfg, tune=False)
    b.integrate(4, use_conv=False)
    torch.cuda.synchronize()
    out[name + '_y'] = b.y.cpu().numpy()
    out[name + '_k'] = b.k.cpu().numpy()
    b.close()
site = C.c_int(-1)
rc = lib.vk_debug_checks(C.byref(site), 0)
np.savez(sys.argv[2], **out)
print(json.dumps({'rc_reset': rc0, 'rc': rc, 'site': site.value, 'cases': [n for n, _ in cases]}))
'''


@gpu
@pytest.mark.timeout(360)
def test_checked_build_agrees_bitwise_and_finds_no_index_out_of_range(tmp_path):
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    if not os.path.exists(_lib.CHECKED_LIB_PATH):
        pytest.fail('%s is missing' % _lib.CHECKED_LIB_PATH)
    res = {}
    for tag, path in (('normal', _lib.LIB_PATH), ('checked', _lib.CHECKED_LIB_PATH)):
        out = str(tmp_path / (tag + '.npz'))
        print('[checked build] %s library: %s' % (tag, path), flush=True)
        p = subprocess.run([sys.executable, '-c', RUN % {'root': ROOT}, path, out], capture_output=True, text=True,
                           timeout=150)
        assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
        info = json.loads(p.stdout.strip().splitlines()[-1])
        res[tag] = (info, np.load(out))
    info_n, arr_n = res['normal']
    info_c, arr_c = res['checked']
    print(info_n, info_c)
    assert info_n['rc'] == -1004                   # VK_ENOTCHECKED: the normal build has no checks
    assert info_c['rc_reset'] == 0 and info_c['rc'] == 0
    assert info_c['site'] == 0, 'table index out of range at site %d' % info_c['site']
    assert sorted(arr_n.files) == sorted(arr_c.files)
    for k in arr_n.files:
        assert np.array_equal(arr_n[k], arr_c[k]), k
