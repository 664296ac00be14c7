"""Real VULCAN networks on the device: the 2024 S-N-C-H-O photochemical
network (89 species, 575 pairs) and the C-H-O thermochemical network (41
species, 207 pairs), from their committed device tables
(tests/golden/make_network_tables.py; their forward rates are pinned to the
reference's ReadRate in tests/test_network_text.py).

They exercise the rate forms the synthetic networks lack: Lindemann
three-body rows with a high-pressure limit, k0-only three-body rows and the
special OH + CH3 + M -> CH3OH + M form.  Per network: device rates against
the oracle on a 60-layer column (T 2000 -> 500 K, P 1e3 -> 1e-2 bar,
equilibrium start, photolysis J0 = 1e-5 /s attenuated by exp(-tau)), then
one Ros2 attempt at dt = 1e-6 s against the oracle's attempt (per-step
parity bar of tests/test_device.py).
"""
import copy

import numpy as np
import pytest

from _golden_util import network_fixture
from oracle import ros2_oracle as O
from vulcan_amd import network as nw
from vulcan_amd import network_text as NT
from vulcan_amd import workloads as W
from vulcan_amd.config import Config
from test_device import assert_step_parity

gpu = pytest.mark.gpu

NETWORKS = ['net_SNCHO_photo_network_2024.npz', 'net_CHO_thermo_network.npz']


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def real_case(name, nz=60):
    net = network_fixture(name)
    photo = {int(j): 1e-5 for j, k in zip(net.vulcan_index, net.kind) if k == nw.KIND_PHOTO}
    if photo:
        NT.set_photolysis(net, photo)
    rng = np.random.default_rng(5)
    return W._case(net, nz, 2000., 500., 1e6, 1e10, rng, Config(), tau=bool(photo))


@gpu
@pytest.mark.parametrize('name', NETWORKS)
def test_real_network_rates_attempt_and_steps(name):
    torch = _torch()
    from vulcan_amd.device import ColumnBatch, DevicePlan
    c = real_case(name)
    net = c.net
    b = ColumnBatch(DevicePlan(net), [c.col], [c.y0], c.cfg)
    torch.cuda.synchronize()
    kd = b.k.cpu().numpy()
    ko = O.rates(net, c.col.Tco, c.col.M, c.col.tau)
    nzm = ko != 0
    assert np.array_equal(kd == 0, ko == 0)
    rel = np.max(np.abs(kd[nzm] - ko[nzm]) / np.abs(ko[nzm]))
    per_kind = {}
    for kind in (nw.KIND_ARRH, nw.KIND_LIND, nw.KIND_K0, nw.KIND_CH3OH, nw.KIND_PHOTO):
        cols = np.repeat(net.kind == kind, 2)
        m = nzm & cols[None, :]
        if m.any():
            per_kind[kind] = float(np.max(np.abs(kd[m] - ko[m]) / np.abs(ko[m])))
    print('%s: ni %d, rates max rel %.3e, per kind %s' % (name, net.ni, rel, per_kind))
    assert rel < 1e-12
    for kind in (nw.KIND_LIND, nw.KIND_K0, nw.KIND_CH3OH):
        assert kind in per_kind, kind
    # one attempt from the equilibrium start, on the device's own rates
    dt = 1e-6
    st0 = O.new_state(net, c.y0, kd, c.cfg)
    st0.dt = dt
    stx = copy.deepcopy(st0)
    O.ros2_solver(net, c.col, st0, c.cfg)
    O.ros2_solver(net, c.col, stx, c.cfg, exact=True)
    st = b.host_state()
    st['dt'] = dt
    b.set_host_state(st)
    b.attempt()
    assert_step_parity(b.y.cpu().numpy(), st0.y, stx.y, '%s attempt dt=%g' % (name, dt), stx.exact_berr)
    dl = float(b.host_state()['delta'][0])
    assert abs(dl - st0.delta) <= 1e-6 * st0.delta
    b.close()
