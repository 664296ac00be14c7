"""The learned warm start against the plain method (the equilibrium start).

A small instance of the warm-start study that runs in seconds on the
device: a 30-species network with photolysis (so that chemical
equilibrium is not the steady state), 20-layer columns drawn from a narrow
range of T profiles, Kzz and abundances, 96 training columns integrated to
VULCAN's steady state from equilibrium by the batched device solver, a
two-hidden-layer MLP trained on them with the library's BF16 kernels, and
24 held-out columns run from both starts under the same convergence
criteria (count_min lowered to 20 so that the step count reflects the
approach to steady state rather than the count floor).  The full-size
study (93 species, 150 layers) is vulcan_amd/warmstart_study.py.
"""
import os

import numpy as np
import pytest

from vulcan_amd import warmstart as WS
from vulcan_amd.config import Config
from vulcan_amd.network import synthetic_network
from vulcan_amd.workloads import SOLAR, iso_gradient_profile

gpu = pytest.mark.gpu
NZ = 20
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                     'vulcan_amd', 'weights', 'warmstart_ni93_nz150.npz')


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def narrow_specs(n, seed):
    out = []
    for s in np.random.SeedSequence(seed).spawn(n):
        rng = np.random.default_rng(s)
        T_top, T_bot = rng.uniform(900., 1100.), rng.uniform(1500., 1700.)
        logK = rng.uniform(7.0, 8.0, 3)
        logab = np.array([np.log10(SOLAR[k]) for k in WS.ABUND_KEYS]) + rng.uniform(-0.1, 0.1, 4)
        out.append(WS.ColumnSpec(T=iso_gradient_profile(NZ, T_top, T_bot, WS.P_B, WS.P_T), logK=logK, logab=logab))
    return out


def _cases(net, specs, cfg, P=None):
    out = []
    for i, s in enumerate(specs):
        if P is None:
            out.append(WS.make_case(net, s, cfg=cfg))
        else:
            xp = 10.0 ** P[i].reshape(NZ, net.ni)
            xp = xp / xp.sum(axis=1, keepdims=True)
            out.append(WS.make_case(net, s, cfg=cfg, ymix0=WS.project(net, xp, WS.element_vector(net, s.logab))))
    return out


def test_features_and_targets_shapes():
    specs = narrow_specs(3, 11)
    X = WS.features(specs)
    assert X.shape == (3, 2 * NZ + 7)
    y = np.full((NZ, 5), 0.2)
    assert np.allclose(WS.targets(y[None]), np.log10(0.2))


@gpu
def test_trained_warm_start_reaches_steady_state_sooner():
    _torch()
    from vulcan_amd.warmstart_study import predict, train
    net = synthetic_network(30, 150, 15, 6, seed=1001)
    cfg = Config(count_min=20)
    train_specs = narrow_specs(96, 101)
    res = WS.run_to_steady(net, _cases(net, train_specs, cfg), max_attempts=3000, poll=16, tune=False)
    print('training columns converged: %d of %d' % (res.done.sum(), len(train_specs)))
    assert res.done.sum() >= 32
    X = WS.features(train_specs)[res.done]
    Y = np.stack([WS.targets(ym[None])[0] for ym in res.ymix])[res.done]
    mlp, norm, loss = train([X.shape[1], 256, 256, Y.shape[1]], X, Y, epochs=600, batch=64, lr=2e-3)
    test_specs = narrow_specs(24, 202)
    P = predict(mlp, norm, WS.features(test_specs))
    cold = WS.run_to_steady(net, _cases(net, test_specs, cfg), max_attempts=3000, poll=4, tune=False)
    warm = WS.run_to_steady(net, _cases(net, test_specs, cfg, P), max_attempts=3000, poll=4, tune=False)
    both = cold.done & warm.done
    print('converged cold %d warm %d of %d; median attempts (both converged) cold %.0f warm %.0f; loss %.3e'
          % (cold.done.sum(), warm.done.sum(), len(test_specs), np.median(cold.attempts[both]),
             np.median(warm.attempts[both]), loss))
    assert warm.done.sum() >= cold.done.sum() and both.sum() >= 8
    assert np.median(warm.attempts[both]) < np.median(cold.attempts[both])
    # the two starts reach the same steady state (within the convergence tolerance)
    dev = [np.max(np.abs(np.log10(np.maximum(cold.ymix[c], 1e-30)) - np.log10(np.maximum(warm.ymix[c], 1e-30)))
                  [cold.ymix[c] > 1e-10]) for c in np.nonzero(both)[0]]
    assert np.median(dev) < 0.1


def test_bf16_bits_round_to_nearest_even():
    import torch
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, -0.0], np.float32)
    assert [int(b) for b in WS.bf16_bits(x)] == [0x3F80, 0x3F80, 0x3F82, 0xC020, 0x0000, 0x8000]
    rng = np.random.default_rng(3)
    w = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-30, 30, 20000)).astype(np.float32)
    ref = torch.from_numpy(w).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(WS.bf16_bits(w), ref)
    back = torch.from_numpy(ref.view(np.int16)).view(torch.bfloat16).float().numpy()
    assert np.array_equal(WS.bf16_to_f32(ref), back)


class _Net:
    """What save_model reads of a network: its widths and host arrays."""
    def __init__(self, sizes, seed):
        rng = np.random.default_rng(seed)
        self.sizes = list(sizes)
        self.W = [rng.standard_normal((a, b)).astype(np.float32) for a, b in zip(sizes[:-1], sizes[1:])]
        self.b = [rng.standard_normal(b).astype(np.float32) for b in sizes[1:]]

    def state_dict(self):
        return {'sizes': self.sizes, 'W': self.W, 'b': self.b}


def test_network_file_round_trip(tmp_path):
    net = _Net([7, 16, 16, 5], 4)
    rng = np.random.default_rng(5)
    norm = (rng.standard_normal(7), rng.uniform(1, 2, 7), rng.standard_normal(5), rng.uniform(1, 2, 5))
    path = str(tmp_path / 'net.npz')
    WS.save_model(path, net, norm)
    d = WS.read_model(path)
    assert d['sizes'] == net.sizes
    for i in range(3):
        assert np.array_equal(d['W%d' % i], WS.bf16_bits(net.W[i]))
        assert np.array_equal(d['b%d' % i], net.b[i])
        assert np.max(np.abs(WS.bf16_to_f32(d['W%d' % i]) - net.W[i]) / np.abs(net.W[i])) <= 2.0 ** -8
    for k, v in zip(('xm', 'xs', 'ym', 'ys'), norm):
        assert np.array_equal(d[k], v)
    # a layer whose shape disagrees with the widths is refused
    net.W[1] = net.W[1][:, :8]
    WS.save_model(path, net, norm)
    with pytest.raises(ValueError):
        WS.read_model(path)


def test_committed_network_file():
    if not os.path.exists(MODEL):
        pytest.skip('no committed warm-start network')
    d = WS.read_model(MODEL)
    assert d['sizes'] == [2 * WS.NZ + 7] + [1024] * 4 + [93 * WS.NZ]
    for k in ('xm', 'xs', 'ym', 'ys'):
        assert np.all(np.isfinite(d[k]))
    assert np.all(d['xs'] > 0) and np.all(d['ys'] > 0)
    for i in range(5):
        assert np.all(np.isfinite(WS.bf16_to_f32(d['W%d' % i]))) and np.all(np.isfinite(d['b%d' % i]))


@gpu
def test_committed_network_loads_and_seeds_valid_starts():
    """The device network holds exactly the stored BF16 weights, and its
    prediction for held-out columns, projected, is a composition with each
    column's element inventory."""
    _torch()
    if not os.path.exists(MODEL):
        pytest.skip('no committed warm-start network')
    from vulcan_amd.warmstart_study import predict
    from vulcan_amd.workloads import full_network
    mlp, norm = WS.load_model(MODEL)
    d = WS.read_model(MODEL)
    for i, wb in enumerate(mlp.Wb):
        assert np.array_equal(wb.cpu().numpy().view(np.uint16), d['W%d' % i])
    net = full_network()
    specs = WS.sample_specs(3, 6006, WS.NZ)
    P = predict(mlp, norm, WS.features(specs))
    assert P.shape == (3, WS.NZ * net.ni) and np.all(np.isfinite(P))
    for s, p in zip(specs, P):
        c = WS.warm_case((net, s, p))
        y = c.y0 / c.col.M[:, None]
        assert np.all(y >= 0) and np.allclose(y.sum(axis=1), 1.0, rtol=1e-10)
