"""The learned warm start against the plain method (the equilibrium start).

A small instance of the warm-start study that runs in seconds: the
30-species network, 20-layer columns, 96 training columns integrated to
VULCAN's steady state from equilibrium by the batched device solver, a
two-hidden-layer MLP trained on them with the library's BF16 kernels, and
16 held-out columns run from both starts under the same convergence
criteria (count_min lowered to 20 so that the step count reflects the
approach to steady state rather than the count floor).  The full-size
study (93 species, 150 layers, 4096 training columns) is
vulcan_amd/warmstart_study.py.
"""
import numpy as np
import pytest

from vulcan_amd import warmstart as WS
from vulcan_amd.config import Config
from vulcan_amd.workloads import small_network

gpu = pytest.mark.gpu
NZ = 20


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


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
    specs = WS.sample_specs(3, 11, NZ)
    X = WS.features(specs)
    assert X.shape == (3, 2 * NZ + 7)
    y = np.full((NZ, 5), 0.2)
    assert np.allclose(WS.targets(y[None]), np.log10(0.2))


@gpu
def test_trained_warm_start_reaches_steady_state_sooner():
    _torch()
    from vulcan_amd.warmstart_study import predict, train
    net = small_network()
    cfg = Config(count_min=20)
    train_specs = WS.sample_specs(96, 101, NZ)
    res = WS.run_to_steady(net, _cases(net, train_specs, cfg), max_attempts=4000, poll=16)
    assert res.done.sum() >= 64
    X = WS.features(train_specs)[res.done]
    Y = np.stack([WS.targets(ym[None])[0] for ym in res.ymix])[res.done]
    mlp, norm, loss = train([X.shape[1], 256, 256, Y.shape[1]], X, Y, epochs=400, batch=64, lr=2e-3)
    test_specs = WS.sample_specs(16, 202, NZ)
    P = predict(mlp, norm, WS.features(test_specs))
    cold = WS.run_to_steady(net, _cases(net, test_specs, cfg), max_attempts=4000, poll=4)
    warm = WS.run_to_steady(net, _cases(net, test_specs, cfg, P), max_attempts=4000, poll=4)
    both = cold.done & warm.done
    print('converged cold %d warm %d; median attempts cold %.0f warm %.0f; training loss %.3e'
          % (cold.done.sum(), warm.done.sum(), np.median(cold.attempts[both]), np.median(warm.attempts[both]), loss))
    assert warm.done.sum() >= cold.done.sum() and both.sum() >= 12
    assert np.median(warm.attempts[both]) < np.median(cold.attempts[both])
    # the two starts reach the same steady state (within the convergence tolerance)
    dev = [np.max(np.abs(np.log10(np.maximum(cold.ymix[c], 1e-30)) - np.log10(np.maximum(warm.ymix[c], 1e-30)))
                  [cold.ymix[c] > 1e-10]) for c in np.nonzero(both)[0]]
    assert np.median(dev) < 0.1
