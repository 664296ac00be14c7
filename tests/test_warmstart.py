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
import numpy as np
import pytest

from vulcan_amd import warmstart as WS
from vulcan_amd.config import Config
from vulcan_amd.network import synthetic_network
from vulcan_amd.workloads import SOLAR, iso_gradient_profile

gpu = pytest.mark.gpu
NZ = 20


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
