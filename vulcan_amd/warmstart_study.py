"""Train the warm-start MLP on solver-generated steady states and compare
time-to-steady-state of warm and cold starts on held-out columns.

    python -m vulcan_amd.warmstart_study --train 4096 --test 512 --out ws.json

1. ``--train`` seeded columns (sampler in vulcan_amd/warmstart.py) are run
   from their equilibrium cold start to VULCAN's steady state (``conv()``)
   by the batched device integrator; converged columns become training
   pairs (features -> log10 mixing ratios).
2. The MLP (307 -> 4 x 1024 SiLU -> ni*nz) is trained on the device with
   the library's BF16-MFMA kernels and Adam.
3. ``--test`` held-out columns run twice to steady state: from the
   equilibrium start and from the projected MLP prediction.  Reported:
   median wall time to steady state per column (batched run, completion
   observed at the host poll), accepted steps and attempts, and how many
   columns met the convergence criteria.
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np
import torch

from . import warmstart as WS
from .config import Config
from .mlp import MLP
from .workloads import full_network


def log(msg):
    print('[warmstart] ' + msg, flush=True)


def generate(net, specs, batch, max_attempts, poll, plan):
    feats, targs, res_all = [], [], []
    for b0 in range(0, len(specs), batch):
        chunk = specs[b0:b0 + batch]
        cases = [WS.make_case(net, s) for s in chunk]
        t0 = time.perf_counter()
        res = WS.run_to_steady(net, cases, max_attempts=max_attempts, poll=poll, plan=plan, log=log)
        log('batch %d-%d: %d/%d converged, median steps %d, %.1f s' % (
            b0, b0 + len(chunk), int(res.done.sum()), len(chunk), int(np.median(res.count)),
            time.perf_counter() - t0))
        X = WS.features(chunk)
        Y = np.stack([WS.targets(ym[None])[0] for ym in res.ymix])
        feats.append(X[res.done])
        targs.append(Y[res.done])
        res_all.append(res)
    return np.concatenate(feats), np.concatenate(targs), res_all


def train(net_sizes, X, Y, epochs, batch, lr, seed=0):
    dev = torch.device('cuda')
    xm, xs = X.mean(0), X.std(0) + 1e-6
    ym, ys = Y.mean(0), Y.std(0) + 1e-3
    Xn = torch.as_tensor(((X - xm) / xs).astype(np.float32), device=dev)
    Yn = torch.as_tensor(((Y - ym) / ys).astype(np.float32), device=dev)
    mlp = MLP(net_sizes, seed=seed)
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    steps = 0
    total = epochs * max(1, n // batch)
    for ep in range(epochs):
        perm = rng.permutation(n)
        losses = []
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(perm[b0:b0 + batch], device=dev)
            frac = steps / max(total, 1)
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * frac))
            loss, _ = mlp.loss_and_backward(Xn[idx].contiguous(), Yn[idx].contiguous())
            mlp.adam_step(lr=lr_t, wd=1e-5)
            losses.append(loss)
            steps += 1
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            log('epoch %d: loss %.4e' % (ep, float(np.mean(losses))))
    return mlp, (xm, xs, ym, ys)


def predict(mlp, norm, X):
    xm, xs, ym, ys = norm
    Xn = torch.as_tensor(((X - xm) / xs).astype(np.float32), device='cuda')
    out = []
    for b0 in range(0, X.shape[0], 256):
        out.append(mlp.forward(Xn[b0:b0 + 256].contiguous()).cpu().numpy())
    return np.concatenate(out) * ys + ym


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--train', type=int, default=4096)
    ap.add_argument('--test', type=int, default=512)
    ap.add_argument('--batch', type=int, default=512)
    ap.add_argument('--max-attempts', type=int, default=12000)
    ap.add_argument('--poll', type=int, default=16)
    ap.add_argument('--epochs', type=int, default=300)
    ap.add_argument('--mb', type=int, default=256)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--seed', type=int, default=5005)
    ap.add_argument('--out', default='warmstart.json')
    ap.add_argument('--save-model', default=None)
    a = ap.parse_args(argv)
    torch.cuda.synchronize()
    net = full_network()
    from .device import DevicePlan
    plan = DevicePlan(net)
    log('generating %d training steady states' % a.train)
    t0 = time.perf_counter()
    specs = WS.sample_specs(a.train, a.seed)
    X, Y, _ = generate(net, specs, a.batch, a.max_attempts, a.poll, plan)
    t_gen = time.perf_counter() - t0
    log('training on %d converged columns' % X.shape[0])
    sizes = [X.shape[1], 1024, 1024, 1024, 1024, Y.shape[1]]
    t1 = time.perf_counter()
    mlp, norm = train(sizes, X, Y, a.epochs, a.mb, a.lr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    if a.save_model:
        np.savez_compressed(a.save_model, **{'W%d' % i: w for i, w in enumerate(mlp.state_dict()['W'])},
                            **{'b%d' % i: b for i, b in enumerate(mlp.state_dict()['b'])},
                            xm=norm[0], xs=norm[1], ym=norm[2], ys=norm[3], sizes=np.array(sizes))
    test = WS.sample_specs(a.test, a.seed + 1)
    Xt = WS.features(test)
    t2 = time.perf_counter()
    P = predict(mlp, norm, Xt)
    t_pred = time.perf_counter() - t2
    cold = [WS.make_case(net, s) for s in test]
    warm = []
    for s, p in zip(test, P):
        xp = 10.0 ** p.reshape(WS.NZ, net.ni)
        xp = xp / xp.sum(axis=1, keepdims=True)
        xw = WS.project(net, xp, WS.element_vector(net, s.logab))
        warm.append(WS.make_case(net, s, ymix0=xw))
    log('held-out cold runs')
    rc = WS.run_to_steady(net, cold, max_attempts=a.max_attempts, poll=a.poll, plan=plan, log=log)
    log('held-out warm runs')
    rw = WS.run_to_steady(net, warm, max_attempts=a.max_attempts, poll=a.poll, plan=plan, log=log)
    both = rc.done & rw.done
    # agreement of the two steady states (same criteria, same element inventory)
    dev = []
    for c in np.nonzero(both)[0]:
        lc = np.log10(np.maximum(rc.ymix[c], 1e-30))
        lw = np.log10(np.maximum(rw.ymix[c], 1e-30))
        sig = rc.ymix[c] > 1e-12
        dev.append(float(np.max(np.abs(lc - lw)[sig])))
    rep = {
        'train_columns': a.train, 'train_converged': int(X.shape[0]), 'gen_seconds': t_gen,
        'train_seconds': t_train, 'predict_seconds_512': t_pred, 'epochs': a.epochs,
        'test_columns': a.test, 'cold_converged': int(rc.done.sum()), 'warm_converged': int(rw.done.sum()),
        'median_wall_cold_s': float(np.nanmedian(rc.wall[both])) if both.any() else None,
        'median_wall_warm_s': float(np.nanmedian(rw.wall[both])) if both.any() else None,
        'median_steps_cold': float(np.median(rc.count[both])) if both.any() else None,
        'median_steps_warm': float(np.median(rw.count[both])) if both.any() else None,
        'median_attempts_cold': float(np.median(rc.attempts[both])) if both.any() else None,
        'median_attempts_warm': float(np.median(rw.attempts[both])) if both.any() else None,
        'median_max_log10_diff_steady_states': float(np.median(dev)) if dev else None,
    }
    if rep['median_wall_cold_s'] and rep['median_wall_warm_s']:
        rep['speedup_median_wall'] = rep['median_wall_cold_s'] / rep['median_wall_warm_s']
        rep['speedup_median_attempts'] = rep['median_attempts_cold'] / rep['median_attempts_warm']
    log(json.dumps(rep))
    with open(a.out, 'w') as f:
        json.dump(rep, f, indent=1)
    return rep


if __name__ == '__main__':
    main()
