"""Train the warm-start MLP on solver-generated steady states and compare
time-to-steady-state of warm and cold starts on held-out columns.

    python -m vulcan_amd.warmstart_study generate --first 0 --count 1024 --out gen0.npz
    python -m vulcan_amd.warmstart_study generate --first 1024 --count 2048 --data gen0.npz --out gen1.npz
    python -m vulcan_amd.warmstart_study train-eval --data gen0.npz gen1.npz --test 512 --out report.json
    python -m vulcan_amd.warmstart_study train --data gen0.npz gen1.npz --out model.npz
    python -m vulcan_amd.warmstart_study evaluate --model model.npz --test 512 --out report.json

1. ``generate``: seeded columns (sampler in vulcan_amd/warmstart.py; column
   i of the training stream is the same whatever the chunking) run to
   VULCAN's steady state (``conv()``) in batches on the device; the
   converged ones become training pairs (features -> log10 mixing ratios,
   stored as float16).  The first chunk starts cold (chemical equilibrium);
   with ``--data`` the MLP is first trained on the steady states gathered
   so far and the new columns start from its projected prediction
   (bootstrapped generation: the solver still decides convergence, so every
   training target is a steady state by VULCAN's criteria, and far more of
   them converge within the attempt cap than from equilibrium).
2. ``train-eval``: the MLP (307 -> 4 x 1024 SiLU -> ni*nz) is trained on
   the device with the library's BF16-MFMA kernels and Adam; ``--test``
   held-out columns (a different seed stream) then run twice to steady
   state in one batch each: from the equilibrium start and from the
   projected MLP prediction.  Reported: median wall time to steady state
   per column (completion observed at the host poll), accepted steps and
   attempts, convergence counts and the agreement of the two steady states.
3. ``train`` trains as ``train-eval`` does and saves the network
   (``warmstart.save_model``: BF16 weights, FP32 biases, the feature and
   target normalisation); ``evaluate`` runs ``train-eval``'s comparison
   with a saved network.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import warmstart as WS
from .mlp import MLP
from .workloads import _pool_map, full_network


def log(msg):
    print('[warmstart] ' + msg, file=sys.stderr, flush=True)


def _cases(fn, net, items):
    out = _pool_map(fn, [(net,) + tuple(it) for it in items])
    for c in out:
        c.net = net
    return out


def load_data(files):
    parts = [np.load(f) for f in files]
    X = np.concatenate([p['X'][p['done']] for p in parts]).astype(np.float64)
    Y = np.concatenate([p['Y'][p['done']] for p in parts]).astype(np.float64)
    info = {'generated': int(sum(len(p['done']) for p in parts)),
            'attempts_converged': np.concatenate([p['attempts'][p['done']] for p in parts]),
            'nz': int(parts[0]['nz'])}
    return X, Y, info


def generate(first, count, seed, nz, batch, max_attempts, poll, out, data=None, epochs=200, hidden=1024,
             layers=4, min_train=64, time_budget=1e9):
    net = full_network()
    from .device import DevicePlan
    plan = DevicePlan(net)
    specs = WS.sample_specs(first + count, seed, nz)[first:]
    mlp = norm = None
    if data:
        Xd, Yd, info = load_data(data)
        if Xd.shape[0] >= min_train:
            log('training the warm-start MLP on %d steady states' % Xd.shape[0])
            mlp, norm, _ = train([Xd.shape[1]] + [hidden] * layers + [Yd.shape[1]], Xd, Yd, epochs, 256, 1e-3)
    X = WS.features(specs).astype(np.float32)
    Y = np.zeros((count, nz * net.ni), np.float16)
    rec = {k: np.zeros(count, t) for k, t in (('done', bool), ('count', np.int32), ('attempts', np.int32),
                                             ('wall', np.float32), ('t_model', np.float64), ('warm', bool))}
    t_all = time.perf_counter()
    n_done = 0
    for b0 in range(0, count, batch):
        if time.perf_counter() - t_all > time_budget:
            log('time budget reached after %d columns' % b0)
            break
        chunk = specs[b0:b0 + batch]
        t0 = time.perf_counter()
        if mlp is not None:
            P = predict(mlp, norm, WS.features(chunk))
            cases = _cases(WS.warm_case, net, list(zip(chunk, P)))
        else:
            cases = _cases(WS.cold_case, net, [(s,) for s in chunk])
        log('batch %d-%d: %d %s starts built in %.1f s' % (first + b0, first + b0 + len(chunk), len(chunk),
                                                           'warm' if mlp is not None else 'cold',
                                                           time.perf_counter() - t0))
        res = WS.run_to_steady(net, cases, max_attempts=max_attempts, poll=poll, plan=plan, log=log)
        sl = slice(b0, b0 + len(chunk))
        rec['done'][sl] = res.done
        rec['count'][sl] = res.count
        rec['attempts'][sl] = res.attempts
        rec['wall'][sl] = res.wall
        rec['t_model'][sl] = res.t_model
        rec['warm'][sl] = mlp is not None
        Y[sl] = np.stack([WS.targets(ym[None])[0] for ym in res.ymix]).astype(np.float16)
        n_done = b0 + len(chunk)
        log('batch %d-%d: %d/%d converged, median attempts of converged %s, %.1f s' % (
            first + b0, first + n_done, int(res.done.sum()), len(chunk),
            int(np.median(res.attempts[res.done])) if res.done.any() else None, time.perf_counter() - t0))
    sl = slice(0, n_done)
    Y[~rec['done']] = 0            # only converged columns are training pairs; zero rows compress away
    np.savez_compressed(out, index=np.arange(first, first + n_done, dtype=np.int32), X=X[sl], Y=Y[sl], seed=seed,
                        nz=nz, **{k: v[sl] for k, v in rec.items()})
    log('generated %d columns (%d converged) in %.1f s -> %s' % (n_done, int(rec['done'][sl].sum()),
                                                                time.perf_counter() - t_all, out))


def train(sizes, X, Y, epochs, batch, lr, seed=0):
    dev = torch.device('cuda')
    xm, xs = X.mean(0), X.std(0) + 1e-6
    ym, ys = Y.mean(0), Y.std(0) + 1e-3
    Xn = torch.as_tensor(((X - xm) / xs).astype(np.float32), device=dev)
    Yn = torch.as_tensor(((Y - ym) / ys).astype(np.float32), device=dev)
    mlp = MLP(sizes, seed=seed)
    n = X.shape[0]
    batch = min(batch, n)
    rng = np.random.default_rng(seed)
    steps, total = 0, epochs * max(1, n // batch)
    losses = []
    for ep in range(epochs):
        perm = rng.permutation(n)
        losses = []
        for b0 in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(perm[b0:b0 + batch], device=dev)
            lr_t = lr * 0.5 * (1 + np.cos(np.pi * steps / max(total, 1)))
            loss, _ = mlp.loss_and_backward(Xn[idx].contiguous(), Yn[idx].contiguous())
            mlp.adam_step(lr=lr_t, wd=1e-5)
            losses.append(loss)
            steps += 1
        if ep % max(1, epochs // 10) == 0 or ep == epochs - 1:
            log('epoch %d: loss %.4e' % (ep, float(np.mean(losses))))
    return mlp, (xm, xs, ym, ys), float(np.mean(losses))


def predict(mlp, norm, X):
    xm, xs, ym, ys = norm
    Xn = torch.as_tensor(((X - xm) / xs).astype(np.float32), device='cuda')
    out = []
    for b0 in range(0, X.shape[0], 256):
        out.append(mlp.forward(Xn[b0:b0 + 256].contiguous()).cpu().numpy())
    return np.concatenate(out) * ys + ym


def compare(net, specs, P, max_attempts, poll, plan=None):
    """Cold (equilibrium) and warm (projected prediction P, log10 mixing
    ratios) runs of the same columns to steady state, one batch each."""
    cold = _cases(WS.cold_case, net, [(s,) for s in specs])
    warm = _cases(WS.warm_case, net, [(s, p) for s, p in zip(specs, P)])
    log('cold runs (%d columns)' % len(specs))
    rc = WS.run_to_steady(net, cold, max_attempts=max_attempts, poll=poll, plan=plan, log=log)
    log('warm runs (%d columns)' % len(specs))
    rw = WS.run_to_steady(net, warm, max_attempts=max_attempts, poll=poll, plan=plan, log=log)
    both = rc.done & rw.done
    dev = []
    for c in np.nonzero(both)[0]:
        lc = np.log10(np.maximum(rc.ymix[c], 1e-30))
        lw = np.log10(np.maximum(rw.ymix[c], 1e-30))
        sig = rc.ymix[c] > 1e-12
        dev.append(float(np.max(np.abs(lc - lw)[sig])))
    med = lambda a: float(np.median(a[both])) if both.any() else None
    rep = {
        'test_columns': len(specs), 'cold_converged': int(rc.done.sum()), 'warm_converged': int(rw.done.sum()),
        'median_wall_cold_s': med(rc.wall), 'median_wall_warm_s': med(rw.wall),
        'median_steps_cold': med(rc.count), 'median_steps_warm': med(rw.count),
        'median_attempts_cold': med(rc.attempts), 'median_attempts_warm': med(rw.attempts),
        'batch_wall_cold_s': float(np.nanmax(rc.wall)) if rc.done.any() else None,
        'batch_wall_warm_s': float(np.nanmax(rw.wall)) if rw.done.any() else None,
        'median_max_log10_diff_steady_states': float(np.median(dev)) if dev else None,
    }
    if rep['median_wall_cold_s'] and rep['median_wall_warm_s']:
        rep['speedup_median_wall'] = rep['median_wall_cold_s'] / rep['median_wall_warm_s']
        rep['speedup_median_attempts'] = rep['median_attempts_cold'] / rep['median_attempts_warm']
    return rep


def train_eval(data, test, test_seed, hidden, layers, epochs, mb, lr, max_attempts, poll, out, save_model=None):
    net = full_network()
    X, Y, info = load_data(data)
    nz = info['nz']
    n_gen = info['generated']
    gen_attempts = info['attempts_converged']
    log('training on %d converged columns of %d generated' % (X.shape[0], n_gen))
    sizes = [X.shape[1]] + [hidden] * layers + [Y.shape[1]]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mlp, norm, final_loss = train(sizes, X, Y, epochs, mb, lr)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    if save_model:
        WS.save_model(save_model, mlp, norm)
    specs = WS.sample_specs(test, test_seed, nz)
    Xt = WS.features(specs)
    t2 = time.perf_counter()
    P = predict(mlp, norm, Xt)
    t_pred = time.perf_counter() - t2
    from .device import DevicePlan
    rep = compare(net, specs, P, max_attempts, poll, plan=DevicePlan(net))
    rep.update({'train_columns_generated': n_gen, 'train_columns_converged': int(X.shape[0]),
                'train_median_attempts_cold': float(np.median(gen_attempts)) if len(gen_attempts) else None,
                'mlp_sizes': sizes, 'epochs': epochs, 'final_loss_normalised': final_loss,
                'train_seconds': t_train, 'predict_seconds': t_pred, 'nz': nz, 'ni': net.ni,
                'test_seed': test_seed})
    log(json.dumps(rep))
    with open(out, 'w') as f:
        json.dump(rep, f, indent=1)
    return rep


def train_only(data, hidden, layers, epochs, mb, lr, out):
    """Train on the converged columns of `data` and save the network."""
    X, Y, info = load_data(data)
    log('training on %d converged columns of %d generated' % (X.shape[0], info['generated']))
    sizes = [X.shape[1]] + [hidden] * layers + [Y.shape[1]]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    mlp, norm, final_loss = train(sizes, X, Y, epochs, mb, lr)
    torch.cuda.synchronize()
    WS.save_model(out, mlp, norm)
    log('trained %s in %.1f s, final normalised loss %.4e -> %s' % (sizes, time.perf_counter() - t1, final_loss,
                                                                   out))
    return final_loss


def evaluate(model, test, test_seed, max_attempts, poll, nz=WS.NZ):
    """Cold and warm runs to steady state of `test` held-out columns with a
    saved network (the comparison of train-eval without the training)."""
    net = full_network()
    mlp, norm = WS.load_model(model)
    specs = WS.sample_specs(test, test_seed, nz)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P = predict(mlp, norm, WS.features(specs))
    t_pred = time.perf_counter() - t0
    from .device import DevicePlan
    rep = compare(net, specs, P, max_attempts, poll, plan=DevicePlan(net))
    rep.update({'predict_seconds': t_pred, 'mlp_sizes': list(mlp.sizes), 'nz': nz, 'ni': net.ni,
                'test_seed': test_seed, 'max_attempts': max_attempts})
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest='cmd', required=True)
    g = sub.add_parser('generate')
    g.add_argument('--first', type=int, default=0)
    g.add_argument('--count', type=int, default=4096)
    g.add_argument('--seed', type=int, default=5005)
    g.add_argument('--nz', type=int, default=WS.NZ)
    g.add_argument('--batch', type=int, default=1024)
    g.add_argument('--max-attempts', type=int, default=8000)
    g.add_argument('--poll', type=int, default=16)
    g.add_argument('--out', required=True)
    g.add_argument('--data', nargs='*', default=None, help='earlier outputs: train on them and start warm')
    g.add_argument('--epochs', type=int, default=200)
    g.add_argument('--min-train', type=int, default=64)
    g.add_argument('--time-budget', type=float, default=1e9, help='seconds; no new batch starts after it')
    t = sub.add_parser('train-eval')
    t.add_argument('--data', nargs='+', required=True)
    t.add_argument('--test', type=int, default=512)
    t.add_argument('--test-seed', type=int, default=6006)
    t.add_argument('--hidden', type=int, default=1024)
    t.add_argument('--layers', type=int, default=4)
    t.add_argument('--epochs', type=int, default=300)
    t.add_argument('--mb', type=int, default=256)
    t.add_argument('--lr', type=float, default=1e-3)
    t.add_argument('--max-attempts', type=int, default=8000)
    t.add_argument('--poll', type=int, default=16)
    t.add_argument('--out', required=True)
    t.add_argument('--save-model', default=None)
    o = sub.add_parser('train')
    o.add_argument('--data', nargs='+', required=True)
    o.add_argument('--hidden', type=int, default=1024)
    o.add_argument('--layers', type=int, default=4)
    o.add_argument('--epochs', type=int, default=300)
    o.add_argument('--mb', type=int, default=256)
    o.add_argument('--lr', type=float, default=1e-3)
    o.add_argument('--out', required=True, help='.npz of the trained network')
    e = sub.add_parser('evaluate')
    e.add_argument('--model', required=True)
    e.add_argument('--test', type=int, default=512)
    e.add_argument('--test-seed', type=int, default=6006)
    e.add_argument('--max-attempts', type=int, default=8000)
    e.add_argument('--poll', type=int, default=16)
    e.add_argument('--out', required=True)
    a = ap.parse_args(argv)
    if a.cmd == 'generate':
        generate(a.first, a.count, a.seed, a.nz, a.batch, a.max_attempts, a.poll, a.out, data=a.data,
                 epochs=a.epochs, min_train=a.min_train, time_budget=a.time_budget)
    elif a.cmd == 'train':
        train_only(a.data, a.hidden, a.layers, a.epochs, a.mb, a.lr, a.out)
    elif a.cmd == 'evaluate':
        rep = evaluate(a.model, a.test, a.test_seed, a.max_attempts, a.poll)
        log(json.dumps(rep))
        with open(a.out, 'w') as f:
            json.dump(rep, f, indent=1)
    else:
        train_eval(a.data, a.test, a.test_seed, a.hidden, a.layers, a.epochs, a.mb, a.lr, a.max_attempts, a.poll,
                   a.out, a.save_model)


if __name__ == '__main__':
    main()
