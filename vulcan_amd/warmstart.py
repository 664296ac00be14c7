"""Learned warm start for the Ros2 integration.

An MLP maps a column's T profile, log P profile, log Kzz at three anchor
layers and four log element abundances (C/H, O/H, N/H, He/H) to the
steady-state log mixing ratios of every species in every layer.  The
prediction only seeds the integration: it is renormalised per layer and
projected onto the column's element inventory (a Kullback-Leibler
projection, the same element-potential iteration as the equilibrium
solver), and the plain device Ros2 integration then runs until VULCAN's own
``conv()`` criteria hold.  The cold start is the reference's default,
chemical equilibrium (ini_mix = 'EQ').

Columns: nz = 150, P log-spaced 1e3 -> 1e-2 bar, isothermal-plus-gradient T
with top and bottom T uniform on [400, 2500] K, log10 Kzz uniform on
[5, 11] at the bottom, middle and top interfaces (log-linear in between),
element abundances log-uniform within a factor of 10 of vulcan_cfg.py's
O/H, C/H, N/H, He/H, photolysis J0 exp(-tau) with tau log-spaced 1e-4 (top)
-> 1e2 (bottom).
"""
from __future__ import annotations

import dataclasses
import time
from typing import List, Optional, Sequence

import numpy as np

from .atmosphere import build_column
from .config import Config
from .equilibrium import equilibrium
from .network import Network
from .workloads import SOLAR, ColumnCase, iso_gradient_profile

NZ = 150
P_B, P_T = 1e9, 1e4
ABUND_KEYS = ('C', 'O', 'N', 'He')


@dataclasses.dataclass
class ColumnSpec:
    T: np.ndarray          # (nz,)
    logK: np.ndarray       # (3,) log10 Kzz at bottom, middle, top interface
    logab: np.ndarray      # (4,) log10 of C/H, O/H, N/H, He/H


def sample_specs(n: int, seed: int, nz: int = NZ) -> List[ColumnSpec]:
    out = []
    for s in np.random.SeedSequence(seed).spawn(n):
        rng = np.random.default_rng(s)
        T_top, T_bot = rng.uniform(400., 2500., 2)
        logK = rng.uniform(5., 11., 3)
        logab = np.array([np.log10(SOLAR[k]) for k in ABUND_KEYS]) + rng.uniform(-1., 1., 4)
        out.append(ColumnSpec(T=iso_gradient_profile(nz, T_top, T_bot, P_B, P_T), logK=logK, logab=logab))
    return out


def kzz_profile(logK: np.ndarray, nz: int) -> np.ndarray:
    x = np.linspace(0.0, 1.0, nz - 1)
    return 10.0 ** np.interp(x, [0.0, 0.5, 1.0], logK)


def element_vector(net: Network, logab: np.ndarray) -> np.ndarray:
    ab = dict(SOLAR)
    for k, v in zip(ABUND_KEYS, logab):
        ab[k] = 10.0 ** v
    return np.array([ab.get(e, 0.0) for e in net.elements])


def make_case(net: Network, spec: ColumnSpec, cfg: Optional[Config] = None, ymix0=None) -> ColumnCase:
    cfg = cfg or Config()
    nz = len(spec.T)
    if ymix0 is None:
        ymix0 = cold_start(net, spec)
    col = build_column(net, nz, spec.T[0], spec.T[-1], P_b=P_B, P_t=P_T, ymix0=ymix0, cfg=cfg,
                       tau_top=1e-4, tau_bot=1e2, T_profile=spec.T)
    col.Kzz = kzz_profile(spec.logK, nz)
    return ColumnCase(net=net, col=col, y0=col.M[:, None] * ymix0, cfg=cfg)


def cold_case(args) -> ColumnCase:
    """(net, spec) -> the column started from chemical equilibrium."""
    net, spec = args
    return make_case(net, spec)


def warm_case(args) -> ColumnCase:
    """(net, spec, log10 mixing ratios) -> the column started from the
    prediction, renormalised per layer and projected onto the column's
    element inventory."""
    net, spec, logx = args
    xp = 10.0 ** logx.reshape(len(spec.T), net.ni).astype(np.float64)
    xp = xp / xp.sum(axis=1, keepdims=True)
    return make_case(net, spec, ymix0=project(net, xp, element_vector(net, spec.logab)))


def cold_start(net: Network, spec: ColumnSpec) -> np.ndarray:
    pco = np.logspace(np.log10(P_B), np.log10(P_T), len(spec.T))
    x = equilibrium(net, spec.T, pco, element_vector(net, spec.logab))
    x = np.maximum(x, 1e-30)
    return x / x.sum(axis=1, keepdims=True)


def project(net: Network, x_pred: np.ndarray, b: np.ndarray, iters: int = 100) -> np.ndarray:
    """KL projection of per-layer mole fractions onto element ratios b:
    x_i ∝ x_pred_i exp(sum_e a_ie pi_e), normalised, with the element totals
    matching b (same Newton iteration as the equilibrium solver with
    mu_i = -ln x_pred_i)."""
    A = net.compo.astype(np.float64)
    nz = x_pred.shape[0]
    present = np.asarray(b) > 0
    ok = np.all(A[:, ~present] == 0, axis=1)
    As = A[ok][:, present]
    bb = np.broadcast_to(np.asarray(b, dtype=np.float64)[present], (nz, present.sum()))
    bb = bb / bb.sum(axis=1, keepdims=True)
    E = As.shape[1]
    mu0 = -np.log(np.maximum(x_pred[:, ok], 1e-300))
    lnn = -mu0 + np.log(0.1)
    lnN = np.full(nz, np.log(0.1))
    for _ in range(iters):
        n = np.exp(lnn)
        N = np.exp(lnN)
        mu = mu0 + lnn - lnN[:, None]
        G = np.zeros((nz, E + 1, E + 1))
        G[:, :E, :E] = np.einsum('zj,jk,ji->zki', n, As, As)
        G[:, :E, E] = n @ As
        G[:, E, :E] = n @ As
        G[:, E, E] = n.sum(axis=1) - N
        rhs = np.zeros((nz, E + 1))
        rhs[:, :E] = bb - n @ As + np.einsum('zj,jk,zj->zk', n, As, mu)
        rhs[:, E] = N - n.sum(axis=1) + (n * mu).sum(axis=1)
        sol = np.linalg.solve(G, rhs[..., None])[..., 0]
        dlnn = -mu + sol[:, :E] @ As.T + sol[:, E][:, None]
        big = np.maximum(np.abs(dlnn).max(axis=1), np.abs(sol[:, E]) * 5)
        lam = np.minimum(1.0, 2.0 / np.maximum(big, 1e-300))
        lnn = np.maximum(lnn + lam[:, None] * dlnn, -745.0)
        lnN = lnN + lam * sol[:, E]
        if np.max(np.abs(lam[:, None] * dlnn)) < 1e-12:
            break
    x = np.zeros_like(x_pred)
    x[:, ok] = np.exp(lnn - lnN[:, None])
    x = np.maximum(x, 1e-30)
    return x / x.sum(axis=1, keepdims=True)


# --------------------------------------------------------------------------
# features / targets
# --------------------------------------------------------------------------

def features(specs: Sequence[ColumnSpec]) -> np.ndarray:
    """T (nz), log10 P (nz), log10 Kzz at the three anchors, four log10 abundances."""
    nz = len(specs[0].T)
    logP = np.log10(np.logspace(np.log10(P_B), np.log10(P_T), nz))
    return np.stack([np.concatenate([s.T, logP, s.logK, s.logab]) for s in specs]).astype(np.float64)


LOGX_FLOOR = -30.0


def targets(ymix: np.ndarray) -> np.ndarray:
    return np.log10(np.maximum(ymix, 10.0 ** LOGX_FLOOR)).reshape(ymix.shape[0], -1)


# --------------------------------------------------------------------------
# a trained network on disk: plain arrays in an .npz (no pickled objects)
# --------------------------------------------------------------------------

def bf16_bits(w: np.ndarray) -> np.ndarray:
    """float32 -> BF16 bit patterns (uint16), rounding to nearest even as the
    device conversion does (finite inputs)."""
    u = np.ascontiguousarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def bf16_to_f32(bits: np.ndarray) -> np.ndarray:
    return (bits.astype(np.uint32) << 16).view(np.float32)


def save_model(path: str, mlp, norm) -> None:
    """Layer widths, weights as the BF16 values the GEMMs read, FP32 biases
    and the FP64 feature and target normalisation (mean, std)."""
    sd = mlp.state_dict()
    arrs = {'sizes': np.asarray(sd['sizes'], np.int64)}
    for i, (w, b) in enumerate(zip(sd['W'], sd['b'])):
        arrs['W%d' % i] = bf16_bits(w)
        arrs['b%d' % i] = np.asarray(b, np.float32)
    for k, v in zip(('xm', 'xs', 'ym', 'ys'), norm):
        arrs[k] = np.asarray(v, np.float64)
    np.savez_compressed(path, **arrs)


def read_model(path: str) -> dict:
    """The arrays of a saved network, checked against its layer widths."""
    with np.load(path, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    sizes = [int(s) for s in d['sizes']]
    for i, (fi, fo) in enumerate(zip(sizes[:-1], sizes[1:])):
        if d['W%d' % i].shape != (fi, fo) or d['b%d' % i].shape != (fo,):
            raise ValueError('%s: layer %d does not have the shape its widths give' % (path, i))
    if d['xm'].shape != (sizes[0],) or d['ym'].shape != (sizes[-1],):
        raise ValueError('%s: normalisation does not match the layer widths' % path)
    d['sizes'] = sizes
    return d


def load_model(path: str):
    """(MLP on the device, normalisation) of a saved network."""
    from .mlp import MLP
    d = read_model(path)
    n = len(d['sizes']) - 1
    mlp = MLP(d['sizes'])
    mlp.load_state_dict({'W': [bf16_to_f32(d['W%d' % i]) for i in range(n)],
                         'b': [d['b%d' % i] for i in range(n)]})
    return mlp, (d['xm'], d['xs'], d['ym'], d['ys'])


# --------------------------------------------------------------------------
# batched runs to steady state with per-column completion times
# --------------------------------------------------------------------------

@dataclasses.dataclass
class RunResult:
    done: np.ndarray        # end_case == 1 (converged)
    end_case: np.ndarray
    count: np.ndarray       # accepted steps
    attempts: np.ndarray
    t_model: np.ndarray
    wall: np.ndarray        # seconds from start until the column was seen done
    ymix: List[np.ndarray]


def run_to_steady(net: Network, cases: Sequence[ColumnCase], max_attempts: int = 20000, poll: int = 16,
                  plan=None, log=None, tune: bool = True) -> RunResult:
    """All columns in one batch; the device loop stops a column at conv()
    (or count_max / runtime); the host polls every `poll` attempts."""
    import torch
    from .device import ColumnBatch, DevicePlan
    plan = plan or DevicePlan(net)
    cfg = cases[0].cfg
    b = ColumnBatch(plan, [c.col for c in cases], [c.y0 for c in cases], cfg, history=cfg.conv_step + 1, tune=tune)
    nc = len(cases)
    wall = np.full(nc, np.nan)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    attempts = 0
    while attempts < max_attempts:
        b.integrate(poll)
        attempts += poll
        st = b.host_state()
        now = time.perf_counter() - t0
        newly = (st['done'] == 1) & np.isnan(wall)
        wall[newly] = now
        if log and attempts % 512 < poll:
            log('  attempts %d: %d/%d done, %.1f s' % (attempts, int(st['done'].sum()), nc, now))
        if st['done'].all():
            break
    st = b.host_state()
    ym = [b.column('ymix', c) for c in range(nc)]
    res = RunResult(done=st['end_case'] == 1, end_case=st['end_case'].copy(), count=st['count'].copy(),
                    attempts=st['attempts'].copy(), t_model=st['t'].copy(), wall=wall, ymix=ym)
    b.close()
    return res
