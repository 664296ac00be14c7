"""Seeded synthetic workloads.

* ``small``    one column, nz=60, ni=30, 150 reversible bimolecular pairs,
               T linear 2000 K (bottom) -> 500 K (top), P 1e3 -> 1e-2 bar,
               Kzz = 1e7 cm2/s, no photolysis, dt0 = 1e-10 s.
* ``full``     one column, nz=150, ni=93, 600 pairs (15 % termolecular with
               Troe falloff) plus 20 photolysis channels with
               J = J0 exp(-tau), tau log-spaced 1e-4 (top) -> 1e2 (bottom);
               Kzz log-linear 1e6 (bottom) -> 1e10 (top) cm2/s.
* ``batched``  256 columns sharing the ``full`` network; nz uniform on
               [80, 200]; isothermal-plus-gradient T with top and bottom T
               uniform on [400, 2500] K; constant Kzz log-uniform on
               [1e5, 1e11] cm2/s.
* ``mixed``    512 columns over three networks (ni, pairs) = (40, 200),
               (93, 600), (150, 1100; three extra elements); nz uniform on
               [100, 150].

``small`` starts from mixing ratios log-uniform on [1e-12, 1e-1], normalised
to 1.  The others start, as VULCAN does by default (ini_mix = 'EQ'), from
chemical equilibrium at each layer's T and P (vulcan_amd/equilibrium.py) for
the element abundances of vulcan_cfg.py:30-34 (O/H, C/H, N/H, He/H; extra
elements at the S/H value).  ``y = n_0 * ymix``.
"""
from __future__ import annotations

import dataclasses
import os
from typing import List, Optional

import numpy as np

from .atmosphere import Column, build_column, initial_mixing
from .config import Config
from .equilibrium import equilibrium
from .network import Network, synthetic_network

# vulcan_cfg.py:30-34 (relative to H)
SOLAR = {'H': 1.0, 'O': 5.37E-4, 'C': 2.95E-4, 'N': 7.08E-5, 'He': 0.0838, 'S': 1.41E-5,
         'X': 1.41E-5, 'Y': 1.41E-5, 'Z': 1.41E-5}


def element_vector(net: Network, abund=None) -> np.ndarray:
    """Element abundances relative to H; zero for elements no species of the
    network carries (their conservation row would be empty)."""
    abund = dict(SOLAR, **(abund or {}))
    present = np.asarray(net.compo).any(axis=0)
    return np.array([abund.get(e, 0.0) if present[q] else 0.0 for q, e in enumerate(net.elements)])

SEED_SMALL, SEED_FULL, SEED_BATCH, SEED_MIXED = 1001, 1002, 1003, 1004


@dataclasses.dataclass
class ColumnCase:
    net: Network
    col: Column
    y0: np.ndarray
    cfg: Config


def small_network() -> Network:
    return synthetic_network(30, 150, 0, 0, seed=SEED_SMALL)


def full_network() -> Network:
    return synthetic_network(93, 510, 90, 20, seed=SEED_FULL)


def mixed_networks() -> List[Network]:
    return [synthetic_network(40, 200, 0, 0, seed=SEED_MIXED),
            full_network(),
            synthetic_network(150, 935, 165, 0, seed=SEED_MIXED + 2, extra_elements=3)]


def _case(net, nz, T_bot, T_top, Kzz_bot, Kzz_top, rng, cfg, tau=False, T_profile=None, start='eq',
          abund=None):
    if start == 'random':
        ymix0 = initial_mixing(net.ni, nz, rng)
    else:
        Tco = np.array(T_profile, dtype=np.float64) if T_profile is not None else np.linspace(T_bot, T_top, nz)
        pco = np.logspace(np.log10(1e9), np.log10(1e4), nz)
        ymix0 = equilibrium(net, Tco, pco, element_vector(net, abund))
        ymix0 = np.maximum(ymix0, 1e-30)
        ymix0 = ymix0 / ymix0.sum(axis=1, keepdims=True)
    col = build_column(net, nz, T_bot, T_top, P_b=1e9, P_t=1e4, Kzz_bot=Kzz_bot,
                       Kzz_top=Kzz_top, ymix0=ymix0, cfg=cfg,
                       tau_top=1e-4 if tau else None, tau_bot=1e2 if tau else None,
                       T_profile=T_profile)
    y0 = col.M[:, None] * ymix0
    return ColumnCase(net=net, col=col, y0=y0, cfg=cfg)


def small_case(nz: int = 60, use_moldiff: bool = True, net: Optional[Network] = None) -> ColumnCase:
    cfg = Config(use_moldiff=use_moldiff)
    rng = np.random.default_rng(np.random.SeedSequence(SEED_SMALL).spawn(1)[0])
    return _case(net or small_network(), nz, 2000., 500., 1e7, None, rng, cfg, start='random')


def full_case(nz: int = 150, net: Optional[Network] = None) -> ColumnCase:
    cfg = Config()
    rng = np.random.default_rng(np.random.SeedSequence(SEED_FULL).spawn(1)[0])
    return _case(net or full_network(), nz, 2000., 500., 1e6, 1e10, rng, cfg, tau=True)


def iso_gradient_profile(nz: int, T_top: float, T_bot: float, P_b=1e9, P_t=1e4, P_iso=1e5):
    """T isothermal at T_top above P_iso, log-linear in P down to T_bot at P_b."""
    pco = np.logspace(np.log10(P_b), np.log10(P_t), nz)
    f = np.clip((np.log10(pco) - np.log10(P_iso)) / (np.log10(P_b) - np.log10(P_iso)), 0.0, 1.0)
    return T_top + (T_bot - T_top) * f


def _pool_map(fn, items, workers=None):
    """Map over case descriptions, in worker processes (spawned: no state is
    inherited from the caller, which may hold a GPU context) when there are
    many; results come back in order.  `fn` must live in an importable
    module (the workers do not import the caller's main script)."""
    workers = workers or min(16, os.cpu_count() or 1, len(items))
    if workers <= 1 or len(items) < 16:
        return [fn(it) for it in items]
    import multiprocessing as mp
    import sys
    # a spawned worker re-imports the caller's main module unless it has no
    # name or path; hide them so that a script which has set up the GPU is
    # not run again in every worker (the workers only need this module)
    main = sys.modules['__main__']
    saved = {k: main.__dict__[k] for k in ('__spec__', '__file__') if k in main.__dict__}
    main.__dict__.pop('__file__', None)
    main.__spec__ = None
    from concurrent.futures import ProcessPoolExecutor
    try:
        # a worker failure raises here instead of leaving the map waiting
        with ProcessPoolExecutor(workers, mp_context=mp.get_context('spawn')) as ex:
            return list(ex.map(fn, items, chunksize=max(1, len(items) // (4 * workers))))
    finally:
        main.__dict__.pop('__spec__', None)
        main.__dict__.update(saved)


def _batched_one(args):
    s, nz_range, net = args
    rng = np.random.default_rng(s)
    nz = int(rng.integers(nz_range[0], nz_range[1] + 1))
    T_top, T_bot = rng.uniform(400., 2500., 2)
    K = 10.0 ** rng.uniform(5, 11)
    prof = iso_gradient_profile(nz, T_top, T_bot)
    return _case(net, nz, T_bot, T_top, K, None, rng, Config(), tau=True, T_profile=prof)


def batched_cases(ncols: int = 256, net: Optional[Network] = None, nz_range=(80, 200),
                  seed: int = SEED_BATCH) -> List[ColumnCase]:
    net = net or full_network()
    out = _pool_map(_batched_one, [(s, nz_range, net) for s in np.random.SeedSequence(seed).spawn(ncols)])
    for c in out:
        c.net = net
    return out


def batched_case(index: int, ncols: int = 256, net: Optional[Network] = None, nz_range=(80, 200),
                 seed: int = SEED_BATCH) -> ColumnCase:
    """Column `index` of batched_cases(ncols) alone (same seed stream)."""
    net = net or full_network()
    c = _batched_one((np.random.SeedSequence(seed).spawn(ncols)[index], nz_range, net))
    c.net = net
    return c


def _mixed_one(args):
    s, net = args
    rng = np.random.default_rng(s)
    nz = int(rng.integers(100, 151))
    T_top, T_bot = rng.uniform(400., 2500., 2)
    K = 10.0 ** rng.uniform(5, 11)
    prof = iso_gradient_profile(nz, T_top, T_bot)
    return _case(net, nz, T_bot, T_top, K, None, rng, Config(), tau=False, T_profile=prof)


def mixed_cases(ncols: int = 512, nets: Optional[List[Network]] = None,
                seed: int = SEED_MIXED) -> List[ColumnCase]:
    """Column c uses network c % 3."""
    nets = nets or mixed_networks()
    seeds = np.random.SeedSequence(seed).spawn(ncols)
    out = _pool_map(_mixed_one, [(s, nets[c % len(nets)]) for c, s in enumerate(seeds)])
    for c, case in enumerate(out):
        case.net = nets[c % len(nets)]
    return out
