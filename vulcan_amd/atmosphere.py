"""Setup of one atmospheric column (host side, runs once per column).

These are the inputs every hot-path kernel reads.  The formulas follow the
reference's setup code:

* pressure grid ``pco`` log-spaced from ``P_b`` to ``P_t`` (store.py:116) and
  interface pressures ``pico`` (build_atm.py:310-325);
* ``M = n_0 = pco / (kb * Tco)`` (build_atm.py:475-476);
* molecular diffusion ``Dzz``, thermal-diffusion factors ``alpha`` and molar
  masses ``ms`` for an H2 background (build_atm.py:660-709);
* mean molar mass, gravity, scale height, layer thickness and interface
  quantities ``dzi``, ``Ti``, ``Hpi`` (build_atm.py:515-615).

The per-step recomputation of the last group (``update_mu_dz``, op.py:944-984)
runs on the device; this module provides the initial state only.
"""
from __future__ import annotations

import dataclasses
from typing import Optional

import numpy as np

from .config import Config, kb, Navo
from .network import Network


@dataclasses.dataclass
class Column:
    pco: np.ndarray
    pico: np.ndarray
    Tco: np.ndarray
    Kzz: np.ndarray      # (nz-1,) on interfaces
    vz: np.ndarray       # (nz-1,)
    M: np.ndarray        # (nz,)   n_0 = M
    Dzz: np.ndarray      # (nz-1, ni)
    alpha: np.ndarray    # (ni,)
    ms: np.ndarray       # (ni,)
    mu: np.ndarray
    g: np.ndarray
    Hp: np.ndarray
    dz: np.ndarray
    zco: np.ndarray
    dzi: np.ndarray
    Ti: np.ndarray
    Hpi: np.ndarray
    gs: float
    Rp: float
    pref_indx: int
    tau: Optional[np.ndarray] = None   # (nz,) photolysis optical depth
    # optional transport terms and boundary conditions (store.py:131-141)
    vm: Optional[np.ndarray] = None        # (nz, ni) molecular-diffusion velocity (use_vm_mol)
    top_flux: Optional[np.ndarray] = None  # (ni,) constant top flux, cm^-2 s^-1 (use_topflux, diff_esc)
    bot_flux: Optional[np.ndarray] = None  # (ni,) bottom flux (use_botflux)
    bot_vdep: Optional[np.ndarray] = None  # (ni,) deposition velocity, cm s^-1 (use_botflux)

    @property
    def nz(self) -> int:
        return int(self.Tco.shape[0])

    @property
    def n_0(self) -> np.ndarray:
        return self.M


def f_pico(pco: np.ndarray) -> np.ndarray:
    """Interface pressures (build_atm.py:310-325)."""
    pco_up1 = np.roll(pco, 1)
    pi = (pco * pco_up1) ** 0.5
    pi[0] = pco[0] ** 1.5 * pco[1] ** (-0.5)
    return np.append(pi, pco[-1] ** 1.5 * pco[-2] ** (-0.5))


def mol_diff_H2(Tco, n_0, species, mass, use_moldiff=True):
    """Dzz on interfaces, alpha and ms for an H2 background (build_atm.py:660-709)."""
    nz, ni = Tco.shape[0], len(species)
    ms = np.asarray(mass, dtype=np.float64).copy()
    alpha = np.zeros(ni)
    Dzz = np.zeros((nz - 1, ni))
    if not use_moldiff:
        return Dzz, alpha, ms
    Tco_i = np.delete((Tco + np.roll(Tco, 1)) * 0.5, 0)
    n0_i = np.delete((n_0 + np.roll(n_0, 1)) * 0.5, 0)
    if 'H' in species:
        alpha[species.index('H')] = -0.1
    if 'He' in species:
        alpha[species.index('He')] = 0.145
    for i in range(ni):
        if ms[i] > 4.:
            alpha[i] = 0.25
    for i in range(ni):
        mi = ms[i]
        Dzz[:, i] = 2.2965E17 * Tco_i ** 0.765 / n0_i * (16.04 / mi * (mi + 2.016) / 18.059) ** 0.5
    return Dzz, alpha, ms


def mol_velocity_H2(Tco, n_0, ms, alpha, g, Hp, dz) -> np.ndarray:
    """Molecular-diffusion velocity on the layers for use_vm_mol
    (build_atm.py:731-740), with the layer-centred H2-background Dzz
    (build_atm.py:666, 705).  The temperature step of the top layer wraps
    round to the bottom layer (np.roll), as the reference's does."""
    nz, ni = Tco.shape[0], ms.shape[0]
    Dzz_cen = np.zeros((nz, ni))
    for i in range(ni):
        mi = ms[i]
        Dzz_cen[:, i] = 2.2965E17 * Tco ** 0.765 / n_0 * (16.04 / mi * (mi + 2.016) / 18.059) ** 0.5
    delta_T = np.roll(Tco, -1) - Tco
    delta_T[0] = delta_T[1]
    return -Dzz_cen * (ms[np.newaxis, :] * g[:, np.newaxis] / (Navo * kb * Tco[:, np.newaxis])
                       - 1. / Hp[:, np.newaxis] + alpha / Tco[:, np.newaxis] * (delta_T[:, np.newaxis]) / dz[:, np.newaxis])


def read_bc_flux(species, top_file=None, bot_file=None):
    """Boundary-condition files (build_atm.py:743-773): the top file has
    lines ``species flux``, the bottom file ``species flux v_dep [fix_mix]``
    ('#' starts a comment).  Returns (top_flux, bot_flux, bot_vdep,
    bot_fix_sp), each (ni,), zero where a file names no value."""
    ni = len(species)
    index = {s: i for i, s in enumerate(species)}
    top, bot, vdep, fix = np.zeros(ni), np.zeros(ni), np.zeros(ni), np.zeros(ni)

    def rows(path):
        with open(path) as f:
            for line in f:
                if not line.startswith('#') and line.strip():
                    yield line.split()
    if top_file:
        for li in rows(top_file):
            top[index[li[0]]] = float(li[1])
    if bot_file:
        for li in rows(bot_file):
            bot[index[li[0]]] = float(li[1])
            vdep[index[li[0]]] = float(li[2])
            if len(li) > 3:
                fix[index[li[0]]] = float(li[3])
    return top, bot, vdep, fix


def thermal_alpha(species, mass) -> np.ndarray:
    """Thermal diffusion factors for an H2 background (build_atm.py:685-690)."""
    return mol_diff_H2(np.array([1000., 1000.]), np.array([1., 1.]), list(species), mass, True)[1]


def mu_g_dz(ymix, ms, Tco, pico, gs, Rp, pref_indx):
    """Mean mass, gravity, scale height and layer geometry (build_atm.py:515-615)."""
    nz = Tco.shape[0]
    mu = np.zeros(nz)
    for i in range(ymix.shape[1]):
        mu += ms[i] * ymix[:, i]
    g = np.zeros(nz)
    Hp = np.zeros(nz)
    dz = np.zeros(nz)
    zco = np.zeros(nz + 1)
    for i in range(pref_indx, nz):
        if i == pref_indx:
            g[i] = gs
            Hp[i] = kb * Tco[i] / (mu[i] / Navo * gs)
        else:
            g[i] = gs * (Rp / (Rp + zco[i])) ** 2
            Hp[i] = kb * Tco[i] / (mu[i] / Navo * g[i])
        dz[i] = Hp[i] * np.log(pico[i] / pico[i + 1])
        zco[i + 1] = zco[i] + dz[i]
    if pref_indx != 0:
        for i in range(pref_indx - 1, -1, -1):
            g[i] = gs * (Rp / (Rp + zco[i + 1])) ** 2
            Hp[i] = kb * Tco[i] / (mu[i] / Navo * g[i])
            dz[i] = Hp[i] * np.log(pico[i] / pico[i + 1])
            zco[i] = zco[i + 1] - dz[i]
    dzi = (0.5 * (dz + np.roll(dz, 1)))[1:]
    Ti = (0.5 * (Tco + np.roll(Tco, -1)))[:-1]
    Hpi = (0.5 * (Hp + np.roll(Hp, -1)))[:-1]
    return mu, g, Hp, dz, zco, dzi, Ti, Hpi


def reference_level(pico: np.ndarray, P_b: float, rocky: bool = False) -> int:
    """Index of the interface closest to 1 bar for gas giants (build_atm.py:527-530)."""
    if not rocky and P_b >= 1e6:
        return min(range(len(pico)), key=lambda i: abs(np.log10(pico[i]) - 6.))
    return 0


def build_column(net: Network, nz: int, T_bot: float, T_top: float,
                 P_b: float = 1e9, P_t: float = 1e4,
                 Kzz_bot: float = 1e7, Kzz_top: Optional[float] = None,
                 ymix0: Optional[np.ndarray] = None, cfg: Optional[Config] = None,
                 tau_top: Optional[float] = None, tau_bot: Optional[float] = None,
                 T_profile: Optional[np.ndarray] = None) -> Column:
    """Column with T linear in layer index from ``T_bot`` (layer 0) to ``T_top``.

    ``Kzz`` is constant (``Kzz_top is None``) or log-linear from the bottom to
    the top interface.  ``ymix0`` fixes the mean molar mass used for the
    initial geometry (the reference uses the initial composition,
    build_atm.py:536-541).
    """
    cfg = cfg or Config()
    pco = np.logspace(np.log10(P_b), np.log10(P_t), nz)
    pico = f_pico(pco)
    Tco = np.array(T_profile, dtype=np.float64) if T_profile is not None else np.linspace(T_bot, T_top, nz)
    if Kzz_top is None:
        Kzz = np.repeat(float(Kzz_bot), nz - 1)
    else:
        Kzz = 10.0 ** np.linspace(np.log10(Kzz_bot), np.log10(Kzz_top), nz - 1)
    vz = np.zeros(nz - 1)
    M = pco / (kb * Tco)
    Dzz, alpha, ms = mol_diff_H2(Tco, M, net.species, net.mass, cfg.use_moldiff)
    if ymix0 is None:
        ymix0 = np.full((nz, net.ni), 1.0 / net.ni)
    pref = reference_level(pico, P_b, cfg.rocky)
    mu, g, Hp, dz, zco, dzi, Ti, Hpi = mu_g_dz(ymix0, ms, Tco, pico, cfg.gs, cfg.Rp, pref)
    tau = None
    if tau_top is not None:
        tau = 10.0 ** np.linspace(np.log10(tau_bot), np.log10(tau_top), nz)
    return Column(pco=pco, pico=pico, Tco=Tco, Kzz=Kzz, vz=vz, M=M, Dzz=Dzz,
                  alpha=alpha, ms=ms, mu=mu, g=g, Hp=Hp, dz=dz, zco=zco, dzi=dzi,
                  Ti=Ti, Hpi=Hpi, gs=cfg.gs, Rp=cfg.Rp, pref_indx=pref, tau=tau)


def initial_mixing(ni: int, nz: int, rng: np.random.Generator,
                   lo: float = 1e-12, hi: float = 1e-1) -> np.ndarray:
    """Layer-uniform mixing ratios log-uniform on [lo, hi], normalised to 1."""
    x = 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), ni)
    x = x / x.sum()
    return np.tile(x, (nz, 1))
