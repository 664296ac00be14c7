"""VULCAN `.vul` output for device runs (Output.save_out, op.py:3216-3256).

A `.vul` file is the pickle (protocol 4) of

    {'variable': var_save, 'atm': vars(atm), 'parameter': vars(para)}

where var_save holds 'species', 'nr', the keys of Variables.var_save
(store.py:86-88, without the photochemistry ones: radiative transfer is not
run here) and, with save_evolution, every save_evo_frq-th entry of the keys
of Variables.var_evol_save (store.py:92; op.py:3241-3247).  VULCAN's
plot_py readers index it as data['variable']['ymix'][:, species.index(sp)],
data['atm']['pco'], data['atm']['zco'], data['atm']['zmco'],
data['atm']['n_0'], data['variable']['t_time'] and
np.array(data['variable']['y_time'])[:, layer, species.index(sp)]
(plot_vulcan.py:64-86, plot_evolution.py:39-59).

`column_result` builds the (var, atm, para) objects of one column of a
ColumnBatch after integration, with the names and shapes of VULCAN's
Variables / AtmData / Parameters (store.py:21-208); `save_vul` writes them.
"""
from __future__ import annotations

import os
import pickle
from types import SimpleNamespace
from typing import Optional, Sequence

import numpy as np

# Variables.var_save without the photochemistry extension (store.py:86-90)
VAR_SAVE = ['k', 'y', 'ymix', 'y_ini', 't', 'dt', 'longdy', 'longdydt', 'atom_ini', 'atom_sum', 'atom_loss',
            'atom_conden', 'aflux_change', 'Rf']
# Variables.var_evol_save (store.py:92)
VAR_EVOL_SAVE = ['y_time', 't_time']


def save_dict(var, atm, para, species: Sequence[str], nr: int, save_evolution: bool = False,
              save_evo_frq: int = 10) -> dict:
    """The dictionary Output.save_out pickles (op.py:3237-3254)."""
    var_save = {'species': list(species), 'nr': int(nr)}
    for key in VAR_SAVE:
        var_save[key] = getattr(var, key)
    if save_evolution:
        for key in VAR_EVOL_SAVE:
            var_save[key] = np.array(getattr(var, key))[::save_evo_frq]
    return {'variable': var_save, 'atm': dict(vars(atm)), 'parameter': dict(vars(para))}


def save_vul(path: str, var, atm, para, species: Sequence[str], nr: int, save_evolution: bool = False,
             save_evo_frq: int = 10, human_readable: bool = False) -> str:
    """Write `path` as VULCAN does (pickle protocol 4, or the str() of the
    dictionary for output_humanread, op.py:3249-3254).  Returns the path."""
    d = save_dict(var, atm, para, species, nr, save_evolution, save_evo_frq)
    out_dir = os.path.dirname(os.path.abspath(path))
    os.makedirs(out_dir, exist_ok=True)
    if human_readable:
        with open(path, 'w') as f:
            f.write(str(d))
    else:
        with open(path, 'wb') as f:
            pickle.dump(d, f, protocol=4)
    return path


def _atom_dict(atoms, values) -> dict:
    return {a: float(v) for a, v in zip(atoms, values)}


def column_result(batch, c: int, col, y_ini: np.ndarray, y_time: Optional[np.ndarray] = None,
                  t_time: Optional[np.ndarray] = None):
    """(var, atm, para) of column `c` of a ColumnBatch, in VULCAN's layout.

    `col` is the column's atmosphere (vulcan_amd.atmosphere.Column) and
    `y_ini` its initial number densities (nz, ni).  The arrays that the
    integration updates on the device (update_mu_dz: g, Hp, dz, mu, zco,
    dzi, Hpi) are read back from the batch; y_time / t_time, when given,
    become the evolution arrays.
    """
    net = batch.plan.net
    ni, nz = net.ni, col.nz
    st = batch.host_state()[c]
    o0 = int(batch.layout.off[c])
    y = batch.column('y', c)
    ymix = batch.column('ymix', c)
    k = batch.column('k', c)
    atoms = net.atoms()
    amat = net.atom_matrix()
    atom_sum = (y @ amat).sum(axis=0)        # sum over layers and species (op.py:2480-2484)
    var = SimpleNamespace(
        k={j + 1: k[:, j].copy() for j in range(net.nr)},          # VULCAN's k dict, 1..nr
        y=y, ymix=ymix, y_ini=np.array(y_ini, dtype=np.float64),
        t=float(st['t']), dt=float(st['dt']), longdy=float(st['longdy']), longdydt=float(st['longdydt']),
        atom_ini=_atom_dict(atoms, batch.atom_ini.cpu().numpy()[c]),
        atom_sum=_atom_dict(atoms, atom_sum),
        atom_loss=_atom_dict(atoms, batch.atom_loss.cpu().numpy()[c]),
        atom_conden={a: 0.0 for a in atoms},
        aflux_change=0.0,
        Rf={2 * p + 1: net.equation(p) for p in range(net.npairs)},
        y_time=[] if y_time is None else y_time, t_time=[] if t_time is None else t_time)
    zco = batch.zco.cpu().numpy()
    zo = o0 + c                      # zco holds nz + 1 entries per column
    zc = zco[zo:zo + nz + 1]
    M = np.array(col.M, dtype=np.float64)
    atm = SimpleNamespace(
        pco=np.array(col.pco), pico=np.array(col.pico), dz=batch.column('dz', c),
        dzi=batch.column('dzi', c)[:nz - 1], zco=zc, zmco=0.5 * (zc[:-1] + zc[1:]), Tco=np.array(col.Tco),
        Kzz=np.array(col.Kzz), vz=np.array(col.vz), M=M, n_0=M.copy(), Hp=batch.column('Hp', c),
        mu=batch.column('mu', c), ms=np.array(col.ms), Dzz=np.array(col.Dzz), alpha=np.array(col.alpha),
        gs=float(col.gs), g=batch.column('g', c), Ti=np.array(col.Ti), Hpi=batch.column('Hpi', c)[:nz - 1],
        top_flux=np.zeros(ni), bot_flux=np.zeros(ni), bot_vdep=np.zeros(ni), bot_fix_sp=np.zeros(ni),
        sat_p={}, sat_mix={}, conden_min_lev={}, r_p={}, rho_p={})
    para = SimpleNamespace(
        nega_y=float(st['nega_y']), small_y=float(st['small_y']), delta=float(st['delta']),
        count=int(st['count']), nega_count=int(st['nega_count']), loss_count=int(st['loss_count']),
        delta_count=int(st['delta_count']), end_case=int(st['end_case']), solver_str='solver',
        switch_final_photo_frq=False, where_varies_most=np.zeros((nz, ni)), pic_count=0,
        fix_species_start=False)
    return var, atm, para
