"""Builds oracle inputs from the reference-written fixtures in tests/golden."""
import os

import numpy as np

from vulcan_amd import network as nw
from vulcan_amd.atmosphere import Column
from vulcan_amd.config import Config

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name='ref_chemfuns_nz8.npz'):
    return np.load(os.path.join(HERE, 'golden', name))


def network_from(d) -> nw.Network:
    R = d['reac'].shape[0]
    net = nw.Network(species=[str(s) for s in d['species']], elements=[str(e) for e in d['elements']],
                     compo=d['compo'], mass=d['mass'], nasa9=d['nasa9'], reac=d['reac'],
                     prod=d['prod'], kind=np.zeros(R, np.int32), par=np.zeros((R, nw.NPAR)),
                     rev=np.ones(R, np.int32), atom_list=[str(a) for a in d['atom_list']])
    net.validate()
    return net


def column_from(d) -> Column:
    return Column(pco=d['pco'].copy(), pico=d['pico'].copy(), Tco=d['Tco'].copy(), Kzz=d['Kzz'].copy(),
                  vz=np.zeros(d['Kzz'].shape), M=d['M'].copy(), Dzz=d['Dzz'].copy(),
                  alpha=d['alpha'].copy(), ms=d['ms'].copy(), mu=np.zeros(d['Tco'].shape),
                  g=d['g0'].copy(), Hp=d['Hp0'].copy(), dz=d['dz0'].copy(), zco=d['zco0'].copy(),
                  dzi=d['dzi0'].copy(), Ti=d['Ti0'].copy(), Hpi=d['Hpi0'].copy(),
                  gs=float(d['gs']), Rp=float(d['Rp']), pref_indx=int(d['pref_indx']))


def traj_config(d) -> Config:
    upd, cmin, tmin, cstep, cmax = d['traj_cfg']
    return Config(update_frq=int(upd), count_min=int(cmin), trun_min=float(tmin),
                  conv_step=int(cstep), count_max=int(cmax))


def rel_err(a, b, floor=0.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide='ignore', invalid='ignore'):
        r = np.where(den > 0, np.abs(a - b) / den, np.abs(a - b))
    return float(np.max(r)) if r.size else 0.0


def network_fixture(name) -> nw.Network:
    """A real VULCAN network's device tables (tests/golden/make_network_tables.py)."""
    d = load(name)
    net = nw.Network(species=[str(s) for s in d['species']], elements=[str(e) for e in d['elements']],
                     compo=d['compo'], mass=d['mass'], nasa9=d['nasa9'], reac=d['reac'], prod=d['prod'],
                     kind=d['kind'], par=d['par'].copy(), rev=d['rev'], labels=[str(s) for s in d['labels']],
                     atom_list=[str(a) for a in d['atom_list']])
    net.validate()
    net.vulcan_index = d['vulcan_index']
    return net
