"""Reader for VULCAN's network text format, NASA9 files and all_compose.txt.

It replaces the reference's code generation (make_chem_funs.py) and rate
reader (op.ReadRate) with device tables:

* reaction lines ``id [ A + B + M -> C + M ] A n E [Ainf ninf Einf]`` are
  numbered 1, 3, 5, ... in file order (make_chem_funs.py:17-110);
* section headers switch the rate form exactly as op.py:94-141 does:
  ``# 3-body`` (Lindemann with a high-pressure limit),
  ``# 3-body reactions without high-pressure rates`` (k0 only),
  ``# special`` (the hard-coded OH + CH3 + M -> CH3OH + M form),
  ``# condensation`` (rates zero unless condensation is on),
  ``# radiative``, ``# photo`` and ``# ionisation`` (rates zero until
  photolysis rates are prescribed), ``# reverse stops`` (pairs from that
  point on have no reverse rate, op.py:287-292);
* species are numbered by first appearance (make_chem_funs.py:176-181);
  photolysis reactions are part of the network only when ``use_photo``
  (make_chem_funs.py:75-78);
* ``2*X`` stoichiometry is expanded into repeated terms.
"""
from __future__ import annotations

import os
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import network as nw


def _parse_equation(eq: str):
    reac, prod = [], []
    side = reac
    for term in eq.split():
        if term == '+':
            continue
        if term == '->':
            side = prod
            continue
        parts = term.split('*')
        if len(parts) == 1:
            side.append(parts[0])
        else:
            side.extend([parts[1]] * int(parts[0]))
    return reac, prod


def parse_network(text: str, use_photo: bool = True, remove_list: Sequence[int] = ()):
    """Returns (equations, kinds, params, rev) lists in reaction order plus the
    species list; ``params`` rows are [A n E Ainf ninf Einf Fc J0]."""
    i = 1
    re_tri = re_tri_k0 = special = conden = photo = ion = False
    stop_rev = None
    eqs, kinds, pars, is_photo, index = [], [], [], [], []
    for line in text.splitlines(True):
        if line.startswith('# 3-body'):
            re_tri = True
        if line.startswith('# 3-body reactions without high-pressure rates'):
            re_tri_k0 = True
        elif line.startswith('# special'):
            re_tri = re_tri_k0 = False
            special = True
        elif line.startswith('# condensation'):
            re_tri = re_tri_k0 = special = False
            conden = True
        elif line.startswith('# radiative'):
            re_tri = re_tri_k0 = special = conden = False
        elif line.startswith('# photo'):
            re_tri = re_tri_k0 = special = conden = False
            photo = True
        elif line.startswith('# ionisation'):
            re_tri = re_tri_k0 = special = conden = photo = False
            ion = True
        elif line.startswith('# reverse stops'):
            stop_rev = i
        if line.startswith('#') or not line.strip():
            continue
        eq = line.partition('[')[-1].rpartition(']')[0].strip()
        cols = line.partition(']')[-1].split()
        par = np.zeros(nw.NPAR)
        if special:
            if eq != 'OH + CH3 + M -> CH3OH + M':
                raise NotImplementedError('special-form reaction without a known expression: ' + eq)
            kind = nw.KIND_CH3OH
        elif conden:
            kind = nw.KIND_ARRH          # k = 0 placeholder (op.py:215-224)
        elif photo or ion:
            kind = nw.KIND_PHOTO         # J prescribed later (op.py:2786)
        else:
            par[:3] = [float(c) for c in cols[:3]]
            if re_tri and not re_tri_k0:
                par[3:6] = [float(c) for c in cols[3:6]]
                kind = nw.KIND_LIND
            elif re_tri_k0:
                kind = nw.KIND_K0
            else:
                kind = nw.KIND_ARRH
        if (photo or ion) and not use_photo:
            i += 2
            continue
        eqs.append(eq)
        kinds.append(kind)
        pars.append(par)
        is_photo.append(photo or ion)
        index.append(i)
        i += 2
    if stop_rev is None:
        stop_rev = i
    rev = [1 if (j < stop_rev and j not in remove_list and j + 1 not in remove_list) else 0 for j in index]
    for r, j in enumerate(index):
        if j in remove_list:
            pars[r] = np.zeros(nw.NPAR)
            kinds[r] = nw.KIND_ARRH
    return eqs, kinds, pars, rev, index


def read_nasa9(thermo_dir: str, species: Sequence[str]) -> np.ndarray:
    """thermo/NASA9/<sp>.txt, flattened: low [0:10], high [10:20] (chem_funs.py:2773-2778)."""
    out = np.zeros((len(species), 20))
    for s, sp in enumerate(species):
        a = np.loadtxt(os.path.join(thermo_dir, 'NASA9', sp + '.txt')).flatten()
        out[s] = a[:20]
    return out


def read_compose(path: str):
    """all_compose.txt: species, element counts, molar mass (build_atm.py:18-25)."""
    with open(path) as f:
        header = f.readline().split()
    elements = header[1:-1]
    rows = {}
    with open(path) as f:
        f.readline()
        for line in f:
            p = line.split()
            if not p:
                continue
            rows[p[0]] = ([int(float(x)) for x in p[1:1 + len(elements)]], float(p[1 + len(elements)]))
    return elements, rows


def load_network(network_file: str, thermo_dir: str, compose_file: Optional[str] = None,
                 use_photo: bool = True, atom_list: Sequence[str] = ('H', 'O', 'C', 'N', 'S'),
                 remove_list: Sequence[int] = ()) -> nw.Network:
    with open(network_file) as f:
        text = f.read()
    eqs, kinds, pars, rev, index = parse_network(text, use_photo, remove_list)
    species: List[str] = []
    sides = []
    for eq in eqs:
        r, p = _parse_equation(eq)
        for t in r + p:
            if t != 'M' and t not in species:
                species.append(t)
        sides.append((r, p))
    sp = {s: k for k, s in enumerate(species)}
    R = len(eqs)
    reac = np.full((R, nw.MAX_TERMS), nw.TERM_NONE, np.int32)
    prod = np.full((R, nw.MAX_TERMS), nw.TERM_NONE, np.int32)
    for q, (r, p) in enumerate(sides):
        if len(r) > nw.MAX_TERMS or len(p) > nw.MAX_TERMS:
            raise NotImplementedError('more than %d terms on one side: %s' % (nw.MAX_TERMS, eqs[q]))
        reac[q, :len(r)] = [nw.TERM_M if t == 'M' else sp[t] for t in r]
        prod[q, :len(p)] = [nw.TERM_M if t == 'M' else sp[t] for t in p]
    compose_file = compose_file or os.path.join(thermo_dir, 'all_compose.txt')
    elements, rows = read_compose(compose_file)
    compo = np.array([rows[s][0] for s in species], np.int32)
    mass = np.array([rows[s][1] for s in species])
    atoms = [a for a in atom_list if a in elements and compo[:, elements.index(a)].any()]
    net = nw.Network(species=species, elements=elements, compo=compo, mass=mass,
                     nasa9=read_nasa9(thermo_dir, species), reac=reac, prod=prod,
                     kind=np.array(kinds, np.int32), par=np.array(pars), rev=np.array(rev, np.int32),
                     labels=eqs, atom_list=atoms)
    net.validate()
    net.vulcan_index = np.array(index)
    return net


def set_photolysis(net: nw.Network, J0: Dict[int, float]) -> None:
    """Prescribe J0 (1/s) of photolysis rows by VULCAN index; k = J0 exp(-tau)."""
    idx = {int(j): r for r, j in enumerate(getattr(net, 'vulcan_index'))}
    for j, v in J0.items():
        r = idx[int(j)]
        if net.kind[r] != nw.KIND_PHOTO:
            raise ValueError('R%d is not a photolysis reaction' % j)
        net.par[r, nw.P_J0] = float(v)
