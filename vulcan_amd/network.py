"""Chemical network description and the device tables derived from it.

A network is a list of reversible reaction *pairs*.  Pair ``p`` owns two rate
rows, forward ``k[:, 2p]`` and reverse ``k[:, 2p+1]``; VULCAN numbers the same
rows 1-based as ``j = 2p+1`` (forward, odd) and ``j+1`` (reverse, even)
(make_chem_funs.py:183-184, op.py:189).  Species are ordered by first
appearance when a network is parsed from text, as in make_chem_funs.py:176-181.

Each side of a pair is a list of up to four *terms* in the order written in the
network line.  A term is a species index, or ``TERM_M`` for the third body
``M``.  The rate of a pair is::

    v_p = k_f * prod(reactant terms) - k_r * prod(product terms)

evaluated left to right with ``M`` substituted by the layer's number density,
which is how the generated ``v_j`` lambdas of chem_funs.py evaluate
(make_chem_funs.py:293-315).  ``dydt[i]`` accumulates ``-v_p`` once per
reactant occurrence of ``i`` and ``+v_p`` once per product occurrence, in
reaction order (make_chem_funs.py:263-291).
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence

import numpy as np

TERM_M = -1       # the third body M
TERM_NONE = -2    # unused term slot
MAX_TERMS = 4

# Forward-rate forms (op.py:145-207) plus prescribed photolysis (op.py:2786).
KIND_ARRH = 0     # k = A T^n exp(-E/T)                         op.py:166
KIND_LIND = 1     # k0/(1 + k0 M/kinf), Lindemann three-body    op.py:176-182
KIND_K0 = 2       # three-body without high-pressure limit      op.py:185-186
KIND_TROE = 3     # Lindemann times the natural-log broadening  op.py:203-207
KIND_PHOTO = 4    # k = J0 exp(-tau(z)) (prescribed photolysis)
KIND_CH3OH = 5    # hard-coded OH + CH3 + M -> CH3OH + M        op.py:197-207

# par columns
P_A, P_N, P_E, P_AINF, P_NINF, P_EINF, P_FC, P_J0 = range(8)
NPAR = 8

ATOMIC_MASS = {
    'H': 1.008, 'O': 15.999, 'C': 12.011, 'N': 14.007, 'He': 4.0026,
    'S': 32.06, 'P': 30.974, 'Si': 28.085, 'Cl': 35.45, 'Na': 22.99,
    'K': 39.098, 'Fe': 55.845, 'Ar': 39.948, 'Ti': 47.867, 'V': 50.942,
    'Mg': 24.305, 'Ca': 40.078, 'e': 0.0,
    # stand-in elements for the "sulfur-like" extra elements of large
    # synthetic networks
    'X': 32.06, 'Y': 30.974, 'Z': 35.45,
}


@dataclasses.dataclass
class Network:
    species: List[str]
    elements: List[str]
    compo: np.ndarray          # (ni, ne) int  atoms of each element per species
    mass: np.ndarray           # (ni,)        molar mass (g/mol)
    nasa9: np.ndarray          # (ni, 20)     NASA9 low [0:10], high [10:20]
    reac: np.ndarray           # (R, 4) int32 reactant terms (TERM_M/TERM_NONE)
    prod: np.ndarray           # (R, 4) int32 product terms
    kind: np.ndarray           # (R,)  int32 forward-rate form
    par: np.ndarray            # (R, 8) float64 rate parameters
    rev: np.ndarray            # (R,)  int32 1 = reverse rate from K_eq
    labels: List[str] = dataclasses.field(default_factory=list)
    atom_list: Optional[List[str]] = None   # elements checked by the loss monitor

    @property
    def ni(self) -> int:
        return len(self.species)

    @property
    def npairs(self) -> int:
        return int(self.reac.shape[0])

    @property
    def nr(self) -> int:
        """Number of rate rows (VULCAN's ``nr``: forward plus reverse)."""
        return 2 * self.npairs

    def atoms(self) -> List[str]:
        return list(self.atom_list) if self.atom_list is not None else [
            e for e in self.elements if e not in ('He', 'e')]

    def atom_matrix(self) -> np.ndarray:
        """(ni, n_atom) float composition restricted to the monitored atoms."""
        cols = [self.elements.index(a) for a in self.atoms()]
        return self.compo[:, cols].astype(np.float64)

    def equation(self, p: int) -> str:
        def side(t):
            return ' + '.join('M' if x == TERM_M else self.species[x]
                              for x in t if x != TERM_NONE)
        return side(self.reac[p]) + ' -> ' + side(self.prod[p])

    def validate(self) -> None:
        ni = self.ni
        assert self.compo.shape[0] == ni and self.mass.shape == (ni,)
        assert self.nasa9.shape == (ni, 20)
        for arr in (self.reac, self.prod):
            assert arr.shape == (self.npairs, MAX_TERMS)
            assert np.all((arr >= TERM_NONE) & (arr < ni))
        assert self.kind.shape == (self.npairs,) and self.rev.shape == (self.npairs,)
        assert self.par.shape == (self.npairs, NPAR)


# ---------------------------------------------------------------------------
# Tables consumed by the device kernels
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class NetworkTables:
    """Integer/float tables uploaded once per network.

    chem CSR: for species ``i`` the entries ``chem_ptr[i]:chem_ptr[i+1]`` list
    ``(pair, sign)`` per occurrence, in the reference's summation order.

    Jacobian: ``slot_pair/slot_sp`` enumerate the distinct (pair, species)
    derivative slots ``dv_p/dy_m``; the dense ni x ni entry ``(i, m)`` gathers
    ``sum coeff * dv[slot]`` over ``jac_ptr[e]:jac_ptr[e+1]`` with
    ``e = jac_map[i*ni+m]`` (``-1`` = structural zero).
    """
    chem_ptr: np.ndarray
    chem_pair: np.ndarray
    chem_sign: np.ndarray
    slot_pair: np.ndarray
    slot_sp: np.ndarray
    jac_map: np.ndarray
    jac_ptr: np.ndarray
    jac_slot: np.ndarray
    jac_coef: np.ndarray
    nslot: int
    njnz: int


def build_tables(net: Network) -> NetworkTables:
    ni, R = net.ni, net.npairs
    chem_lists: List[list] = [[] for _ in range(ni)]
    for p in range(R):
        for t in net.reac[p]:
            if t >= 0:
                chem_lists[t].append((p, -1))
        for t in net.prod[p]:
            if t >= 0:
                chem_lists[t].append((p, +1))
    chem_ptr = np.zeros(ni + 1, np.int32)
    for i in range(ni):
        chem_ptr[i + 1] = chem_ptr[i] + len(chem_lists[i])
    flat = [e for lst in chem_lists for e in lst]
    chem_pair = np.array([e[0] for e in flat], np.int32)
    chem_sign = np.array([e[1] for e in flat], np.int32)

    slot_pair, slot_sp = [], []
    slot_of = {}
    for p in range(R):
        seen = []
        for t in list(net.reac[p]) + list(net.prod[p]):
            if t >= 0 and t not in seen:
                seen.append(t)
        for m in seen:
            slot_of[(p, m)] = len(slot_pair)
            slot_pair.append(p)
            slot_sp.append(m)

    # net stoichiometric coefficient of species i in pair p
    entries = {}
    for p in range(R):
        coef = {}
        for t in net.reac[p]:
            if t >= 0:
                coef[t] = coef.get(t, 0) - 1
        for t in net.prod[p]:
            if t >= 0:
                coef[t] = coef.get(t, 0) + 1
        sp_in = sorted(set(int(t) for t in list(net.reac[p]) + list(net.prod[p]) if t >= 0))
        for i, c in coef.items():
            if c == 0:
                continue
            for m in sp_in:
                entries.setdefault((int(i), m), []).append((slot_of[(p, m)], c))
    keys = sorted(entries.keys())
    jac_map = -np.ones(ni * ni, np.int32)
    jac_ptr = np.zeros(len(keys) + 1, np.int32)
    jac_slot, jac_coef = [], []
    for e, key in enumerate(keys):
        jac_map[key[0] * ni + key[1]] = e
        lst = entries[key]
        jac_ptr[e + 1] = jac_ptr[e] + len(lst)
        jac_slot.extend(s for s, _ in lst)
        jac_coef.extend(float(c) for _, c in lst)
    return NetworkTables(
        chem_ptr=chem_ptr, chem_pair=chem_pair, chem_sign=chem_sign,
        slot_pair=np.array(slot_pair, np.int32), slot_sp=np.array(slot_sp, np.int32),
        jac_map=jac_map, jac_ptr=jac_ptr,
        jac_slot=np.array(jac_slot, np.int32), jac_coef=np.array(jac_coef, np.float64),
        nslot=len(slot_pair), njnz=len(keys))


# ---------------------------------------------------------------------------
# Seeded synthetic networks
# ---------------------------------------------------------------------------

def _formula(vec: Sequence[int], elements: Sequence[str]) -> str:
    s = ''
    for n, e in zip(vec, elements):
        if n == 1:
            s += e
        elif n > 1:
            s += e + str(n)
    return s


def _species_pool(elements: Sequence[str], max_count: Sequence[int], max_atoms: int):
    pool = []
    ranges = [range(m + 1) for m in max_count]
    import itertools
    for vec in itertools.product(*ranges):
        tot = sum(vec)
        if 1 <= tot <= max_atoms:
            pool.append(tuple(vec))
    return pool


def synthetic_network(ni: int, n_bi: int, n_tb: int = 0, n_photo: int = 0,
                      seed: int = 1001, extra_elements: int = 0,
                      hs: float = 4000.0, ss: float = 2.0,
                      bond: float = 2.5e4) -> Network:
    """Seeded element-conserving network.

    * ``ni`` species (one of them inert ``He``), drawn from small compositions
      over H, O, C, N (plus ``extra_elements`` stand-in elements X, Y, Z).
    * ``n_bi`` reversible A + B -> C + D pairs, drawn uniformly from all
      element-conserving quadruples: A log-uniform on [1e-14, 1e-9] cm3/s,
      n uniform on [-2, 2], E uniform on [0, 30000] K.
    * ``n_tb`` termolecular A + B + M -> AB + M pairs with Troe falloff:
      k0 log-uniform on [1e-33, 1e-28] cm6/s, kinf log-uniform on
      [1e-12, 1e-9] cm3/s, Fc uniform on [0.3, 0.8].
    * ``n_photo`` photolysis channels AB -> A + B, J0 log-uniform on
      [1e-6, 1e-3] 1/s, no reverse rate (they sit after "reverse stops").
    * NASA9-style Gibbs coefficients (identical low and high sets, so dG(T)
      is smooth): element group values that cancel in conserving reactions,
      a bond enthalpy of ``bond`` K per bond, and per-species perturbations
      of ``hs`` K (enthalpy) and ``ss`` (entropy).  Each A + B -> C + D pair
      is oriented so that its forward direction is exothermic; then
      k_r = k_f / K_eq stays within the entropy factor of k_f, as detailed
      balance with physical activation energies would give.
    """
    rng = np.random.default_rng(np.random.SeedSequence(seed))
    base = ['H', 'O', 'C', 'N']
    extra = ['X', 'Y', 'Z'][:extra_elements]
    elements = base + extra + ['He']
    nel = len(elements)
    max_count = [6, 3, 3, 2] + [2] * len(extra)
    pool = _species_pool(base + extra, max_count, 6)
    must = [(1, 0, 0, 0), (2, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1),
            (2, 1, 0, 0), (0, 1, 1, 0), (4, 0, 1, 0), (3, 0, 0, 1)]
    must = [m + (0,) * len(extra) for m in must]
    for e in range(len(extra)):
        v = [0] * (4 + len(extra))
        v[4 + e] = 1
        must.append(tuple(v))
        v2 = list(v)
        v2[0] = 2
        must.append(tuple(v2))
    # small molecules first: fill by total atom count, random within a level,
    # so that many element-conserving reactions exist among the species
    # pure-hydrogen species stop at H2, so that H2 is the bulk gas under
    # hydrogen-dominated element abundances
    rest = [v for v in pool if v not in must and not (v[0] > 2 and sum(v) == v[0])]
    keys = rng.random(len(rest))
    rest = [v for _, v in sorted(zip(keys, rest), key=lambda t: (sum(t[1]), t[0]))]
    n_chem = ni - 1
    chosen = must[:n_chem] + rest[:max(0, n_chem - len(must))]
    assert len(chosen) == n_chem, "not enough candidate compositions"
    comps = [tuple(v) + (0,) for v in chosen] + [tuple([0] * (nel - 1) + [1])]
    comp_arr = np.array(comps, np.int64)
    names = [_formula(v, elements) for v in comps]
    sp_index = {c: i for i, c in enumerate(comps)}

    # element-conserving candidate reactions
    sums = {}
    for a in range(n_chem):
        for b in range(a, n_chem):
            s = tuple(comp_arr[a] + comp_arr[b])
            sums.setdefault(s, []).append((a, b))
    quads = []
    for s, pairs in sums.items():
        for x in range(len(pairs)):
            for y in range(x + 1, len(pairs)):
                quads.append(pairs[x] + pairs[y])
    assert len(quads) >= n_bi, f"only {len(quads)} conserving quadruples"
    pick = rng.choice(len(quads), size=n_bi, replace=False)
    triples = []
    for a in range(n_chem):
        for b in range(a, n_chem):
            s = tuple(comp_arr[a] + comp_arr[b])
            if s in sp_index:
                triples.append((a, b, sp_index[s]))
    need = n_tb + n_photo
    assert len(triples) >= max(n_tb, n_photo), "not enough association triples"
    tb_pick = rng.choice(len(triples), size=n_tb, replace=False) if n_tb else []
    ph_pick = rng.choice(len(triples), size=n_photo, replace=False) if n_photo else []

    # NASA9-style coefficients a = [a0..a6, 0, b1, b2], g/RT = h/RT - s/R.
    # Element group values cancel in conserving reactions; a bond term of
    # ``bond`` K per bond (atoms - 1) makes associations exothermic; the
    # per-species perturbations set the spread of K_eq.
    grp_a2 = rng.uniform(1.0, 2.0, nel)
    grp_a8 = rng.uniform(-2e4, 2e4, nel)
    grp_a9 = rng.uniform(5.0, 15.0, nel)
    nasa9 = np.zeros((len(comps), 20))
    for i, c in enumerate(comp_arr):
        a = np.zeros(10)
        a[2] = 1.5 + c @ grp_a2 + rng.uniform(-0.5, 0.5)
        a[3] = rng.uniform(-1e-4, 1e-4)
        a[4] = rng.uniform(-2e-8, 2e-8)
        a[8] = c @ grp_a8 - bond * max(int(c.sum()) - 1, 0) + rng.uniform(-hs, hs)
        a[9] = c @ grp_a9 + rng.uniform(-ss, ss)
        nasa9[i, :10] = a
        nasa9[i, 10:] = a
    h0 = nasa9[:, 8]

    R = n_bi + n_tb + n_photo
    reac = np.full((R, MAX_TERMS), TERM_NONE, np.int32)
    prod = np.full((R, MAX_TERMS), TERM_NONE, np.int32)
    kind = np.zeros(R, np.int32)
    par = np.zeros((R, NPAR))
    rev = np.ones(R, np.int32)
    p = 0
    for q in pick:
        a, b, c, d = quads[q]
        if rng.random() < 0.5:
            a, b = b, a
        if rng.random() < 0.5:
            c, d = d, c
        # forward = exothermic direction, so k_r = k_f / K_eq stays below k_f
        # up to the entropy factor
        if h0[c] + h0[d] > h0[a] + h0[b]:
            a, b, c, d = c, d, a, b
        reac[p, :2] = (a, b)
        prod[p, :2] = (c, d)
        kind[p] = KIND_ARRH
        par[p, P_A] = 10.0 ** rng.uniform(-14, -9)
        par[p, P_N] = rng.uniform(-2, 2)
        par[p, P_E] = rng.uniform(0, 30000)
        p += 1
    for q in tb_pick:
        a, b, c = triples[q]
        reac[p, :3] = (a, b, TERM_M)
        prod[p, :2] = (c, TERM_M)
        kind[p] = KIND_TROE
        par[p, P_A] = 10.0 ** rng.uniform(-33, -28)
        par[p, P_AINF] = 10.0 ** rng.uniform(-12, -9)
        par[p, P_FC] = rng.uniform(0.3, 0.8)
        p += 1
    for q in ph_pick:
        a, b, c = triples[q]
        reac[p, 0] = c
        prod[p, :2] = (a, b)
        kind[p] = KIND_PHOTO
        par[p, P_J0] = 10.0 ** rng.uniform(-6, -3)
        rev[p] = 0
        p += 1

    mass = np.array([sum(n * ATOMIC_MASS[e] for n, e in zip(c, elements)) for c in comps])
    net = Network(species=names, elements=elements, compo=comp_arr.astype(np.int32),
                  mass=mass, nasa9=nasa9, reac=reac, prod=prod, kind=kind, par=par,
                  rev=rev, atom_list=base + extra)
    net.labels = [net.equation(i) for i in range(R)]
    net.validate()
    return net
