"""Photochemistry radiative transfer: host set-up and the device update.

Set-up (host, once per run) restates ReadRate.make_bins_read_cross
(op.py:495-781) and Atm.read_sflux (build_atm.py:617-657):

* wavelength bins from the shortest to the longest wavelength any
  photolysed species absorbs at, bounded by the stellar spectrum (2 to
  700 nm at most) and the longest photodissociation threshold, uniform at
  ``dbin1`` below ``dbin_12trans`` and ``dbin2`` above it;
* absorption, photodissociation-branch (dissociation cross section times
  the branching ratio) and Rayleigh cross sections linearly interpolated on
  the bins (zero outside the tabulated range, branching ratios held at
  their end values), and the per-layer temperature-dependent tables of the
  species in ``T_cross_sp`` (log10-linear in T between the bracketing
  tables);
* the stellar flux scaled to the planet's orbit and interpolated on the
  bins.

The update (device, every ``update_photo_frq`` steps) is ``vk_photo_update``:
optical depth, two-stream fluxes, actinic flux and its change, and the
photolysis rates written into the rate array (op.py:2580-2787).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from . import network as nw

R_SUN = 6.957E10            # phy_const.py:9 (cm)
AU = 1.49597871E13          # phy_const.py:8 (cm)
HC = 1.98644582E-9          # phy_const.py:7 (erg nm)


@dataclasses.dataclass
class PhotoConfig:
    """vulcan_cfg.py's photochemistry knobs (names as there)."""
    cross_folder: str = 'thermo/photo_cross/'
    sflux_file: str = 'atm/stellar_flux/sflux-HD189_Moses11.txt'
    r_star: float = 0.805
    orbit_radius: float = 0.03142
    sl_angle: float = 48 / 180. * 3.14159
    f_diurnal: float = 1.
    scat_sp: Sequence[str] = ('H2', 'He')
    T_cross_sp: Sequence[str] = ()
    edd: float = 0.5
    dbin1: float = 0.1
    dbin2: float = 2.
    dbin_12trans: float = 240.
    ini_update_photo_frq: int = 100
    final_update_photo_frq: int = 5
    flux_atol: float = 1.
    remove_list: Sequence[int] = ()


@dataclasses.dataclass
class PhotoTables:
    """Device inputs of the radiative transfer (per-layer tables stacked over
    the columns of a batch)."""
    bins: np.ndarray
    din12: int
    sflux_top: np.ndarray
    abs_sp: List[int]
    abs_T: List[int]
    cross: np.ndarray            # [nabs][nbin]
    cross_T: np.ndarray          # [nT][L][nbin]
    scat_sp: List[int]
    cross_scat: np.ndarray       # [nscat][nbin]
    br_col: List[int]            # rate column (VULCAN index - 1) each branch's J goes to
    br_T: List[int]
    cross_J: np.ndarray          # [nbranch][nbin]
    cross_J_T: np.ndarray        # [nTJ][L][nbin]
    dbin1: float
    dbin2: float
    cos_sl: float
    edd: float
    f_diurnal: float
    flux_atol: float
    hc: float = HC
    # (species, branch) of each device branch, and the branch count of every
    # photolysed species (VULCAN's var.n_branch): J_sp and the .vul keys
    branches: Optional[List[Tuple[str, int]]] = None
    n_branch: Optional[Dict[str, int]] = None


def photo_branches(network_file: str) -> Dict[Tuple[str, int], int]:
    """(species, branch) -> VULCAN reaction index of the photodissociation
    rows (the ``# photo`` section: ``id [ A -> B + C ] A branch``,
    op.py:229-244)."""
    out = {}
    photo = False
    i = 1
    for line in open(network_file):
        if line.startswith('# photo'):
            photo = True
        elif line.startswith('# ionisation'):
            photo = False
        if line.startswith('#') or not line.strip():
            continue
        if photo:
            cols = line.partition(']')[-1].split()
            out[(cols[0], int(cols[1]))] = int(line.partition('[')[0].strip())
        i += 2
    return out


def _interp(x_new, x, y, below=0.0, above=0.0):
    """Linear interpolation with constant values outside [x[0], x[-1]] (the
    reference's interp1d(..., bounds_error=False, fill_value=...))."""
    order = np.argsort(x, kind='mergesort')
    x, y = np.asarray(x, dtype=np.float64)[order], np.asarray(y, dtype=np.float64)[order]
    out = np.interp(x_new, x, y)
    out[x_new < x[0]] = below
    out[x_new > x[-1]] = above
    return out


def _read_csv(path, ncols):
    return np.loadtxt(path, delimiter=',', skiprows=1, usecols=range(ncols), ndmin=2)


def _read_branch(path):
    with open(path) as f:
        f.readline()
        names = [n.strip() for n in f.readline().lstrip('#').split(',')]
    data = np.loadtxt(path, delimiter=',', skiprows=2, ndmin=2)
    return {n: data[:, q] for q, n in enumerate(names)}


def read_sflux(path: str):
    d = np.loadtxt(path, skiprows=1, ndmin=2)
    return d[:, 0], d[:, 1]


def build_tables(net: nw.Network, network_file: str, Tco_layers: np.ndarray, pcfg: PhotoConfig,
                 root: str = '.') -> PhotoTables:
    """The radiative-transfer tables of `net` for columns whose stacked layer
    temperatures are `Tco_layers` ([L]); data paths relative to `root` (the
    VULCAN directory)."""
    species = list(net.species)
    branches = photo_branches(os.path.join(root, network_file) if not os.path.isabs(network_file) else network_file)
    photo_sp = sorted({sp for sp, _ in branches})
    n_branch = {sp: max(b for s, b in branches if s == sp) for sp in photo_sp}
    folder = os.path.join(root, pcfg.cross_folder)
    # thresholds (the first column names, the second the wavelength)
    labels, lmd = [], []
    with open(os.path.join(folder, 'thresholds.txt')) as f:
        f.readline()
        for line in f:
            p = line.split()
            if len(p) >= 2 and not p[0].startswith('#'):
                labels.append(p[0]); lmd.append(float(p[1]))
    threshold = {s: v for s, v in zip(labels, lmd) if s in species}
    cross_raw, ratio_raw, cross_T_raw, T_lists = {}, {}, {}, {}
    bin_min, bin_max, diss_max = np.inf, -np.inf, -np.inf
    for sp in photo_sp:
        cross_raw[sp] = _read_csv(os.path.join(folder, sp, sp + '_cross.csv'), 3)
        ratio_raw[sp] = _read_branch(os.path.join(folder, sp, sp + '_branch.csv'))
        if sp in pcfg.T_cross_sp:
            tl = []
            for fn in os.listdir(os.path.join(folder, sp)):
                if fn.startswith(sp) and fn.endswith('K.csv'):
                    tl.append(int(fn.replace(sp, '').replace('_cross_', '').replace('K.csv', '')))
            for tt in tl:
                cross_T_raw[(sp, tt)] = _read_csv(os.path.join(folder, sp, '%s_cross_%dK.csv' % (sp, tt)), 3)
            cross_T_raw[(sp, 300)] = cross_raw[sp]
            T_lists[sp] = np.array(tl + [300])
        c = cross_raw[sp]
        if c[0, 1] == 0 or c[-1, 1] == 0:
            raise ValueError('zeros at the ends of the cross-section table of ' + sp)
        bin_min, bin_max = min(bin_min, c[0, 0]), max(bin_max, c[-1, 0])
        diss_max = max(diss_max, threshold[sp])
    sfl, sfv = read_sflux(os.path.join(root, pcfg.sflux_file))
    def_min, def_max = max(sfl[0], 2.), min(sfl[-1], 700.)
    bin_min = max(bin_min, def_min)
    bin_max = min(bin_max, def_max, diss_max)
    if pcfg.dbin_12trans >= bin_min and pcfg.dbin_12trans <= bin_max:
        bins = np.concatenate((np.arange(bin_min, pcfg.dbin_12trans, pcfg.dbin1),
                               np.arange(pcfg.dbin_12trans, bin_max, pcfg.dbin2)))
    else:
        bins = np.arange(bin_min, bin_max, pcfg.dbin1)
    nb = len(bins)
    hit = np.nonzero(bins == pcfg.dbin_12trans)[0]
    din12 = int(hit[-1]) if hit.size else -1
    scale = (pcfg.r_star * R_SUN / (AU * pcfg.orbit_radius)) ** 2
    sflux_top = _interp(bins, sfl, sfv * scale)
    L = len(Tco_layers)
    cross = np.zeros((len(photo_sp), nb))
    T_sp = [sp for sp in photo_sp if sp in pcfg.T_cross_sp]
    cross_T = np.zeros((len(T_sp), L, nb))
    blist = [(sp, b) for sp in photo_sp for b in range(1, n_branch[sp] + 1)
             if branches[(sp, b)] not in pcfg.remove_list]
    cross_J = np.zeros((len(blist), nb))
    TJ = [q for q, (sp, _) in enumerate(blist) if sp in pcfg.T_cross_sp]
    cross_J_T = np.zeros((len(TJ), L, nb))
    jcol = {}
    for a, sp in enumerate(photo_sp):
        c = cross_raw[sp]
        cross[a] = _interp(bins, c[:, 0], c[:, 1])
        diss = _interp(bins, c[:, 0], c[:, 2])
        rr = ratio_raw[sp]
        ratio = {b: _interp(bins, rr['lambda'], rr['br_ratio_%d' % b], rr['br_ratio_%d' % b][0],
                            rr['br_ratio_%d' % b][-1]) for b in range(1, n_branch[sp] + 1)}
        for q, (s, b) in enumerate(blist):
            if s == sp:
                cross_J[q] = diss * ratio[b]
                jcol[(s, b)] = q
        if sp in pcfg.T_cross_sp:
            t = T_sp.index(sp)
            _cross_T_tables(sp, t, bins, Tco_layers, T_lists[sp], cross_T_raw, cross[a], cross_J, cross_T,
                            cross_J_T, [(TJ.index(q), q, ratio[b]) for q, (s, b) in enumerate(blist) if s == sp])
    scat = [species.index(sp) for sp in pcfg.scat_sp]
    cross_scat = np.zeros((len(scat), nb))
    for q, sp in enumerate(pcfg.scat_sp):
        d = np.loadtxt(os.path.join(folder, 'rayleigh', sp + '_scat.txt'), skiprows=1, ndmin=2)
        cross_scat[q] = _interp(bins, d[:, 0], d[:, 1])
    # the reference indexes k by the id printed in the network file
    # (pho_rate_index, op.py:241-244) rather than by the line's position; an
    # id beyond the network's rates only fills the reference's dict
    keep = [q for q, (sp, b) in enumerate(blist) if 1 <= branches[(sp, b)] <= net.nr]
    br_col = [branches[blist[q]] - 1 for q in keep]
    return PhotoTables(bins=bins, din12=din12, sflux_top=sflux_top, abs_sp=[species.index(sp) for sp in photo_sp],
                       abs_T=[T_sp.index(sp) if sp in T_sp else -1 for sp in photo_sp], cross=cross,
                       cross_T=cross_T, scat_sp=scat, cross_scat=cross_scat, br_col=br_col,
                       br_T=[TJ.index(q) if q in TJ else -1 for q in keep], cross_J=cross_J[keep],
                       cross_J_T=cross_J_T, dbin1=pcfg.dbin1, dbin2=pcfg.dbin2, cos_sl=float(np.cos(pcfg.sl_angle)),
                       edd=pcfg.edd, f_diurnal=pcfg.f_diurnal, flux_atol=pcfg.flux_atol,
                       branches=[blist[q] for q in keep], n_branch=dict(n_branch))


def _log10_or(v):
    with np.errstate(divide='ignore'):
        r = np.log10(v)
    return np.where(np.isinf(r), -100., r)


def _cross_T_tables(sp, t, bins, Tco_layers, T_list, raw, room, cross_J, cross_T, cross_J_T, jbr):
    """Per-layer temperature-dependent absorption and branch cross sections
    of species sp (op.py:637-745)."""
    tmax, tmin = T_list.max(), T_list.min()
    for lev, Tz in enumerate(Tco_layers):
        lo_set, hi_set = T_list[T_list <= Tz], T_list[T_list > Tz]
        if lo_set.size and hi_set.size:
            Tlo, Thi = lo_set.max(), hi_set.min()
            a, b = raw[(sp, Tlo)], raw[(sp, Thi)]
            ld_min, ld_max = max(a[0, 0], b[0, 0]), min(a[-1, 0], b[-1, 0])
            out = (bins < ld_min) | (bins > ld_max)
            lo, hi = _log10_or(_interp(bins, a[:, 0], a[:, 1])), _log10_or(_interp(bins, b[:, 0], b[:, 1]))
            jl, jh = _log10_or(_interp(bins, a[:, 0], a[:, 2])), _log10_or(_interp(bins, b[:, 0], b[:, 2]))
            xT = np.array([Tlo, Thi])
            lt = np.array([np.interp(Tz, xT, [lo[n], hi[n]]) for n in range(len(bins))])
            ljt = np.array([np.interp(Tz, xT, [jl[n], jh[n]]) for n in range(len(bins))])
            ct = np.where(lt == -100., 0.0, 10 ** lt)
            cross_T[t, lev] = np.where(out, room, ct)
            for tj, q, ratio in jbr:
                cj = np.where(ljt == -100., 0.0, 10 ** ljt * ratio)
                cross_J_T[tj, lev] = np.where(out, cross_J[q], cj)
        else:
            edge = tmin if not T_list[T_list < Tz].size else tmax
            if edge == 300:
                cross_T[t, lev] = room
                for tj, q, _ in jbr:
                    cross_J_T[tj, lev] = cross_J[q]
            else:
                a = raw[(sp, edge)]
                out = (bins < a[0, 0]) | (bins > a[-1, 0])
                cross_T[t, lev] = np.where(out, room, _interp(bins, a[:, 0], a[:, 1]))
                dj = _interp(bins, a[:, 0], a[:, 2])
                for tj, q, ratio in jbr:
                    cross_J_T[tj, lev] = np.where(out, cross_J[q], dj * ratio)


def tables_from_var(var, species: Sequence[str], nr: int, cfg) -> PhotoTables:
    """PhotoTables from the arrays VULCAN's own set-up left in ``var``
    (make_bins_read_cross, op.py:495-781; read_sflux, build_atm.py:617-657):
    bins, sflux_top, sflux_din12_indx, cross, cross_T, cross_J, cross_J_T,
    cross_scat, photo_sp, pho_rate_index, dbin1, dbin2.  One column: the
    temperature-dependent tables are [nz][nbin]."""
    species = list(species)
    photo_sp = sorted(var.photo_sp)
    T_sp = [sp for sp in photo_sp if sp in cfg.T_cross_sp]
    blist = [b for b in sorted(var.pho_rate_index)
             if var.pho_rate_index[b] not in cfg.remove_list and 1 <= var.pho_rate_index[b] <= nr]
    TJ = [q for q, (sp, _) in enumerate(blist) if sp in T_sp]
    nb = len(var.bins)
    stackT = lambda rows: np.stack(rows) if rows else np.zeros((0, 1, nb))
    return PhotoTables(bins=np.asarray(var.bins, dtype=np.float64), din12=int(var.sflux_din12_indx),
                       sflux_top=np.asarray(var.sflux_top, dtype=np.float64),
                       abs_sp=[species.index(sp) for sp in photo_sp],
                       abs_T=[T_sp.index(sp) if sp in T_sp else -1 for sp in photo_sp],
                       cross=np.stack([var.cross[sp] for sp in photo_sp]),
                       cross_T=stackT([var.cross_T[sp] for sp in T_sp]),
                       scat_sp=[species.index(sp) for sp in cfg.scat_sp],
                       cross_scat=np.stack([var.cross_scat[sp] for sp in cfg.scat_sp]),
                       br_col=[var.pho_rate_index[b] - 1 for b in blist],
                       br_T=[TJ.index(q) if q in TJ else -1 for q in range(len(blist))],
                       cross_J=np.stack([var.cross_J[b] for b in blist]),
                       cross_J_T=stackT([var.cross_J_T[blist[q]] for q in TJ]),
                       dbin1=float(var.dbin1), dbin2=float(var.dbin2), cos_sl=float(np.cos(cfg.sl_angle)),
                       edd=float(cfg.edd), f_diurnal=float(cfg.f_diurnal), flux_atol=float(cfg.flux_atol),
                       branches=list(blist),
                       n_branch=dict(getattr(var, 'n_branch', None) or
                                     {sp: max(b for s, b in var.pho_rate_index if s == sp) for sp in photo_sp}))


class DevicePhoto:
    """vk_photo for the columns of a device batch (plan + layout)."""

    def __init__(self, plan, layout, tables: PhotoTables):
        lib = _lib.load()
        self._lib = lib
        self.tables = tables
        t = tables
        self._keep = []

        def ip(a):
            arr = np.ascontiguousarray(a, dtype=np.int32)
            self._keep.append(arr)
            return arr.ctypes.data_as(_lib.c_int_p)

        def dp(a):
            arr = np.ascontiguousarray(a, dtype=np.float64)
            self._keep.append(arr)
            return arr.ctypes.data_as(_lib.c_double_p)

        d = _lib.PhotoDesc(nbin=len(t.bins), nabs=len(t.abs_sp), nscat=len(t.scat_sp), nbranch=len(t.br_col),
                           nT=t.cross_T.shape[0], nTJ=t.cross_J_T.shape[0], din12=int(t.din12),
                           abs_sp=ip(t.abs_sp), abs_T=ip(t.abs_T), scat_sp=ip(t.scat_sp), br_col=ip(t.br_col),
                           br_T=ip(t.br_T), bins=dp(t.bins), sflux_top=dp(t.sflux_top), cross=dp(t.cross),
                           cross_T=dp(t.cross_T), cross_scat=dp(t.cross_scat), cross_J=dp(t.cross_J),
                           cross_J_T=dp(t.cross_J_T), dbin1=t.dbin1, dbin2=t.dbin2, cos_sl=t.cos_sl, edd=t.edd,
                           f_diurnal=t.f_diurnal, flux_atol=t.flux_atol, hc=t.hc)
        h = C.c_void_p()
        _lib.check(lib.vk_photo_create(plan.handle, layout.handle, C.byref(d), C.byref(h)), 'vk_photo_create')
        self.handle = h
        self.plan, self.layout = plan, layout
        self._keep = []

    def update(self, y, ymix, dz, k, state=None, force=True, ini_frq=100, final_frq=5, yconv_min=0.1):
        from .device import dptr, stream_handle
        _lib.check(self._lib.vk_photo_update(self.handle, self.plan.handle, self.layout.handle, dptr(y), dptr(ymix),
                                             dptr(dz), dptr(k), dptr(state), int(force), int(ini_frq),
                                             int(final_frq), float(yconv_min), stream_handle()), 'vk_photo_update')

    def read(self, what: int) -> np.ndarray:
        nb = len(self.tables.bins)
        rows = self.layout.L + (0 if what == VK_PHOTO_AFLUX else self.layout.ncols)
        out = np.zeros((rows, nb))
        _lib.check(self._lib.vk_photo_read(self.handle, int(what), out.ctypes.data), 'vk_photo_read')
        return out

    def close(self):
        if getattr(self, 'handle', None):
            self._lib.vk_photo_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


VK_PHOTO_TAU, VK_PHOTO_SFLUX, VK_PHOTO_DFLUX_U, VK_PHOTO_DFLUX_D, VK_PHOTO_AFLUX = range(5)
