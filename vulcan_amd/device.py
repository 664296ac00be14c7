"""Device-side objects: network plan, column layout and a batch of columns.

PyTorch-ROCm provides the device buffers and the stream; every computation
runs in the HIP library through the C ABI (vulcan_amd/_lib.py).  There is no
CPU fallback: without a HIP device or without the built library these
classes raise.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .atmosphere import Column, mol_velocity_H2, thermal_alpha
from .config import Config
from .network import Network, NetworkTables, build_tables

F64 = torch.float64


def _require_device():
    if not torch.cuda.is_available():
        raise _lib.VulcanAmdError('no HIP device visible: the vulcan_amd kernels need an MI355X (gfx950)')


FACTOR_NVARIANTS = 5     # VK_FACTOR_NVARIANTS


def stream_handle(stream: Optional[torch.cuda.Stream] = None) -> C.c_void_p:
    s = stream or torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def dptr(t: Optional[torch.Tensor]) -> Optional[int]:
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous()
    return t.data_ptr()


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_lib.c_int_p)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_lib.c_double_p)


class DevicePlan:
    """Network tables on the device (vk_plan)."""

    def __init__(self, net: Network, alpha: Optional[np.ndarray] = None):
        _require_device()
        lib = _lib.load()
        self.net = net
        self.tables: NetworkTables = build_tables(net)
        tb = self.tables
        keep = dict(
            reac=np.ascontiguousarray(net.reac, np.int32), prod=np.ascontiguousarray(net.prod, np.int32),
            kind=np.ascontiguousarray(net.kind, np.int32), par=np.ascontiguousarray(net.par, np.float64),
            rev=np.ascontiguousarray(net.rev, np.int32), nasa9=np.ascontiguousarray(net.nasa9, np.float64),
            atom_mat=np.ascontiguousarray(net.atom_matrix(), np.float64),
            ms=np.ascontiguousarray(net.mass, np.float64),
            alpha=np.ascontiguousarray(alpha if alpha is not None else thermal_alpha(net.species, net.mass),
                                       np.float64),
            chem_ptr=tb.chem_ptr.astype(np.int32), chem_pair=tb.chem_pair.astype(np.int32),
            chem_sign=tb.chem_sign.astype(np.int32), slot_pair=tb.slot_pair.astype(np.int32),
            slot_sp=tb.slot_sp.astype(np.int32), jac_map=tb.jac_map.astype(np.int32),
            jac_ptr=tb.jac_ptr.astype(np.int32), jac_slot=tb.jac_slot.astype(np.int32),
            jac_coef=tb.jac_coef.astype(np.float64))
        self.natom = keep['atom_mat'].shape[1]
        self.alpha = keep['alpha']
        d = _lib.NetworkDesc()
        d.ni, d.npairs, d.natom = net.ni, net.npairs, self.natom
        for k in ('reac', 'prod', 'kind', 'rev', 'chem_ptr', 'chem_pair', 'chem_sign', 'slot_pair', 'slot_sp',
                  'jac_map', 'jac_ptr', 'jac_slot'):
            setattr(d, k, _ip(keep[k]))
        for k in ('par', 'nasa9', 'atom_mat', 'ms', 'alpha', 'jac_coef'):
            setattr(d, k, _dp(keep[k]))
        d.nchem, d.nslot, d.njnz, d.njac = len(keep['chem_pair']), tb.nslot, tb.njnz, len(keep['jac_slot'])
        h = C.c_void_p()
        _lib.check(lib.vk_plan_create(C.byref(d), C.byref(h)), 'vk_plan_create')
        self.handle = h
        self._lib = lib

    def close(self):
        if self.handle:
            self._lib.vk_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceColumns:
    """Stacked column layout (vk_columns)."""

    def __init__(self, nz: Sequence[int]):
        _require_device()
        lib = _lib.load()
        nz = np.ascontiguousarray(nz, np.int32)
        h = C.c_void_p()
        _lib.check(lib.vk_columns_create(len(nz), _ip(nz), C.byref(h)), 'vk_columns_create')
        self.handle = h
        self.nz = nz
        self.off = np.concatenate([[0], np.cumsum(nz)]).astype(np.int64)
        self.L = int(self.off[-1])
        self.ncols = len(nz)
        self._lib = lib

    def close(self):
        if self.handle:
            self._lib.vk_columns_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def species_mask(names: Sequence[str], species: Sequence[str]) -> np.ndarray:
    """Bit mask of species names (bit i % 32 of word i // 32); an unknown
    name raises ValueError, as the reference's ``species.index(sp)`` does."""
    mask = np.zeros(5, np.uint32)
    for sp in names:
        i = list(species).index(sp)
        mask[i // 32] |= np.uint32(1 << (i % 32))
    return mask


def conv_ignore_mask(cfg: Config, species: Sequence[str]) -> np.ndarray:
    """``cfg.conver_ignore`` as a species mask (op.py:1045-1046)."""
    return species_mask(cfg.conver_ignore, species)


def escape_active(cfg: Config) -> bool:
    """Whether diffusion-limited escape (diff_esc, op.py:986-999) changes the
    step: through the top flux of diffdf (use_topflux, op.py:1589-1592) or
    the Jacobian of the _vm operators (op.py:2101-2107).  Otherwise the
    reference's update of top_flux has no effect and the device skips it."""
    return bool(cfg.diff_esc) and bool(cfg.use_topflux or (cfg.use_vm_mol and cfg.use_moldiff))


def params_from(cfg: Config, use_conv: bool = True, use_update_mu: bool = True,
                species: Optional[Sequence[str]] = None) -> _lib.Params:
    p = _lib.Params()
    for name in ('dt_min', 'dt_max', 'dt_var_min', 'dt_var_max', 'rtol', 'atol', 'mtol', 'mtol_conv', 'pos_cut',
                 'nega_cut', 'loss_eps', 'trun_min', 'runtime', 'st_factor', 'yconv_cri', 'slope_cri', 'yconv_min',
                 'flux_cri'):
        setattr(p, name, float(getattr(cfg, name)))
    p.count_min, p.count_max, p.conv_step, p.update_frq = cfg.count_min, cfg.count_max, cfg.conv_step, cfg.update_frq
    p.use_moldiff = int(cfg.use_moldiff)
    p.use_conv = int(use_conv)
    p.use_update_mu = int(use_update_mu)
    p.refine = int(getattr(cfg, 'refine', 1))
    if cfg.conver_ignore:
        if species is None:
            raise ValueError('conver_ignore needs the species list')
        for w, m in enumerate(conv_ignore_mask(cfg, species)):
            p.conv_ignore[w] = int(m)
    return p


class ColumnBatch:
    """Device state of columns sharing one network.

    Per-layer arrays are stacked over columns (column c owns layers
    ``off[c]:off[c+1]``); interface arrays use the first nz-1 slots of each
    column.  ``integrate`` runs the device-resident Ros2 loop; ``attempt``
    runs one plain ``Ros2.solver`` attempt.
    """

    def __init__(self, plan: DevicePlan, cols: List[Column], y0: Sequence[np.ndarray], cfg: Config,
                 history: int = 0, k: Optional[np.ndarray] = None, graph: bool = False, tune: bool = True):
        cfg.check_supported()
        self.plan, self.cfg = plan, cfg
        net = plan.net
        self.ni = net.ni
        self.layout = DeviceColumns([c.nz for c in cols])
        L, ni, nc = self.layout.L, net.ni, len(cols)
        off = self.layout.off
        dev = torch.device('cuda')

        def stack_layers(name):
            return np.concatenate([getattr(c, name) for c in cols])

        def stack_iface(name, width=None):
            out = np.zeros((L,) if width is None else (L, width))
            for c, o in zip(cols, off[:-1]):
                v = getattr(c, name)
                out[o:o + c.nz - 1] = v
            return out

        T = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=F64, device=dev)
        self.Tco, self.M = T(stack_layers('Tco')), T(stack_layers('M'))
        self.g, self.Hp = T(stack_layers('g')), T(stack_layers('Hp'))
        self.Kzz, self.dzi = T(stack_iface('Kzz')), T(stack_iface('dzi'))
        self.Ti, self.Hpi = T(stack_iface('Ti')), T(stack_iface('Hpi'))
        self.Dzz = T(stack_iface('Dzz', ni))
        self.tau = T(stack_layers('tau')) if all(c.tau is not None for c in cols) else None
        # transport terms and boundary conditions (store.py:131-141); an
        # array stays unallocated (NULL on the device) while its term is off
        self.vz = T(stack_iface('vz')) if any(np.any(c.vz) for c in cols) else None
        self.vm = None
        if cfg.use_vm_mol and cfg.use_moldiff:
            # build_atm.py:735-736 from the set-up geometry when the column has none
            self.vm = T(np.concatenate([c.vm if c.vm is not None else
                                        mol_velocity_H2(c.Tco, c.M, c.ms, c.alpha, c.g, c.Hp, c.dz)
                                        for c in cols]))

        def per_column(name):
            return T(np.stack([np.zeros(ni) if getattr(c, name) is None else getattr(c, name) for c in cols]))
        esc = escape_active(cfg)
        self.top_flux = per_column('top_flux') if (cfg.use_topflux or esc) else None
        self.bot_flux = per_column('bot_flux') if cfg.use_botflux else None
        self.bot_vdep = per_column('bot_vdep') if cfg.use_botflux else None
        self.fix_bot = None
        if cfg.use_fix_sp_bot:
            fb = np.full(ni, -1.0)
            for name, mix in cfg.use_fix_sp_bot.items():
                if not mix >= 0:
                    raise ValueError('use_fix_sp_bot[%r] = %r is not a mixing ratio' % (name, mix))
                fb[list(net.species).index(name)] = float(mix)
            self.fix_bot = T(np.tile(fb, (nc, 1)))
        self.dz = T(stack_layers('dz'))
        self.mu = T(stack_layers('mu'))
        self.zco = T(np.concatenate([c.zco for c in cols]))
        self.lnp = T(np.concatenate([[np.log(c.pico[i] / c.pico[i + 1]) for i in range(c.nz)] for c in cols]))
        self.gs = T([c.gs for c in cols])
        self.Rp = T([c.Rp for c in cols])
        self.pref = torch.as_tensor(np.array([c.pref_indx for c in cols], np.int32), device=dev)
        if cfg.use_moldiff and any(not np.array_equal(c.alpha, plan.alpha) for c in cols):
            raise ValueError("column alpha differs from the plan's thermal diffusion factors")
        if any(not np.array_equal(c.ms, net.mass) for c in cols):
            raise ValueError('column ms must equal the network molar masses')
        self.y = T(np.concatenate(y0))
        ys = self.y.sum(dim=1, keepdim=True)
        self.ymix = (self.y / ys).contiguous()
        self.y_prev = self.y.clone()
        self.k = torch.zeros((L, net.nr), dtype=F64, device=dev)
        self.natom = plan.natom
        self.atom_ini = torch.zeros((nc, self.natom), dtype=F64, device=dev)
        self.atom_loss = torch.zeros_like(self.atom_ini)
        self.atom_loss_prev = torch.zeros_like(self.atom_ini)
        st = np.zeros(nc, _lib.COLSTATE_DTYPE)
        st['dt'] = cfg.dttry
        st['longdy'] = st['longdydt'] = 1.0      # store.Variables' initial values (store.py:36-38)
        self.state = torch.as_tensor(st.view(np.uint8).copy(), device=dev)
        self.history = history
        lib = _lib.load()
        self._lib = lib
        h = C.c_void_p()
        _lib.check(lib.vk_integrator_create(plan.handle, self.layout.handle, history, C.byref(h)),
                   'vk_integrator_create')
        self.it = h
        self.set_graph(graph)
        self.atm = _lib.Atm(Tco=dptr(self.Tco), M=dptr(self.M), Kzz=dptr(self.Kzz), Dzz=dptr(self.Dzz),
                            dzi=dptr(self.dzi), Ti=dptr(self.Ti), Hpi=dptr(self.Hpi), g=dptr(self.g),
                            Hp=dptr(self.Hp), tau=dptr(self.tau), vz=dptr(self.vz), vm=dptr(self.vm),
                            top_flux=dptr(self.top_flux), bot_flux=dptr(self.bot_flux),
                            bot_vdep=dptr(self.bot_vdep), fix_bot=dptr(self.fix_bot),
                            use_topflux=int(cfg.use_topflux), use_botflux=int(cfg.use_botflux),
                            max_flux=float(cfg.max_flux))
        if esc:
            for w, m in enumerate(species_mask(cfg.diff_esc, net.species)):
                self.atm.esc[w] = int(m)
        self.geom = _lib.Geom(lnp=dptr(self.lnp), zco=dptr(self.zco), dz=dptr(self.dz), mu=dptr(self.mu),
                              gs=dptr(self.gs), Rp=dptr(self.Rp), pref=dptr(self.pref))
        if k is None:
            self.compute_rates()
        else:
            self.k.copy_(torch.as_tensor(k, dtype=F64))
        self.reset_element_baseline()
        self.factor_ms = self.tune() if tune else None

    # -- setup ------------------------------------------------------------------
    def compute_rates(self):
        _lib.check(self._lib.vk_rates(self.plan.handle, self.layout.handle, C.byref(self.atm), dptr(self.k),
                                      stream_handle()), 'vk_rates')

    def update_phi_esc(self):
        """update_phi_esc (op.py:986-999) now, from the current y."""
        _lib.check(self._lib.vk_update_phi_esc(self.plan.handle, self.layout.handle, C.byref(self.atm),
                                               dptr(self.y), stream_handle()), 'vk_update_phi_esc')

    def reset_element_baseline(self):
        """atom_ini from the current y (build_atm.py:283-291); losses to zero."""
        _lib.check(self._lib.vk_atom_sums(self.plan.handle, self.layout.handle, dptr(self.y), dptr(self.atom_ini),
                                          stream_handle()), 'vk_atom_sums')
        self.atom_loss.zero_()
        self.atom_loss_prev.zero_()

    # -- stepping ------------------------------------------------------------------
    def params(self, use_conv=None, use_update_mu=True):
        if use_conv is None:
            use_conv = self.history > 0
        return params_from(self.cfg, use_conv=use_conv, use_update_mu=use_update_mu, species=self.plan.net.species)

    def attempt(self):
        """One Ros2.solver attempt on every active column (op.py:2860-3007)."""
        prm = self.params(use_conv=False)
        _lib.check(self._lib.vk_ros2_solver(self.it, C.byref(self.atm), C.byref(prm), dptr(self.k), dptr(self.y),
                                            dptr(self.ymix), dptr(self.state), stream_handle()), 'vk_ros2_solver')

    def integrate(self, n_attempts: int, use_conv=None, use_update_mu=True):
        prm = self.params(use_conv=use_conv, use_update_mu=use_update_mu)
        _lib.check(self._lib.vk_integrate(self.it, C.byref(self.atm), C.byref(self.geom), C.byref(prm),
                                          dptr(self.k), dptr(self.y), dptr(self.ymix), dptr(self.y_prev),
                                          dptr(self.atom_ini), dptr(self.atom_loss), dptr(self.atom_loss_prev),
                                          dptr(self.state), int(n_attempts), stream_handle()), 'vk_integrate')

    def tune(self, reps: int = 2):
        """Time every block-inversion variant on this batch's shape and keep
        the fastest in the plan; returns milliseconds per variant (-1: not
        applicable)."""
        ms = (C.c_float * FACTOR_NVARIANTS)()
        _lib.check(self._lib.vk_tune_factor(self.plan.handle, self.layout.handle, int(reps), ms, stream_handle()),
                   'vk_tune_factor')
        return list(ms)

    def attach_photo(self, photo, ini_frq: int, final_frq: int):
        """Refresh the photolysis rates (and the actinic-flux change conv()
        tests) at the top of the loop every ini_frq steps, then every
        final_frq steps once the column is near steady state (op.py:818-830);
        photo is a vulcan_amd.photo.DevicePhoto of this batch's layout, or
        None."""
        self.photo = photo
        _lib.check(self._lib.vk_integrator_set_photo(self.it, photo.handle if photo is not None else None,
                                                     int(ini_frq), int(final_frq)), 'vk_integrator_set_photo')

    def update_photo(self, force: bool = True):
        """compute_tau + compute_flux + compute_J now (vulcan.py:172-174)."""
        self.photo.update(self.y, self.ymix, self.dz, self.k, state=self.state, force=force,
                          yconv_min=self.cfg.yconv_min)

    def set_graph(self, on: bool):
        """Replay a captured HIP graph per attempt in ``integrate``."""
        _lib.check(self._lib.vk_integrator_set_graph(self.it, int(on)), 'vk_integrator_set_graph')

    def set_timing(self, on: bool):
        _lib.check(self._lib.vk_integrator_set_timing(self.it, int(on)), 'vk_integrator_set_timing')

    def factor_stats(self):
        ms = C.c_double()
        n = C.c_longlong()
        _lib.check(self._lib.vk_integrator_stats(self.it, C.byref(ms), C.byref(n)), 'vk_integrator_stats')
        return ms.value, n.value

    def read_history(self, q0: int, count: int):
        """(t_time [ncols][count], y_time[q0:count] as [count-q0][L][ni]) from
        the integrator (the last ``history`` states are kept)."""
        t = np.zeros((self.layout.ncols, count))
        y = np.zeros((count - q0, self.layout.L, self.ni))
        _lib.check(self._lib.vk_integrator_history(self.it, 0, int(count), int(q0), t.ctypes.data, y.ctypes.data),
                   'vk_integrator_history')
        return t, y

    def write_history(self, t_time: np.ndarray, y_time: np.ndarray, q0: int):
        """Load t_time [ncols][count] and y_time[q0:count] ([count-q0][L][ni])
        into the integrator (resuming a run at count > 0)."""
        t = np.ascontiguousarray(t_time, dtype=np.float64)
        y = np.ascontiguousarray(y_time, dtype=np.float64)
        count = t.shape[1]
        assert t.shape[0] == self.layout.ncols and y.shape == (count - q0, self.layout.L, self.ni)
        _lib.check(self._lib.vk_integrator_history(self.it, 1, int(count), int(q0), t.ctypes.data, y.ctypes.data),
                   'vk_integrator_history')

    # -- host views ------------------------------------------------------------------
    def host_state(self) -> np.ndarray:
        return self.state.cpu().numpy().view(_lib.COLSTATE_DTYPE).copy()

    def set_host_state(self, st: np.ndarray):
        self.state.copy_(torch.as_tensor(np.ascontiguousarray(st).view(np.uint8)))

    def column(self, name: str, c: int) -> np.ndarray:
        t = getattr(self, name)
        o0, o1 = int(self.layout.off[c]), int(self.layout.off[c + 1])
        return t[o0:o1].cpu().numpy()

    def close(self):
        if getattr(self, 'it', None):
            self._lib.vk_integrator_destroy(self.it)
            self.it = None
        self.layout.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MixedBatch:
    """Columns over several networks (VULCAN runs one network per column;
    a batch may mix them).  One DevicePlan and ColumnBatch per network, each
    driven on its own default-priority stream so that the groups' attempts
    run concurrently; the block-inversion kernel is tuned per network (per
    ni).  `cases` carry `.net`, `.col`, `.y0`, `.cfg`."""

    def __init__(self, cases, history: int = 0, graph: bool = True, tune: bool = True):
        groups = {}
        for idx, c in enumerate(cases):
            groups.setdefault(id(c.net), []).append(idx)
        self.ncols = len(cases)
        self.groups = []
        for idx in groups.values():
            first = cases[idx[0]]
            plan = DevicePlan(first.net)
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                b = ColumnBatch(plan, [cases[i].col for i in idx], [cases[i].y0 for i in idx], first.cfg,
                                history=history, graph=graph, tune=tune)
            self.groups.append((np.asarray(idx), b, stream))
        torch.cuda.synchronize()

    def integrate(self, n_attempts: int, use_conv=None, use_update_mu=True):
        for _, b, stream in self.groups:
            with torch.cuda.stream(stream):
                b.integrate(n_attempts, use_conv=use_conv, use_update_mu=use_update_mu)

    def synchronize(self):
        torch.cuda.synchronize()

    def host_state(self) -> np.ndarray:
        torch.cuda.synchronize()
        out = np.zeros(self.ncols, _lib.COLSTATE_DTYPE)
        for idx, b, _ in self.groups:
            out[idx] = b.host_state()
        return out

    def group_of(self, c: int):
        """(batch, index within it) holding column c."""
        for idx, b, _ in self.groups:
            hit = np.nonzero(idx == c)[0]
            if hit.size:
                return b, int(hit[0])
        raise IndexError(c)

    def close(self):
        for _, b, _ in self.groups:
            b.close()
        self.groups = []
