"""``op.Ros2`` / ``op.Integration`` counterparts for VULCAN-style host code.

The reference selects its solver by name (``getattr(op, vulcan_cfg.ode_solver)()``,
vulcan.py:162-163) and drives it from ``Integration.__call__`` (op.py:808-936)
through ``naming_solver``, ``one_step``, ``step_size`` and ``solver``
(op.py:3078-3125).  This module offers the same methods with the same
argument meaning on the same ``var``/``atm``/``para`` attribute names
(store.py:21-208):

* ``Ros2.solver(var, atm, para)``: one attempt on the GPU; replaces ``var.y``
  and ``var.ymix`` and sets ``para.delta`` exactly as op.py:2860-3007;
* ``Ros2.one_step`` / ``clip`` / ``loss`` / ``step_ok`` / ``step_reject`` /
  ``reset_y`` / ``step_size``: the reference's host logic around the solver
  (op.py:2447-2534, 3091-3125), kept on the host for this plain path;
* ``Integration(solver)(var, atm, para, make_atm)``: the whole loop
  (backup, one_step, update_mu_dz, renormalisation, save_step, step_size,
  stop/conv) resident on the device, polled every ``poll`` attempts.

``compute_tau`` / ``compute_flux`` / ``compute_J`` run the photochemical
radiative transfer on the device from the tables VULCAN's set-up left in
``var``.  Vertical advection (``atm.vz``), the ``_vm`` operators
(``atm.vm``), top/bottom fluxes (``atm.top_flux``, ``bot_flux``,
``bot_vdep``), fixed bottom mixing ratios and diffusion-limited escape
follow the same ``vulcan_cfg`` switches as the reference.  Condensation,
settling, ions and photoionisation are not part of the accelerated path
and raise ``NotImplementedError``.
Numerical failure is never raised (negative y, delta and element loss are
reported through ``var``/``para``, as in the reference).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .atmosphere import Column
from .chem_funs import ChemFuns, k_dict_to_array
from .config import Config
from .network import Network


def _column_from_atm(atm, nz: int, ni: int) -> Column:
    g = lambda n, d=None: np.asarray(getattr(atm, n), dtype=np.float64) if hasattr(atm, n) else d
    return Column(pco=g('pco', np.zeros(nz)), pico=g('pico', np.ones(nz + 1)), Tco=g('Tco'), Kzz=g('Kzz'),
                  vz=g('vz', np.zeros(nz - 1)), M=g('M'), Dzz=g('Dzz', np.zeros((nz - 1, ni))),
                  alpha=g('alpha', np.zeros(ni)), ms=g('ms'), mu=g('mu', np.zeros(nz)), g=g('g'), Hp=g('Hp'),
                  dz=g('dz', np.zeros(nz)), zco=g('zco', np.zeros(nz + 1)), dzi=g('dzi'),
                  Ti=g('Ti', np.zeros(nz - 1)), Hpi=g('Hpi', np.zeros(nz - 1)), gs=float(getattr(atm, 'gs', 0.0)),
                  Rp=float(getattr(atm, 'Rp', 1.0)), pref_indx=int(getattr(atm, 'pref_indx', 0)),
                  tau=g('tau', None), vm=g('vm', None), top_flux=g('top_flux', None), bot_flux=g('bot_flux', None),
                  bot_vdep=g('bot_vdep', None))


def _sync_bc(b, col: Column, T, pad):
    """Copy the column's transport and boundary arrays into a batch built
    for it (the reference's update_phi_esc rewrites atm.top_flux between
    solver calls)."""
    if b.vz is not None:
        b.vz.copy_(T(pad(col.vz)))
    elif np.any(col.vz):
        raise ValueError('atm.vz became nonzero after the device batch was built without advection')
    if b.vm is not None and col.vm is not None:
        b.vm.copy_(T(col.vm))
    for name in ('top_flux', 'bot_flux', 'bot_vdep'):
        dst, src = getattr(b, name), getattr(col, name)
        if dst is not None and src is not None:
            dst.copy_(T(np.asarray(src, dtype=np.float64)[None]))


class Ros2:
    """op.Ros2 counterpart (op.py:2824-3125) for one network."""

    def __init__(self, net: Network, cfg: Optional[Config] = None):
        self.net = net
        self.cfg = cfg or Config()
        self.cfg.check_supported()
        for sp in self.cfg.conver_ignore:
            if sp not in net.species:
                # the reference fails the same way, at its first conv() test
                # (species.index(sp), op.py:1045-1046)
                raise ValueError('conver_ignore names %r, which is not a species of the network' % (sp,))
        from .device import escape_active
        if escape_active(self.cfg):
            for sp in self.cfg.diff_esc:
                if sp not in net.species:
                    # species.index(sp) in update_phi_esc (op.py:989-993)
                    raise ValueError('diff_esc names %r, which is not a species of the network' % (sp,))
        self.chem = ChemFuns(net)
        self.atom_list = net.atoms()
        self._atom_mat = net.atom_matrix()
        self._batch = None
        self._batch_nz = None

    # -- plugin surface --------------------------------------------------------
    def naming_solver(self, para):
        """op.py:3078-3088."""
        para.solver_str = 'solver'

    # -- photochemistry (op.py:2580-2787) on the device ----------------------------
    def _photo(self, var, atm):
        """The device radiative transfer of this solver's column, built from
        the tables VULCAN's set-up left in var (make_bins_read_cross,
        read_sflux) the first time."""
        from .photo import DevicePhoto, tables_from_var
        nz = np.shape(var.y)[0]
        b = self._device_batch(atm, nz)
        if getattr(self, '_photo_dev', None) is None or self._photo_nz != nz:
            self._photo_dev = DevicePhoto(b.plan, b.layout, tables_from_var(var, self.net.species, self.net.nr,
                                                                            self.cfg))
            self._photo_nz = nz
        return b, self._photo_dev

    def compute_tau(self, var, atm):
        """compute_tau + compute_flux + compute_J in one device update
        (op.py:2580-2787); var.tau and the fluxes are filled here,
        var.k[...] and var.J_sp by compute_J."""
        import torch
        from . import photo as P
        b, ph = self._photo(var, atm)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64))
        b.y.copy_(T(var.y))
        b.ymix.copy_(T(var.ymix))
        b.dz.copy_(T(atm.dz))
        b.k.copy_(T(k_dict_to_array(var.k, np.shape(var.y)[0], self.net.nr)))
        st = b.host_state()
        st['aflux_change'] = 0.0
        b.set_host_state(st)
        ph.update(b.y, b.ymix, b.dz, b.k, state=b.state, force=True)
        var.tau = ph.read(P.VK_PHOTO_TAU)
        var.sflux = ph.read(P.VK_PHOTO_SFLUX)
        var.dflux_u = ph.read(P.VK_PHOTO_DFLUX_U)
        var.dflux_d = ph.read(P.VK_PHOTO_DFLUX_D)
        var.prev_aflux = np.copy(getattr(var, 'aflux', np.zeros(1)))
        var.aflux = ph.read(P.VK_PHOTO_AFLUX)
        var.aflux_change = float(b.host_state()['aflux_change'][0])
        self._photo_k = b.k.cpu().numpy()

    def compute_flux(self, var, atm):
        """Done by compute_tau's device update (op.py:2603-2738)."""

    def compute_J(self, var, atm):
        """Photolysis rates of the last device update into var.k (op.py:2751-2787):
        k = J f_diurnal at the reference's pho_rate_index; var.J_sp[(sp, br)] = J
        for every branch br >= 1 and var.J_sp[(sp, 0)] their sum, in branch
        order.  Branches left out of the device tables (remove_list) keep
        J = 0 in J_sp."""
        t = self._photo_dev.tables
        col = {b: c for b, c in zip(t.branches, t.br_col)}
        for sp_br, idx in var.pho_rate_index.items():
            if sp_br in col:
                var.k[idx] = self._photo_k[:, col[sp_br]].copy()
        nz = np.shape(var.y)[0]
        n_branch = getattr(var, 'n_branch', None) or t.n_branch
        var.J_sp = {(sp, bn): np.zeros(nz) for sp in var.photo_sp for bn in range(n_branch[sp] + 1)}
        for sp in var.photo_sp:
            for bn in range(1, n_branch[sp] + 1):
                if (sp, bn) in col:
                    var.J_sp[(sp, bn)] = self._photo_k[:, col[(sp, bn)]] / self.cfg.f_diurnal
                var.J_sp[(sp, 0)] += var.J_sp[(sp, bn)]

    def compute_Jion(self, var, atm):
        raise NotImplementedError('photoionisation (use_ion) is outside the accelerated path')

    # -- the device attempt ------------------------------------------------------
    def _device_batch(self, atm, nz):
        import torch
        from .device import ColumnBatch
        col = _column_from_atm(atm, nz, self.net.ni)
        if self._batch is None or self._batch_nz != nz:
            from .device import DevicePlan
            plan = DevicePlan(self.net, alpha=col.alpha if self.cfg.use_moldiff else None)
            y0 = np.ones((nz, self.net.ni))
            self._batch = ColumnBatch(plan, [col], [y0], self.cfg, k=np.zeros((nz, self.net.nr)))
            self._batch_nz = nz
        b = self._batch
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64))
        pad = lambda v: np.append(np.asarray(v, dtype=np.float64), 0.0)
        for name in ('Tco', 'M', 'g', 'Hp'):
            getattr(b, name).copy_(T(getattr(col, name)))
        for name in ('Kzz', 'dzi', 'Ti', 'Hpi'):
            getattr(b, name).copy_(T(pad(getattr(col, name))))
        Dz = np.zeros((nz, self.net.ni))
        Dz[:nz - 1] = col.Dzz
        b.Dzz.copy_(T(Dz))
        _sync_bc(b, col, T, pad)
        return b

    def solver(self, var, atm, para):
        """One Ros2 attempt (op.py:2860-3007) on the GPU."""
        import torch
        from . import _lib
        y = np.ascontiguousarray(var.y, dtype=np.float64)
        nz = y.shape[0]
        b = self._device_batch(atm, nz)
        b.k.copy_(torch.as_tensor(k_dict_to_array(var.k, nz, self.net.nr)))
        b.y.copy_(torch.as_tensor(y))
        b.ymix.copy_(torch.as_tensor(np.ascontiguousarray(var.ymix, dtype=np.float64)))
        st = np.zeros(1, _lib.COLSTATE_DTYPE)
        st['dt'] = var.dt
        b.set_host_state(st)
        b.attempt()
        var.y = b.y.cpu().numpy()
        var.ymix = b.ymix.cpu().numpy()
        para.delta = float(b.host_state()['delta'][0])
        return var, para

    # -- host logic around the solver (op.py:2447-2534) ------------------------------
    def clip(self, var, para, atm, pos_cut=None, nega_cut=None):
        cfg = self.cfg
        pos_cut = cfg.pos_cut if pos_cut is None else pos_cut
        nega_cut = cfg.nega_cut if nega_cut is None else nega_cut
        y, ymix = var.y, var.ymix.copy()
        para.small_y += np.abs(np.sum(y[np.logical_and(y < pos_cut, y >= 0)]))
        para.nega_y += np.abs(np.sum(y[np.logical_and(y > nega_cut, y <= 0)]))
        y[np.logical_and(y < pos_cut, y >= nega_cut)] = 0.
        y[np.logical_and(ymix < cfg.mtol, y < 0)] = 0.
        var = self.loss(var)
        var.y, var.ymix = y, y / np.vstack(np.sum(y, axis=1))
        return var, para

    def loss(self, var):
        y = var.y
        for a, atom in enumerate(self.atom_list):
            var.atom_sum[atom] = np.sum([self._atom_mat[i, a] * y[:, i] for i in range(y.shape[1])])
            var.atom_loss[atom] = (var.atom_sum[atom] - var.atom_ini[atom]) / var.atom_ini[atom]
        return var

    def step_ok(self, var, para, loss_eps=None, rtol=None):
        loss_eps = self.cfg.loss_eps if loss_eps is None else loss_eps
        rtol = self.cfg.rtol if rtol is None else rtol
        cur = np.fromiter((var.atom_loss[a] for a in self.atom_list), float)
        prev = np.fromiter((var.atom_loss_prev[a] for a in self.atom_list), float)
        return bool(np.all(var.y >= 0) and np.amax(np.abs(cur - prev)) < loss_eps and para.delta <= rtol)

    def step_reject(self, var, para, loss_eps=None, rtol=None):
        rtol = self.cfg.rtol if rtol is None else rtol
        if para.delta > rtol:
            para.delta_count += 1
        elif np.any(var.y < 0):
            para.nega_count += 1
        else:
            para.loss_count += 1
        var = self.reset_y(var)
        if var.dt < self.cfg.dt_min:
            var.dt = self.cfg.dt_min
            var.y[var.y < 0] = 0.
            return True
        return False

    def reset_y(self, var, dt_reduc=None):
        var.y = var.y_prev
        var.dt *= self.cfg.dt_var_min if dt_reduc is None else dt_reduc
        return var

    def one_step(self, var, atm, para):
        """op.py:3091-3103."""
        while True:
            var, para = getattr(self, para.solver_str or 'solver')(var, atm, para)
            var, para = self.clip(var, para, atm)
            if self.step_ok(var, para):
                break
            elif self.step_reject(var, para):
                break
        return var, para

    def step_size(self, var, para, dt_var_min=None, dt_var_max=None, dt_min=None, dt_max=None):
        """op.py:3105-3125."""
        cfg = self.cfg
        h = var.dt
        delta = para.delta
        rtol = cfg.rtol
        if delta == 0:
            delta = 0.01 * rtol
        h_factor = 0.9 * (rtol / delta) ** 0.5
        h_factor = np.maximum(h_factor, cfg.dt_var_min if dt_var_min is None else dt_var_min)
        h_factor = np.minimum(h_factor, cfg.dt_var_max if dt_var_max is None else dt_var_max)
        h *= h_factor
        h = np.maximum(h, cfg.dt_min if dt_min is None else dt_min)
        h = np.minimum(h, cfg.dt_max if dt_max is None else dt_max)
        var.dt = float(h)
        return var


class Integration:
    """Integration.__call__ (op.py:808-936) resident on the device.

    ``var`` must carry ``y``, ``ymix``, ``k``, ``dt``, ``t``, ``atom_ini`` and
    ``atom_loss``; ``atm`` the column arrays.  A run resumed at
    ``para.count > 0`` needs ``var.t_time`` (``count`` entries) and
    ``var.y_time`` (at least the last ``conv_step + 1`` states), which are
    loaded into the device history that ``conv()`` reads.  On return
    ``var.y``, ``var.ymix``, ``var.t``, ``var.dt``, ``var.longdy``,
    ``var.longdydt``, ``var.atom_loss``, ``var.t_time`` (every step),
    ``var.y_time`` (every step with ``keep_y_time``, else the last
    ``conv_step + 1``), ``para.count``, ``para.end_case``, ``para.delta`` and
    the reject counters are set, and ``atm.g``/``Hp``/``dzi``/``Hpi`` hold
    the last ``update_mu_dz`` geometry.  The host polls the device every
    ``poll`` attempts (at most ``conv_step + 1``, the device ring's length,
    so that every state can be copied out before it is overwritten).
    """

    def __init__(self, odesolver: Ros2, output=None, poll: int = 64, keep_y_time: bool = True):
        self.odesolver = odesolver
        self.output = output
        self.poll = poll
        self.keep_y_time = keep_y_time

    def __call__(self, var, atm, para, make_atm=None):
        import torch
        from .device import ColumnBatch, DevicePlan
        solver = self.odesolver
        cfg = solver.cfg
        y = np.ascontiguousarray(var.y, dtype=np.float64)
        nz, ni = y.shape
        col = _column_from_atm(atm, nz, ni)
        plan = DevicePlan(solver.net, alpha=col.alpha if cfg.use_moldiff else None)
        H = cfg.conv_step + 1
        b = ColumnBatch(plan, [col], [y], cfg, history=H, k=k_dict_to_array(var.k, nz, solver.net.nr))
        b.ymix.copy_(torch.as_tensor(np.ascontiguousarray(var.ymix, dtype=np.float64)))
        ai = np.array([var.atom_ini[a] for a in solver.atom_list])
        b.atom_ini.copy_(torch.as_tensor(ai[None]))
        loss0 = getattr(var, 'atom_loss', None) or {}
        al = np.array([float(loss0.get(a, 0.0)) for a in solver.atom_list])
        b.atom_loss.copy_(torch.as_tensor(al[None]))
        b.atom_loss_prev.copy_(torch.as_tensor(al[None]))   # backup() at the loop's start
        count0 = int(getattr(para, 'count', 0) or 0)
        t_hist = [float(t) for t in getattr(var, 't_time', [])][:count0]
        y_hist = list(getattr(var, 'y_time', []))[:count0]
        if count0 > 0:
            if len(t_hist) != count0 or len(y_hist) < min(count0, H):
                raise ValueError('resuming at para.count = %d needs var.t_time (%d entries) and the last %d '
                                 'states of var.y_time' % (count0, count0, min(count0, H)))
            q0 = count0 - min(count0, H)
            b.write_history(np.array(t_hist)[None], np.stack(y_hist[len(y_hist) - (count0 - q0):]), q0)
        st = b.host_state()
        st['dt'], st['t'], st['count'] = var.dt, var.t, count0
        st['longdy'], st['longdydt'] = getattr(var, 'longdy', 1.0), getattr(var, 'longdydt', 1.0)
        st['aflux_change'] = getattr(var, 'aflux_change', 0.0)
        b.set_host_state(st)
        if cfg.use_photo:
            # the radiative transfer of the solver's own set-up call
            # (vulcan.py:172-174) carries on inside the loop; without one,
            # the tables come from var and the set-up call is made here
            ph = getattr(solver, '_photo_dev', None)
            if ph is None or solver._photo_nz != nz:
                from .photo import DevicePhoto, tables_from_var
                ph = DevicePhoto(plan, b.layout, tables_from_var(var, solver.net.species, solver.net.nr, cfg))
                b.photo = ph
                b.update_photo(force=True)
            b.attach_photo(ph, cfg.ini_update_photo_frq, cfg.final_update_photo_frq)
        poll = max(1, min(self.poll, H))
        seen = count0
        y_time = y_hist if self.keep_y_time else []
        while True:
            b.integrate(poll)
            h = b.host_state()[0]
            cnt = int(h['count'])
            if self.keep_y_time and cnt > seen:
                _, ys = b.read_history(seen, cnt)
                y_time.extend(ys[q] for q in range(cnt - seen))
            seen = cnt
            if h['done']:
                break
        t_all, ys = b.read_history(max(0, cnt - H) if not self.keep_y_time else cnt, cnt)
        var.t_time = [float(t) for t in t_all[0]]
        var.y_time = y_time if self.keep_y_time else [ys[q] for q in range(ys.shape[0])]
        var.y, var.ymix = b.y.cpu().numpy(), b.ymix.cpu().numpy()
        if cfg.use_photo:
            kk = b.k.cpu().numpy()
            for idx in var.pho_rate_index.values():
                if 1 <= idx <= solver.net.nr:
                    var.k[idx] = kk[:, idx - 1].copy()
            var.aflux_change = float(h['aflux_change'])
            b.attach_photo(None, 1, 1)
        var.t, var.dt = float(h['t']), float(h['dt'])
        var.longdy, var.longdydt = float(h['longdy']), float(h['longdydt'])
        para.count, para.end_case, para.delta = int(h['count']), int(h['end_case']), float(h['delta'])
        para.nega_count, para.loss_count, para.delta_count = int(h['nega_count']), int(h['loss_count']), \
            int(h['delta_count'])
        for name in ('g', 'Hp'):
            setattr(atm, name, b.column(name, 0))
        for name in ('dzi', 'Hpi'):
            setattr(atm, name, b.column(name, 0)[:nz - 1])
        if b.top_flux is not None:
            atm.top_flux = b.top_flux.cpu().numpy()[0].copy()     # update_phi_esc's result
        loss = b.atom_loss.cpu().numpy()[0]
        if getattr(var, 'atom_loss', None) is None:
            var.atom_loss = {}
        for a, atom in enumerate(solver.atom_list):
            var.atom_loss[atom] = float(loss[a])
        b.close()
        return var
