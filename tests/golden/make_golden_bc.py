"""Write golden vectors of the reference's boundary conditions and transport
variants.

Run in the development container only (the reference is not available
elsewhere):

    python tests/golden/make_golden_bc.py

Builds the same seeded 8-layer column of the generated network as
``make_golden.py`` (its ``setup()``; network tables and rates are those of
``ref_chemfuns_nz8.npz``) and records, with a signed vertical velocity
``vz``, constant top fluxes, bottom fluxes with deposition velocities, fixed
bottom mixing ratios and diffusion-limited escape switched on:

* the molecular-diffusion velocity ``vm`` of ``mol_diff`` with
  ``use_vm_mol`` (build_atm.py:731-740);
* ``diffdf`` / ``diffdf_no_mol`` / ``diffdf_vm`` (op.py:1438-1694) and
  ``lhs_jac_tot`` / ``lhs_jac_no_mol`` / ``lhs_jac_tot_vm`` as blocks
  (op.py:1973-2166);
* one ``Ros2.solver`` attempt with the flux and fixed-bottom conditions
  (op.py:2860-3007), for the default and the ``_vm`` operators;
* ``update_phi_esc`` (op.py:986-999), with one flux above ``max_flux``;
* two short ``Integration`` runs (op.py:808-936) with these conditions.

Only data (arrays, strings) is written, as compressed ``.npz``.
"""
import contextlib
import io
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_golden import NZ, OUT, setup  # noqa: E402


def to_blocks(a, ni):
    D = np.stack([a[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] for l in range(NZ)])
    L = np.zeros((NZ, ni))
    U = np.zeros((NZ, ni))
    rest = a.copy()
    for l in range(NZ):
        rest[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] = 0
        rows = np.arange(l * ni, (l + 1) * ni)
        if l > 0:
            L[l] = a[rows, rows - ni]
            rest[rows, rows - ni] = 0
        if l < NZ - 1:
            U[l] = a[rows, rows + ni]
            rest[rows, rows + ni] = 0
    assert not rest.any(), 'lhs has entries outside the block-tridiagonal pattern'
    return D, L, U


def main():
    ctx = setup()
    cfg, op, var, atm, para = ctx['cfg'], ctx['op'], ctx['var'], ctx['atm'], ctx['para']
    rng, species, ni, make_atm, ini = ctx['rng'], ctx['species'], ctx['ni'], ctx['make_atm'], ctx['ini']
    geom0 = ctx['geom0']
    y0 = var.y.copy()
    gold = dict(y0=y0, k=ctx['karr'])

    # ---- the molecular-diffusion velocity (use_vm_mol) -------------------------
    cfg.use_vm_mol = True
    make_atm.mol_diff(atm)
    cfg.use_vm_mol = False
    gold['vm'] = atm.vm.copy()
    gold['Dzz_cen'] = atm.Dzz_cen.copy()
    gold['Dzz'] = atm.Dzz.copy()

    # ---- boundary-condition inputs ---------------------------------------------
    vz = rng.uniform(-1., 1., NZ - 1) * 10.0 ** rng.uniform(0., 2., NZ - 1)
    vz[2] = 0.0
    atm.vz = vz.copy()
    top_sp = ['H', 'H2O', 'CO', 'CH4', 'O2']
    bot_sp = ['CO', 'CH4', 'H2O', 'CO2', 'C2H2', 'OH']
    for sp in top_sp:
        atm.top_flux[species.index(sp)] = rng.choice([-1., 1.]) * 10.0 ** rng.uniform(6., 10.)
    for sp in bot_sp:
        atm.bot_flux[species.index(sp)] = 10.0 ** rng.uniform(8., 12.)
        atm.bot_vdep[species.index(sp)] = rng.choice([0., 10.0 ** rng.uniform(-3., 0.)])
    gold['vz'] = vz
    gold['top_flux_in'] = atm.top_flux.copy()
    gold['bot_flux'] = atm.bot_flux.copy()
    gold['bot_vdep'] = atm.bot_vdep.copy()
    fix = {'CO2': 3.0e-4, 'CH4': 2.0e-6}
    gold['fix_sp'] = np.array(list(fix))
    gold['fix_mix'] = np.array(list(fix.values()))
    cfg.diff_esc = ['H', 'H2']
    gold['diff_esc'] = np.array(cfg.diff_esc)

    # ---- operators ---------------------------------------------------------------
    cfg.use_topflux = True
    cfg.use_botflux = True
    solver = op.Ros2()
    gold['diffdf_mol'] = solver.diffdf(y0, atm)
    gold['diffdf_nomol'] = solver.diffdf_no_mol(y0, atm)
    gold['diffdf_vm'] = solver.diffdf_vm(y0, atm)
    var.y = y0.copy()
    var.dt = 3.0e-9
    gold['lhs_dt'] = var.dt
    for name, fn in (('tot', solver.lhs_jac_tot), ('nomol', solver.lhs_jac_no_mol), ('vm', solver.lhs_jac_tot_vm)):
        D, L, U = to_blocks(fn(var, atm), ni)
        gold['lhs_%s_D' % name], gold['lhs_%s_L' % name], gold['lhs_%s_U' % name] = D, L, U

    # ---- one solver attempt per operator set -------------------------------------
    cfg.use_fix_sp_bot = dict(fix)
    for name, vm_mol in (('tot', False), ('vm', True)):
        cfg.use_vm_mol = vm_mol
        solver = op.Ros2()                          # reads use_fix_sp_bot (op.py:1435-1436)
        var.dt = 1e-10
        var.y = y0.copy()
        var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
        var, para = solver.solver(var, atm, para)
        gold['solver_%s_sol' % name] = var.y.copy()
        gold['solver_%s_ymix' % name] = var.ymix.copy()
        gold['solver_%s_delta' % name] = para.delta
    gold['solver_dt'] = 1e-10
    cfg.use_vm_mol = False

    # ---- update_phi_esc: H2's diffusion-limited flux exceeds max_flux ------------
    integ = op.Integration(solver, op.Output())
    var.y = y0 * (1.0 + 0.2 * rng.random(y0.shape))
    cfg.max_flux = 1e10
    integ.update_phi_esc(var, atm)
    gold['phi_y'] = var.y.copy()
    gold['phi_max_flux'] = cfg.max_flux
    gold['phi_top_flux'] = atm.top_flux.copy()
    for sp in cfg.diff_esc:
        i = species.index(sp)
        raw = -atm.Dzz[-1, i] * var.y[-1, i] * (1. / atm.Hp[-1] - atm.ms[i] * atm.g[-1] / (op.Navo * op.kb * atm.Tco[-1]))
        gold['phi_raw_' + sp] = raw
    assert gold['phi_raw_H2'] < -cfg.max_flux < gold['phi_raw_H'], 'max_flux does not split the two species'

    # ---- two short integrations -----------------------------------------------------
    # a: default operators with vz, both flux conditions, fixed bottom mixing
    #    ratios and escape (through use_topflux);
    # b: the _vm operators with vz and escape but no flux conditions: escape then
    #    enters only the Jacobian (op.py:2101-2107)
    runs = (('a', dict(use_vm_mol=False, use_topflux=True, use_botflux=True, use_fix_sp_bot=dict(fix))),
            ('b', dict(use_vm_mol=True, use_topflux=False, use_botflux=False, use_fix_sp_bot={})))
    cfg.update_frq = 10
    cfg.count_min = 20
    cfg.trun_min = 1e-8
    cfg.conv_step = 10
    cfg.count_max = 40
    gold['traj_cfg'] = np.array([cfg.update_frq, cfg.count_min, cfg.trun_min, cfg.conv_step, cfg.count_max])
    for tag, sw in runs:
        for key, v in sw.items():
            setattr(cfg, key, v)
        for key, v in geom0.items():
            setattr(atm, key[:-1], v.copy())
        atm.top_flux = gold['top_flux_in'].copy()
        var.y = y0.copy()
        var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
        var = ini.ele_sum(var)
        var.dt = cfg.dttry
        var.t = 0
        var.longdy, var.longdydt = 1., 1.
        var.y_time, var.t_time, var.atom_loss_time = [], [], []
        var.atom_loss = {a: 0. for a in var.atom_loss}
        para.count = 0
        para.delta_count = para.nega_count = para.loss_count = 0
        para.end_case = 0
        solver = op.Ros2()
        integ = op.Integration(solver, op.Output())
        with contextlib.redirect_stdout(io.StringIO()):
            integ(var, atm, para, make_atm)
        gold['traj%s_y' % tag] = np.stack(var.y_time)
        gold['traj%s_t' % tag] = np.array(var.t_time)
        gold['traj%s_final' % tag] = np.array([var.dt, para.delta, var.longdy, var.longdydt, para.end_case,
                                               para.count, para.delta_count, para.nega_count, para.loss_count])
        gold['traj%s_top_flux' % tag] = atm.top_flux.copy()
        gold['traj%s_switches' % tag] = np.array([int(sw['use_vm_mol']), int(sw['use_topflux']),
                                                  int(sw['use_botflux']), int(bool(sw['use_fix_sp_bot']))])
    out = os.path.join(OUT, 'ref_bc_nz8.npz')
    np.savez_compressed(out, **gold)
    print('wrote %s: traj a %d steps (end %d), traj b %d steps (end %d)'
          % (out, gold['traja_final'][5], gold['traja_final'][4], gold['trajb_final'][5], gold['trajb_final'][4]))


if __name__ == '__main__':
    main()
