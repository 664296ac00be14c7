"""Boundary conditions and transport variants: vertical advection (vz),
constant top flux, bottom flux with deposition velocity, fixed bottom mixing
ratios, diffusion-limited escape and the upwind molecular-diffusion velocity
(use_vm_mol).

CPU pins of the oracle:

* golden vectors the reference wrote (tests/golden/make_golden_bc.py) on the
  seeded 8-layer column of ref_chemfuns_nz8.npz: vm, the three diffdf and
  lhs variants, a solver attempt per operator set, update_phi_esc and two
  41-step Integration runs;
* invariants: with closed boundaries the flux-form operators telescope
  (column integral zero); with flux boundaries the column integral equals
  top_flux + bot_flux - vdep*y[0]; the lhs diffusion part is the derivative
  of diffdf at frozen layer totals.

The device parity tests for the same vectors are marked gpu.
"""
import ctypes as C
import os

import numpy as np
import pytest

from _golden_util import load, network_from, column_from, rel_err
from oracle import ros2_oracle as O
from vulcan_amd import network as nw
from vulcan_amd.atmosphere import mol_velocity_H2, read_bc_flux
from vulcan_amd.config import Config

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope='module')
def bd():
    d0 = load()
    d = load('ref_bc_nz8.npz')
    net = network_from(d0)
    assert np.array_equal(d['y0'], d0['y0']) and np.array_equal(d['k'], d0['k'])
    return d0, d, net, nw.build_tables(net)


def bc_column(d0, d):
    col = column_from(d0)
    assert np.array_equal(col.Dzz, d['Dzz'])
    col.vz = d['vz'].copy()
    col.vm = d['vm'].copy()
    col.top_flux = d['top_flux_in'].copy()
    col.bot_flux = d['bot_flux'].copy()
    col.bot_vdep = d['bot_vdep'].copy()
    return col


def bc_config(d, vm=False, top=True, bot=True, fix=True, **kw):
    return Config(use_vm_mol=vm, use_topflux=top, use_botflux=bot,
                  use_fix_sp_bot=dict(zip([str(s) for s in d['fix_sp']], d['fix_mix'])) if fix else {},
                  diff_esc=[str(s) for s in d['diff_esc']], **kw)


# --------------------------------------------------------------------------
# golden vectors
# --------------------------------------------------------------------------

def test_vm_profile_matches_reference(bd):
    d0, d, net, tb = bd
    col = column_from(d0)
    vm = mol_velocity_H2(col.Tco, col.M, col.ms, col.alpha, col.g, col.Hp, col.dz)
    assert np.array_equal(vm, d['vm'])


@pytest.mark.parametrize('mol,vm,key', [(True, False, 'diffdf_mol'), (False, False, 'diffdf_nomol'),
                                        (True, True, 'diffdf_vm')])
def test_diffdf_with_boundary_terms_bitwise(bd, mol, vm, key):
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    out = O.diffdf(col, d['y0'], mol, bc_config(d, vm=vm))
    assert np.array_equal(out, d[key])


@pytest.mark.parametrize('mol,vm,key', [(True, False, 'tot'), (False, False, 'nomol'), (True, True, 'vm')])
def test_lhs_with_boundary_terms_bitwise(bd, mol, vm, key):
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, vm=vm)
    esc = O.esc_indices(net.species, cfg)
    # with the reference's own Jacobian blocks the blocks are bitwise equal
    D, L, U = O.lhs_blocks(col, d['y0'], float(d['lhs_dt']), -d0['negjac'], mol, cfg, esc)
    assert np.array_equal(D, d['lhs_%s_D' % key])
    assert np.array_equal(L, d['lhs_%s_L' % key])
    assert np.array_equal(U, d['lhs_%s_U' % key])


@pytest.mark.parametrize('vm', [False, True])
def test_solver_attempt_with_boundary_conditions(bd, vm, monkeypatch):
    """Fixed bottom mixing ratios override sol[0] and, with the bottom flux,
    drop layer 0 from delta (op.py:2945-2953); bitwise with the reference's
    Jacobian."""
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, vm=vm)
    monkeypatch.setattr(O, 'jac_blocks', lambda *a, **k: -d0['negjac'])
    st = O.new_state(net, d['y0'].copy(), d['k'], cfg)
    st.dt = float(d['solver_dt'])
    O.ros2_solver(net, col, st, cfg, tb)
    name = 'vm' if vm else 'tot'
    assert np.array_equal(st.y, d['solver_%s_sol' % name])
    assert np.array_equal(st.ymix, d['solver_%s_ymix' % name])
    assert st.delta == float(d['solver_%s_delta' % name])
    for sp, mix in zip(d['fix_sp'], d['fix_mix']):
        assert st.y[0, net.species.index(str(sp))] == mix * col.M[0]


def test_update_phi_esc_bitwise(bd):
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, max_flux=float(d['phi_max_flux']))
    st = O.new_state(net, d['phi_y'].copy(), d['k'], cfg)
    O.update_phi_esc(col, st, cfg, net.species)
    assert np.array_equal(col.top_flux, d['phi_top_flux'])
    # H2's diffusion-limited flux is capped at -max_flux, H's is not
    assert col.top_flux[net.species.index('H2')] == -float(d['phi_max_flux'])
    assert col.top_flux[net.species.index('H')] == float(d['phi_raw_H'])


@pytest.mark.parametrize('tag', ['a', 'b'])
def test_trajectory_with_boundary_conditions(bd, tag):
    """41 accepted steps with rejections, update_mu_dz and update_phi_esc
    every 10 steps: a) default operators with vz, both flux conditions, fixed
    bottom mixing ratios and escape; b) the _vm operators with escape in the
    Jacobian only."""
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    for key in ('g', 'Hp', 'dz', 'zco', 'dzi', 'Ti', 'Hpi'):
        setattr(col, key, d0['trajg0_' + key].copy())
    vm, top, bot, fix = (bool(x) for x in d['traj%s_switches' % tag])
    upd, cmin, tmin, cstep, cmax = d['traj_cfg']
    cfg = bc_config(d, vm=vm, top=top, bot=bot, fix=fix, max_flux=float(d['phi_max_flux']),
                    update_frq=int(upd), count_min=int(cmin), trun_min=float(tmin), conv_step=int(cstep),
                    count_max=int(cmax))
    st = O.new_state(net, d['y0'].copy(), d['k'], cfg)
    O.integrate(net, col, st, cfg, tables=tb)
    fin = d['traj%s_final' % tag]
    assert st.count == int(fin[5]) and st.end_case == int(fin[4])
    assert [st.delta_count, st.nega_count, st.loss_count] == [int(x) for x in fin[6:9]]
    assert rel_err(np.asarray(st.t_time), d['traj%s_t' % tag]) < 1e-9
    ys = np.stack(st.y_time)
    assert rel_err(ys, d['traj%s_y' % tag], floor=1e-6 * d0['M'].min()) < 1e-6
    assert rel_err(col.top_flux, d['traj%s_top_flux' % tag], floor=1e-3) < 1e-6


# --------------------------------------------------------------------------
# invariants
# --------------------------------------------------------------------------

def _weights(col):
    """Cell widths of the flux form: dzi[0], dz_ave[1..nz-2], dzi[nz-2]."""
    dzi = col.dzi
    w = np.empty(col.nz)
    w[0] = dzi[0]
    w[1:-1] = 0.5 * (dzi[:-1] + dzi[1:])
    w[-1] = dzi[-1]
    return w


@pytest.mark.parametrize('mol,vm', [(False, False), (True, True)])
def test_closed_column_operators_telescope(bd, mol, vm):
    """Eddy diffusion, upwind advection and (with vm) molecular diffusion are
    fluxes between layers: sum_j w_j diffdf_j = 0 for every species."""
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, vm=vm, top=False, bot=False, fix=False)
    w = _weights(col)[:, None]
    if vm:
        # the top layer's flux uses the layer row vm[nz-1] where the layer
        # below uses vm[nz-2] for the same interface (op.py:1649-1653 against
        # 1666-1670): the column integral telescopes only when they agree
        f = O.diffdf(col, d['y0'], mol, cfg)
        assert np.all(np.abs(np.sum(w * f, axis=0)) > 1e-12 * np.sum(np.abs(w * f), axis=0))
        col.vm[-1] = col.vm[-2]
    f = O.diffdf(col, d['y0'], mol, cfg)
    tot = np.sum(w * f, axis=0)
    scale = np.sum(np.abs(w * f), axis=0)
    assert np.all(np.abs(tot) <= 1e-12 * scale + 1e-300)
    # and the advection alone moves mass out of the cell it leaves (upwind):
    # a positive vz at interface 0 drains layer 0 only
    col2 = bc_column(d0, d)
    col2.Kzz = np.zeros_like(col2.Kzz)
    col2.vz = np.zeros_like(col2.vz)
    col2.vz[0] = 3.0
    f2 = O.diffdf(col2, d['y0'], False, cfg)
    assert np.allclose(f2[0], -3.0 * d['y0'][0] / col2.dzi[0], rtol=1e-15, atol=0)
    assert np.allclose(f2[1], 3.0 * d['y0'][0] / (0.5 * (col2.dzi[0] + col2.dzi[1])), rtol=1e-15, atol=0)
    assert not f2[2:].any()


def test_boundary_fluxes_enter_the_column_integral(bd):
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    col.vm[-1] = col.vm[-2]
    cfg = bc_config(d, vm=True)
    y = d['y0']
    f = O.diffdf(col, y, True, cfg)
    w = _weights(col)[:, None]
    expect = col.top_flux + col.bot_flux - y[0] * col.bot_vdep
    tot = np.sum(w * f, axis=0)
    scale = np.sum(np.abs(w * f), axis=0) + np.abs(expect)
    assert np.all(np.abs(tot - expect) <= 1e-12 * scale)


@pytest.mark.parametrize('mol,vm', [(True, False), (False, False), (True, True)])
def test_lhs_is_the_frozen_total_derivative_of_diffdf(bd, mol, vm):
    """c0 I - lhs (J = 0) applied to y equals diffdf(y) minus its constant
    flux terms; the deposition velocity is part of both."""
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, vm=vm)
    y = d['y0']
    nz, ni = y.shape
    dt = 1e30                   # c0 = 1/(r dt) negligible: no cancellation against it
    D, L, U = O.lhs_blocks(col, y, dt, np.zeros((nz, ni, ni)), mol, cfg, None)
    A = (1. / (O.R_ROS * dt)) * np.eye(nz * ni) - O.lhs_dense(D, L, U)
    f = O.diffdf(col, y, mol, cfg).copy()
    f[-1] -= col.top_flux / col.dzi[-1]
    f[0] -= col.bot_flux / col.dzi[0]
    Ay = A @ y.ravel()
    assert np.allclose(Ay, f.ravel(), rtol=1e-10, atol=1e-10 * np.abs(f).max())


def test_escape_jacobian_term_is_flux_over_density(bd):
    """The _vm lhs subtracts top_flux/y on the top diagonal for diff_esc
    species with y > 0 (op.py:2101-2107; no 1/dzi factor, as there)."""
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    cfg = bc_config(d, vm=True)
    y = d['y0'].copy()
    esc = O.esc_indices(net.species, cfg)
    Da, _, _ = O.lhs_blocks(col, y, 1e-3, np.zeros((8, net.ni, net.ni)), True, cfg, esc)
    Db, _, _ = O.lhs_blocks(col, y, 1e-3, np.zeros((8, net.ni, net.ni)), True, cfg, None)
    diff = np.diag(Da[-1]) - np.diag(Db[-1])
    expect = np.zeros(net.ni)
    for i in esc:
        expect[i] = -col.top_flux[i] / y[-1, i]
    assert np.allclose(diff, expect, rtol=1e-12, atol=0)
    assert np.array_equal(Da[:-1], Db[:-1])


def test_read_bc_flux(tmp_path):
    species = ['H', 'H2O', 'CO', 'CH4']
    (tmp_path / 'top.txt').write_text('#species flux\nH   1.5E9\n\nH2O  -4E4\n')
    (tmp_path / 'bot.txt').write_text('#species flux v_dep\nCO  3.7e11  0.03\nCH4 1.6e+11 0.  2e-6\n')
    top, bot, vdep, fix = read_bc_flux(species, str(tmp_path / 'top.txt'), str(tmp_path / 'bot.txt'))
    assert np.array_equal(top, [1.5e9, -4e4, 0, 0])
    assert np.array_equal(bot, [0, 0, 3.7e11, 1.6e11])
    assert np.array_equal(vdep, [0, 0, 0.03, 0])
    assert np.array_equal(fix, [0, 0, 0, 2e-6])


# --------------------------------------------------------------------------
# the device against the same vectors (through the C ABI)
# --------------------------------------------------------------------------

gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


def _bc_batch(d0, d, cfg, col=None, history=0):
    from vulcan_amd.device import ColumnBatch, DevicePlan
    net = network_from(d0)
    col = col if col is not None else bc_column(d0, d)
    plan = DevicePlan(net, alpha=d0['alpha'])
    return net, col, ColumnBatch(plan, [col], [d['y0']], cfg, history=history, k=d['k'])


def _step_rel(a, ref):
    ys = ref.sum(axis=1)[:, None]
    return float(np.max(np.abs(a - ref) / np.maximum(np.abs(ref), 1e-10 * ys)))


@gpu
@pytest.mark.parametrize('vm', [False, True])
def test_device_diffdf_and_lhs_with_boundary_terms_bitwise(bd, vm):
    torch = _torch()
    from vulcan_amd import _lib
    from vulcan_amd.device import dptr, stream_handle
    d0, d, net, tb = bd
    cfg = bc_config(d, vm=vm)
    net, col, b = _bc_batch(d0, d, cfg)
    lib = _lib.load()
    out = torch.empty_like(b.y)
    cases = ((1, 'diffdf_vm'),) if vm else ((1, 'diffdf_mol'), (0, 'diffdf_nomol'))
    for mol, key in cases:
        _lib.check(lib.vk_diffdf(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), mol, dptr(out),
                                 stream_handle()), 'vk_diffdf')
        assert np.array_equal(out.cpu().numpy(), d[key]), key
    y = d['y0']
    nz, ni = y.shape
    J = torch.empty((nz, ni, ni), dtype=torch.float64, device='cuda')
    _lib.check(lib.vk_jac(b.plan.handle, b.layout.handle, dptr(b.y), dptr(b.M), dptr(b.k), dptr(J),
                          stream_handle()), 'vk_jac')
    Jo = O.jac_blocks(net, y, col.M, d['k'])
    assert np.array_equal(J.cpu().numpy(), Jo)
    dt = float(d['lhs_dt'])
    dtc = torch.full((1,), dt, dtype=torch.float64, device='cuda')
    Ls = torch.empty((nz, ni), dtype=torch.float64, device='cuda')
    Us = torch.empty_like(Ls)
    esc = O.esc_indices(net.species, cfg)
    for mol, key in (((1, 'vm'),) if vm else ((1, 'tot'), (0, 'nomol'))):
        _lib.check(lib.vk_lhs(b.plan.handle, b.layout.handle, C.byref(b.atm), dptr(b.y), dptr(b.k), mol,
                              dptr(dtc), dptr(J), dptr(Ls), dptr(Us), stream_handle()), 'vk_lhs')
        D, L, U = O.lhs_blocks(col, y, dt, Jo, bool(mol), cfg, esc)
        assert np.array_equal(J.cpu().numpy(), D), key
        assert np.array_equal(Ls.cpu().numpy(), d['lhs_%s_L' % key]), key
        assert np.array_equal(Us.cpu().numpy(), d['lhs_%s_U' % key]), key
        _lib.check(lib.vk_jac(b.plan.handle, b.layout.handle, dptr(b.y), dptr(b.M), dptr(b.k), dptr(J),
                              stream_handle()), 'vk_jac')


@gpu
@pytest.mark.parametrize('vm', [False, True])
def test_device_attempt_with_boundary_conditions(bd, vm):
    _torch()
    d0, d, net, tb = bd
    cfg = bc_config(d, vm=vm)
    net, col, b = _bc_batch(d0, d, cfg)
    st = b.host_state()
    st['dt'] = float(d['solver_dt'])
    b.set_host_state(st)
    b.attempt()
    name = 'vm' if vm else 'tot'
    ref = d['solver_%s_sol' % name]
    sol = b.y.cpu().numpy()
    rel = _step_rel(sol, ref)
    print('attempt with boundary conditions (%s): rel %.3e' % (name, rel))
    assert rel < 1e-10
    dl = float(b.host_state()['delta'][0])
    assert abs(dl - float(d['solver_%s_delta' % name])) <= 1e-9 * float(d['solver_%s_delta' % name])
    for sp, mix in zip(d['fix_sp'], d['fix_mix']):
        i = net.species.index(str(sp))
        assert sol[0, i] == mix * col.M[0] == ref[0, i]


@gpu
def test_device_update_phi_esc_bitwise(bd):
    torch = _torch()
    d0, d, net, tb = bd
    cfg = bc_config(d, max_flux=float(d['phi_max_flux']))
    net, col, b = _bc_batch(d0, d, cfg)
    b.y.copy_(torch.as_tensor(d['phi_y']))
    b.update_phi_esc()
    assert np.array_equal(b.top_flux.cpu().numpy()[0], d['phi_top_flux'])


@gpu
@pytest.mark.parametrize('tag', ['a', 'b'])
def test_device_trajectory_with_boundary_conditions(bd, tag):
    """The reference's 41-step runs with boundary conditions, on the device
    (graph replay): counts, end state and the escape fluxes."""
    _torch()
    d0, d, net, tb = bd
    col = bc_column(d0, d)
    for key in ('g', 'Hp', 'dz', 'zco', 'dzi', 'Ti', 'Hpi'):
        setattr(col, key, d0['trajg0_' + key].copy())
    vm, top, bot, fix = (bool(x) for x in d['traj%s_switches' % tag])
    upd, cmin, tmin, cstep, cmax = d['traj_cfg']
    cfg = bc_config(d, vm=vm, top=top, bot=bot, fix=fix, max_flux=float(d['phi_max_flux']),
                    update_frq=int(upd), count_min=int(cmin), trun_min=float(tmin), conv_step=int(cstep),
                    count_max=int(cmax))
    net, col, b = _bc_batch(d0, d, cfg, col=col, history=cfg.conv_step + 1)
    b.set_graph(True)
    for _ in range(400):
        b.integrate(8)
        if int(b.host_state()['done'][0]):
            break
    h = b.host_state()[0]
    fin = d['traj%s_final' % tag]
    assert int(h['end_case']) == int(fin[4]) and int(h['count']) == int(fin[5])
    assert [int(h['delta_count']), int(h['nega_count']), int(h['loss_count'])] == [int(x) for x in fin[6:9]]
    yref = d['traj%s_y' % tag][-1]
    ys = yref.sum(axis=1)
    yd = b.y.cpu().numpy()
    rel = _step_rel(yd, yref)
    absn = float(np.max(np.abs(yd - yref) / ys[:, None]))
    print('41-step trajectory %s: rel %.3e abs/total %.3e' % (tag, rel, absn))
    assert rel < 1e-7 and absn < 1e-12
    assert abs(h['longdy'] - fin[2]) <= 1e-6 * abs(fin[2])
    if b.top_flux is not None:
        assert rel_err(b.top_flux.cpu().numpy()[0], d['traj%s_top_flux' % tag], floor=1e-3) < 1e-6


@gpu
def test_integration_binding_with_boundary_conditions(bd):
    """vulcan_amd.op.Integration on VULCAN-shaped objects whose atm carries
    vz, top_flux, bot_flux and bot_vdep (store.py:131-141): run a, with
    atm.top_flux written back as update_phi_esc leaves it."""
    _torch()
    from test_boundary import vulcan_objects, _traj_geometry
    from vulcan_amd.op import Integration, Ros2
    d0, d, net, tb = bd
    net, var, atm, para = vulcan_objects(d0)
    _traj_geometry(d0, atm)
    atm.vz = d['vz'].copy()
    atm.top_flux, atm.bot_flux, atm.bot_vdep = d['top_flux_in'].copy(), d['bot_flux'].copy(), d['bot_vdep'].copy()
    upd, cmin, tmin, cstep, cmax = d['traj_cfg']
    cfg = bc_config(d, max_flux=float(d['phi_max_flux']), update_frq=int(upd), count_min=int(cmin),
                    trun_min=float(tmin), conv_step=int(cstep), count_max=int(cmax))
    Integration(Ros2(net, cfg), poll=8)(var, atm, para)
    fin = d['traja_final']
    assert para.end_case == int(fin[4]) and para.count == int(fin[5])
    assert [para.delta_count, para.nega_count, para.loss_count] == [int(x) for x in fin[6:9]]
    assert rel_err(np.asarray(var.t_time), d['traja_t']) < 1e-9
    assert rel_err(np.stack(var.y_time), d['traja_y'], floor=1e-6 * d0['M'].min()) < 1e-6
    assert rel_err(atm.top_flux, d['traja_top_flux'], floor=1e-3) < 1e-6
