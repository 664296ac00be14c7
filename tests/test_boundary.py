"""The drop-in surface on the device: ``vulcan_amd.op.Ros2`` /
``vulcan_amd.op.Integration`` / ``vulcan_amd.chem_funs.ChemFuns`` driven on
VULCAN-shaped ``var`` / ``atm`` / ``para`` objects (store.py:21-208), and the
steady-state decision (``vk_conv``, ``k_conv`` inside ``vk_integrate``)
against what the reference itself decided.

Golden vectors: ``ref_chemfuns_nz8.npz`` (chemdf, neg_symjac, Gibbs, one
solver attempt, one_step with its rejections, a 61-step trajectory) and
``ref_conv_nz8.npz`` (two runs to steady state and crafted conv() histories),
both written by ``tests/golden/make_golden.py`` from the reference.

Steady-state parity.  The step-size controller makes long runs sensitive to
rounding: delta (op.py:2948-2985) is decided by species at mixing ratios of
1e-15 to 1e-21 whose number densities sit just above atol, so two correct
implementations (the oracle's LAPACK band solve, the device's block solves,
the reference's own sympy-rounded Jacobian) agree step by step to 1e-10 of
the layer totals while their delta, and with it dt and the occasional
rejection, drifts apart once dt is large (from about step 690 of these
runs).  The runs are therefore checked (1) exactly where no rounding enters
the decision: conv() on the reference's own states gives the reference's
longdy, longdydt and verdict bit for bit; (2) step by step over the first
600 steps (times to 1e-8); (3) at the end: end_case 1, the step count within
5 % of the reference's, and the converged mixing ratios within yconv_cri
(0.01, relative) of the reference's steady state wherever they exceed 1e-10.
"""
import ctypes as C
import types

import numpy as np
import pytest

from _golden_util import load, network_from, column_from, traj_config, rel_err
from test_conv import check_steady_state, crafted_cases, reference_conv_tests, run_config
from vulcan_amd import _lib
from vulcan_amd.config import Config

gpu = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    return torch


@pytest.fixture(scope='module')
def gd():
    return load()


@pytest.fixture(scope='module')
def cg():
    return load('ref_conv_nz8.npz')


def vulcan_objects(d):
    """store.Variables / AtmData / Parameters-shaped objects for the golden
    column, as VULCAN holds them after its setup (build_atm.py, ele_sum)."""
    net = network_from(d)
    y0 = d['y0'].copy()
    nz = y0.shape[0]
    atm = types.SimpleNamespace(
        pco=d['pco'].copy(), pico=d['pico'].copy(), Tco=d['Tco'].copy(), Kzz=d['Kzz'].copy(),
        vz=np.zeros(nz - 1), M=d['M'].copy(), n_0=d['M'].copy(), Dzz=d['Dzz'].copy(), alpha=d['alpha'].copy(),
        ms=d['ms'].copy(), mu=np.zeros(nz), g=d['g0'].copy(), Hp=d['Hp0'].copy(), dz=d['dz0'].copy(),
        zco=d['zco0'].copy(), dzi=d['dzi0'].copy(), Ti=d['Ti0'].copy(), Hpi=d['Hpi0'].copy(),
        gs=float(d['gs']), Rp=float(d['Rp']), pref_indx=int(d['pref_indx']))
    atoms = net.atoms()
    A = net.atom_matrix()
    var = types.SimpleNamespace(
        y=y0, ymix=y0 / np.vstack(np.sum(y0, axis=1)), k={j + 1: d['k'][:, j].copy() for j in range(net.nr)},
        dt=1e-10, t=0.0, longdy=1.0, longdydt=1.0, y_time=[], t_time=[], aflux_change=0.0,
        atom_ini={a: np.sum([A[i, q] * y0[:, i] for i in range(net.ni)]) for q, a in enumerate(atoms)},
        atom_loss={a: 0.0 for a in atoms}, atom_sum={a: 0.0 for a in atoms}, atom_loss_prev={a: 0.0 for a in atoms})
    var.y_prev = var.y.copy()
    para = types.SimpleNamespace(count=0, delta=0.0, delta_count=0, nega_count=0, loss_count=0, small_y=0.0,
                                 nega_y=0.0, end_case=0, solver_str='solver')
    return net, var, atm, para


def _traj_geometry(d, atm):
    for key in ('g', 'Hp', 'dz', 'zco', 'dzi', 'Ti', 'Hpi'):
        setattr(atm, key, d['trajg0_' + key].copy())


# --------------------------------------------------------------------------
# chem_funs surface
# --------------------------------------------------------------------------

@gpu
def test_chem_funs_surface_against_reference(gd):
    """ChemFuns.chemdf / neg_symjac / symjac / Gibbs with the generated
    module's argument conventions (k as VULCAN's 1-based dict)."""
    _torch()
    from vulcan_amd.chem_funs import ChemFuns, as_module
    d = gd
    net = network_from(d)
    cf = ChemFuns(net)
    kdict = {j + 1: d['k'][:, j] for j in range(net.nr)}
    assert np.array_equal(cf.chemdf(d['y0'], d['M'], kdict), d['chemdf'])
    # the generated Jacobian sums its terms in sympy's order: equal to rounding
    # (row-scaled 2e-15, as the oracle), zero rows alike
    nj = d['negjac']
    scale = np.maximum(np.abs(nj).max(axis=2, keepdims=True), 1e-300)
    J = cf.neg_symjac(d['y0'], d['M'], kdict)
    assert np.max(np.abs(J - nj) / scale) < 2e-15
    assert np.array_equal(np.abs(J).max(axis=2) == 0, np.abs(nj).max(axis=2) == 0)
    assert np.array_equal(cf.symjac(d['y0'], d['M'], d['k']), -J)
    dense = cf.neg_symjac(d['y0'], d['M'], kdict, dense=True)
    ni, nz = net.ni, d['y0'].shape[0]
    assert dense.shape == (ni * nz, ni * nz)
    assert np.array_equal(dense[ni:2 * ni, ni:2 * ni], J[1]) and not dense[:ni, ni:].any()
    with np.errstate(all='ignore'):
        for p in (0, 7, net.npairs - 1):
            K = cf.Gibbs(2 * p + 1, d['Tg'])
            assert rel_err(K, d['Kg'][p]) < 1e-12, p
    mod = as_module(net)
    assert mod.spec_list == [str(s) for s in d['species']] and mod.ni == net.ni and mod.nr == net.nr


# --------------------------------------------------------------------------
# Ros2 plugin surface
# --------------------------------------------------------------------------

@gpu
def test_ros2_solver_one_step_step_size_against_reference(gd):
    """naming_solver / solver / one_step (with the reference's rejections) /
    step_size on VULCAN-shaped objects, as op.Integration calls them."""
    _torch()
    from vulcan_amd.op import Ros2
    d = gd
    net, var, atm, para = vulcan_objects(d)
    solver = Ros2(net, Config(ode_solver='Ros2_amd'))
    solver.naming_solver(para)
    assert para.solver_str == 'solver'
    # one attempt
    var.dt = float(d['solver_dt'])
    var, para = solver.solver(var, atm, para)
    ref = d['solver_sol']
    ys = ref.sum(axis=1)[:, None]
    assert np.max(np.abs(var.y - ref) / np.maximum(np.abs(ref), 1e-10 * ys)) < 1e-10
    assert abs(para.delta - float(d['solver_delta'])) <= 1e-9 * float(d['solver_delta'])
    # one_step from the initial state at dt = 1e-7: the reference rejects
    net, var, atm, para = vulcan_objects(d)
    var.dt = float(d['onestep_dt_in'])
    var.y_prev = var.y.copy()
    var.atom_loss_prev = var.atom_loss.copy()
    var, para = solver.one_step(var, atm, para)
    counts = [para.delta_count, para.nega_count, para.loss_count]
    assert counts == [int(x) for x in d['onestep_counts']]
    assert sum(counts) > 0
    ref = d['onestep_y']
    ys = ref.sum(axis=1)[:, None]
    rel = np.max(np.abs(var.y - ref) / np.maximum(np.abs(ref), 1e-10 * ys))
    print('one_step after %d rejections: rel %.3e' % (sum(counts), rel))
    assert rel < 1e-10
    var = solver.step_size(var, para)
    assert rel_err(var.dt, float(d['onestep_dt_out'])) < 1e-9


# --------------------------------------------------------------------------
# Integration surface
# --------------------------------------------------------------------------

@gpu
def test_integration_reproduces_reference_trajectory(gd):
    """vulcan_amd.op.Integration on the reference's 61-step run (rejections,
    update_mu_dz every 10 steps, stop at count_max): counts, the full t_time
    and y_time histories and the final state."""
    _torch()
    from vulcan_amd.op import Integration, Ros2
    d = gd
    net, var, atm, para = vulcan_objects(d)
    _traj_geometry(d, atm)
    cfg = traj_config(d)
    Integration(Ros2(net, cfg), poll=8)(var, atm, para)
    fin = d['traj_final']
    assert para.end_case == int(fin[4]) and para.count == int(fin[5])
    assert [para.delta_count, para.nega_count, para.loss_count] == [int(x) for x in fin[6:9]]
    assert len(var.t_time) == para.count and len(var.y_time) == para.count
    assert rel_err(np.asarray(var.t_time), d['traj_t']) < 1e-9
    ty = d['traj_y']
    rel = rel_err(np.stack(var.y_time), ty, floor=1e-6 * d['M'].min())
    assert rel < 1e-6, rel
    assert rel_err(var.longdy, fin[2]) < 1e-6 and rel_err(var.dt, fin[0]) < 1e-6
    assert rel_err(atm.g, d['traj_geom'][0]) < 1e-12


@gpu
def test_integration_resumes_from_histories(gd):
    """A run stopped at count_max and resumed from var.t_time / var.y_time
    (para.count > 0) continues as one uninterrupted run does."""
    _torch()
    from vulcan_amd.op import Integration, Ros2
    d = gd
    cfg = traj_config(d)
    net, var, atm, para = vulcan_objects(d)
    _traj_geometry(d, atm)
    Integration(Ros2(net, cfg), poll=8)(var, atm, para)
    straight = (para.count, var.t, var.y.copy(), var.longdy)
    cfg1 = traj_config(d)
    cfg1.count_max = 29
    net, var, atm, para = vulcan_objects(d)
    _traj_geometry(d, atm)
    Integration(Ros2(net, cfg1), poll=8)(var, atm, para)
    assert para.end_case == 3 and para.count == 30
    with pytest.raises(ValueError):
        v2 = types.SimpleNamespace(**vars(var))
        v2.t_time = []
        Integration(Ros2(net, cfg), poll=8)(v2, atm, types.SimpleNamespace(**vars(para)))
    para.end_case = 0
    Integration(Ros2(net, cfg), poll=8)(var, atm, para)
    assert para.count == straight[0] and para.end_case == int(d['traj_final'][4])
    assert rel_err(var.t, straight[1]) < 1e-12
    assert rel_err(var.y, straight[2], floor=1e-6 * d['M'].min()) < 1e-9
    assert rel_err(var.longdy, straight[3]) < 1e-9


@gpu
def test_integration_to_steady_state_with_conver_ignore(gd, cg):
    """Run b of the reference (conver_ignore set, end_case 1) through
    vulcan_amd.op.Integration."""
    _torch()
    from vulcan_amd.op import Integration, Ros2
    ign = [str(s) for s in cg['runb_ignore']]
    net, var, atm, para = vulcan_objects(gd)
    Integration(Ros2(net, run_config(cg, ign)))(var, atm, para)
    check_steady_state(cg, 'b', para.count, para.end_case, np.asarray(var.t_time), var.y)


# --------------------------------------------------------------------------
# the steady-state decision on the device
# --------------------------------------------------------------------------

def _conv_device(torch, species, nz, y_time_of, count, t_time, y, ymix, M, Kzz, Hp, cfg, aflux=0.0, t=None):
    """vk_conv on one column: y_time_of(q) gives the state after step q+1
    (None where conv() does not read it)."""
    from vulcan_amd import network as nw
    from vulcan_amd.device import DeviceColumns, DevicePlan, dptr, params_from, stream_handle
    ni = len(species)
    net = nw.Network(species=list(species), elements=['H'], compo=np.ones((ni, 1), np.int32),
                     mass=np.full(ni, 2.0), nasa9=np.zeros((ni, 20)),
                     reac=np.array([[0, -2, -2, -2]], np.int32), prod=np.array([[1, -2, -2, -2]], np.int32),
                     kind=np.zeros(1, np.int32), par=np.zeros((1, 8)), rev=np.zeros(1, np.int32))
    plan, cols = DevicePlan(net), DeviceColumns([nz])
    dev = torch.device('cuda')
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    H = count
    hist = np.zeros((H, nz, ni))
    for q in range(count):
        s = y_time_of(q)
        if s is not None:
            hist[q % H] = s
    Md, Kd, Hd = T(M), T(np.append(Kzz, 0.0)), T(Hp)
    yd, md, hd, td = T(y), T(ymix), T(hist), T(np.asarray(t_time)[None])
    atm = _lib.Atm(M=dptr(Md), Kzz=dptr(Kd), Hp=dptr(Hd))
    prm = params_from(cfg, use_conv=True, species=species)
    st = np.zeros(1, _lib.COLSTATE_DTYPE)
    st['t'] = t_time[-1] if t is None else t
    st['count'] = count
    st['aflux_change'] = aflux
    sd = torch.as_tensor(st.view(np.uint8).copy(), device=dev)
    _lib.check(_lib.load().vk_conv(plan.handle, cols.handle, C.byref(atm), C.byref(prm), dptr(yd), dptr(md),
                                   dptr(hd), H, dptr(td), count, dptr(sd), stream_handle()), 'vk_conv')
    out = sd.cpu().numpy().view(_lib.COLSTATE_DTYPE)[0]
    return out


@gpu
def test_device_conv_on_crafted_histories(cg):
    """k_conv on histories that reach each branch of conv(): the verdict and
    longdy / longdydt bit for bit as the reference computed them."""
    torch = _torch()
    species = [str(s) for s in cg['case_species']]
    for c in crafted_cases(cg):
        yt = c['y_time']
        y = yt[-1]
        cfg = Config(conv_step=c['conv_step'], conver_ignore=c['ignore'], trun_min=0.0, count_min=0)
        out = _conv_device(torch, species, y.shape[0], lambda q: yt[q], c['count'], c['t_time'], y,
                           y / y.sum(axis=1)[:, None], cg['case_n0'], cg['case_Kzz'], cg['case_Hp'], cfg,
                           aflux=c['aflux'])
        assert (int(out['end_case']) == 1) == c['res'], c['name']
        assert out['longdy'] == c['longdy'], (c['name'], out['longdy'], c['longdy'])
        assert out['longdydt'] == c['longdydt'], c['name']


@gpu
def test_device_conv_on_reference_states(gd, cg):
    """k_conv on the states the reference's last three conv() tests of each
    run read: the same longdy, longdydt and verdict, bit for bit."""
    torch = _torch()
    d = gd
    species = [str(s) for s in d['species']]
    for tag, j, ct in reference_conv_tests(cg):
        cfg = run_config(cg, ct['ignore'])
        cfg.trun_min, cfg.count_min = 0.0, 0
        out = _conv_device(torch, species, d['y0'].shape[0], lambda q: ct['states'].get(q), ct['count'],
                           ct['t_time'], ct['y'], ct['ymix'], d['M'], d['Kzz'], ct['Hp'], cfg, t=ct['t'])
        assert (int(out['end_case']) == 1) == ct['res'], (tag, j)
        assert out['longdy'] == ct['longdy'], (tag, j, out['longdy'], ct['longdy'])
        assert out['longdydt'] == ct['longdydt'], (tag, j)


@gpu
@pytest.mark.parametrize('tag', ['a', 'b'])
def test_device_integration_reaches_reference_steady_state(gd, cg, tag):
    """The device loop (graph replay, conv every step) from the reference's
    initial state to end_case 1 with the default criteria."""
    _torch()
    from vulcan_amd.device import ColumnBatch, DevicePlan
    d = gd
    net, col = network_from(d), column_from(d)
    ign = [str(s) for s in cg['run%s_ignore' % tag]]
    cfg = run_config(cg, ign)
    b = ColumnBatch(DevicePlan(net, alpha=d['alpha']), [col], [d['y0']], cfg, history=cfg.conv_step + 1,
                    k=d['k'], graph=True)
    for _ in range(200):
        b.integrate(64)
        if int(b.host_state()['done'][0]):
            break
    h = b.host_state()[0]
    t, _ = b.read_history(int(h['count']), int(h['count']))
    check_steady_state(cg, tag, int(h['count']), int(h['end_case']), t[0], b.y.cpu().numpy())
    b.close()
