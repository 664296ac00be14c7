"""Write golden input/output vectors from the reference implementation.

Run in the development container only (the reference is not available
elsewhere):

    python tests/golden/make_golden.py

It imports the reference's modules read-only from /root/reference (no file
there is written: bytecode caching is off and the plotting/output switches
are off) and records, for the network of the shipped generated
``chem_funs.py`` (65 species, 814 rate rows):

* ``chem_funs.Gibbs`` K_eq on a temperature grid crossing 1000 K,
* ``chem_funs.chemdf`` and ``chem_funs.neg_symjac`` (block diagonal),
* ``ODESolver.diffdf`` / ``diffdf_no_mol`` and ``lhs_jac_tot`` /
  ``lhs_jac_no_mol`` (as blocks),
* one ``Ros2.solver`` attempt, one ``one_step`` + ``step_size``,
* a short ``Integration`` trajectory with rejections, ``update_mu_dz`` calls
  and the ``conv`` test,
* ``ReadRate.read_rate`` forward rates of ``thermo/CHO_thermo_network.txt``.

Inputs are seeded synthetic columns; outputs are what the reference computes
on them.  Only data (arrays, strings) is written, as compressed ``.npz``.
"""
import contextlib
import io
import os
import sys

import numpy as np

REF = '/root/reference'
OUT = os.path.dirname(os.path.abspath(__file__))
NZ = 8


def main():
    sys.dont_write_bytecode = True
    os.environ.setdefault('MPLBACKEND', 'Agg')
    os.chdir(REF)
    sys.path.insert(0, REF)
    import vulcan_cfg as cfg
    cfg.nz = NZ
    cfg.use_adapt_rtol = False
    cfg.use_photo = False
    cfg.use_live_plot = False
    cfg.use_live_flux = False
    cfg.use_print_prog = False
    cfg.use_print_delta = False
    cfg.plot_TP = False
    cfg.use_moldiff = True
    cfg.conver_ignore = []
    cfg.network = 'thermo/CHO_thermo_network.txt'
    import chem_funs
    import store
    import build_atm
    import op

    rng = np.random.default_rng(20250101)
    species = list(chem_funs.spec_list)
    ni, nr = chem_funs.ni, chem_funs.nr
    R = nr // 2

    # ---- network tables in the repository's term format --------------------
    sp_idx = {s: i for i, s in enumerate(species)}
    reac = np.full((R, 4), -2, np.int32)
    prod = np.full((R, 4), -2, np.int32)
    for p in range(R):
        j = 2 * p + 1
        r_terms, p_terms = chem_funs.re_wM_dict[j]
        for q, t in enumerate(r_terms):
            reac[p, q] = -1 if t == 'M' else sp_idx[t]
        for q, t in enumerate(p_terms):
            prod[p, q] = -1 if t == 'M' else sp_idx[t]
    nasa9 = np.array([chem_funs.nasa9[s][:20] for s in species])  # low [0:10], high [10:20]
    comp_names = list(build_atm.compo.dtype.names)
    elements = [e for e in comp_names if e not in ('species', 'mass')]
    compo = np.array([[build_atm.compo[build_atm.compo_row.index(s)][e] for e in elements]
                      for s in species], np.int32)
    mass = np.array([build_atm.compo[build_atm.compo_row.index(s)]['mass'] for s in species])
    # monitor only elements present in the network (an absent element gives
    # atom_ini = 0 and a NaN loss, which rejects every attempt)
    cfg.atom_list = [a for a in cfg.atom_list if compo[:, elements.index(a)].any()]

    # ---- K_eq from the generated Gibbs ---------------------------------------
    Tg = np.array([250., 400., 600., 800., 999.9, 1000., 1000.1, 1200., 1600., 2000., 2500., 3000.])
    with np.errstate(all='ignore'):
        Kg = np.array([chem_funs.Gibbs(2 * p + 1, Tg) for p in range(R)])

    # ---- column ----------------------------------------------------------------
    var = store.Variables()
    atm = store.AtmData()
    para = store.Parameters()
    para.start_time = 0.0
    make_atm = build_atm.Atm()
    atm = make_atm.f_pico(atm)
    atm.Tco = np.linspace(1800., 900., NZ)
    atm.Kzz = 10.0 ** np.linspace(6., 10., NZ - 1)
    atm.vz = np.zeros(NZ - 1)
    atm.M = atm.pco / (op.kb * atm.Tco)
    atm.n_0 = atm.M.copy()
    x = 10.0 ** rng.uniform(-12, -1, ni)
    x[sp_idx['H2']] = 1.0
    x[sp_idx['He']] = 0.1
    x /= x.sum()
    # per-layer variation so that eddy diffusion is not a pure cancellation
    xl = np.tile(x, (NZ, 1)) * (1.0 + 0.5 * rng.random((NZ, ni)))
    var.ymix = xl / np.vstack(np.sum(xl, axis=1))
    var.y = atm.n_0[:, None] * var.ymix
    make_atm.mol_diff(atm)
    atm = make_atm.f_mu_dz(var, atm, None)
    geom0 = dict(g0=atm.g.copy(), Hp0=atm.Hp.copy(), dz0=atm.dz.copy(), zco0=atm.zco.copy(),
                 dzi0=atm.dzi.copy(), Ti0=atm.Ti.copy(), Hpi0=atm.Hpi.copy())
    ini = build_atm.InitialAbun()
    var = ini.ele_sum(var)

    # rates: forward/reverse consistent with the reference's own K_eq, both
    # bounded by a random base rate (thermodynamically oriented)
    with np.errstate(all='ignore'):
        for p in range(R):
            j = 2 * p + 1
            base = 10.0 ** (rng.uniform(-32, -29) if -1 in reac[p] else rng.uniform(-14, -11))
            K = chem_funs.Gibbs(j, atm.Tco)
            var.k[j] = base * np.minimum(1.0, K)
            var.k[j + 1] = base * np.minimum(1.0, 1.0 / K)
    karr = np.stack([var.k[j] for j in range(1, nr + 1)], axis=1)

    solver = op.Ros2()
    solver.naming_solver(para)
    gold = {}
    y0 = var.y.copy()
    gold['chemdf'] = chem_funs.chemdf(y0, atm.M, var.k)
    nj = chem_funs.neg_symjac(y0, atm.M, var.k)
    blocks = np.stack([nj[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] for l in range(NZ)])
    off = nj.copy()
    for l in range(NZ):
        off[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] = 0
    assert not off.any(), 'neg_symjac has entries outside the diagonal blocks'
    gold['negjac'] = blocks
    gold['diffdf_mol'] = solver.diffdf(y0, atm)
    gold['diffdf_nomol'] = solver.diffdf_no_mol(y0, atm)

    def to_blocks(a):
        D = np.stack([a[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] for l in range(NZ)])
        L = np.zeros((NZ, ni)); U = np.zeros((NZ, ni))
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

    var.dt = 3.0e-9
    for name, fn in (('tot', solver.lhs_jac_tot), ('nomol', solver.lhs_jac_no_mol)):
        D, L, U = to_blocks(fn(var, atm))
        gold['lhs_%s_D' % name], gold['lhs_%s_L' % name], gold['lhs_%s_U' % name] = D, L, U
    gold['lhs_dt'] = var.dt

    # one solver attempt
    var.dt = 1e-10
    var.y = y0.copy()
    var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
    ymix_in = var.ymix.copy()
    var, para = solver.solver(var, atm, para)
    gold['solver_dt'] = 1e-10
    gold['solver_sol'] = var.y.copy()
    gold['solver_ymix'] = var.ymix.copy()
    gold['solver_delta'] = para.delta
    gold['solver_ymix_in'] = ymix_in

    # one_step + step_size from the initial state at a larger dt
    var.y = y0.copy()
    var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
    var.dt = 1e-7
    var.y_prev = var.y.copy()
    var.atom_loss_prev = var.atom_loss.copy()
    para.delta_count = para.nega_count = para.loss_count = 0
    var, para = solver.one_step(var, atm, para)
    gold['onestep_dt_in'] = 1e-7
    gold['onestep_y'] = var.y.copy()
    gold['onestep_ymix'] = var.ymix.copy()
    gold['onestep_delta'] = para.delta
    gold['onestep_counts'] = np.array([para.delta_count, para.nega_count, para.loss_count])
    gold['onestep_atom_loss'] = np.array([var.atom_loss[a] for a in cfg.atom_list if a in var.atom_loss])
    var = solver.step_size(var, para)
    gold['onestep_dt_out'] = var.dt

    # short integration from the initial state (restored geometry)
    var.y = y0.copy()
    var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
    atm = make_atm.f_mu_dz(var, atm, None)
    var = ini.ele_sum(var)
    var.dt = cfg.dttry
    var.t = 0
    var.y_time, var.t_time, var.atom_loss_time = [], [], []
    para.count = 0
    para.delta_count = para.nega_count = para.loss_count = 0
    para.end_case = 0
    cfg.update_frq = 10
    cfg.count_min = 20
    cfg.trun_min = 1e-8
    cfg.conv_step = 10
    cfg.count_max = 60
    out = op.Output()
    integ = op.Integration(solver, out)
    gold['traj_geom0'] = np.stack([atm.g.copy(), atm.Hp.copy()])
    for key in ('g', 'Hp', 'dz', 'zco', 'dzi', 'Ti', 'Hpi'):
        gold['trajg0_' + key] = getattr(atm, key).copy()
    with contextlib.redirect_stdout(io.StringIO()):
        integ(var, atm, para, make_atm)
    gold['traj_y'] = np.stack(var.y_time)
    gold['traj_t'] = np.array(var.t_time)
    gold['traj_final'] = np.array([var.dt, para.delta, var.longdy, var.longdydt, para.end_case,
                                   para.count, para.delta_count, para.nega_count, para.loss_count])
    gold['traj_geom'] = np.stack([atm.g, atm.Hp, np.append(atm.dzi, np.nan), np.append(atm.Hpi, np.nan)])
    gold['traj_cfg'] = np.array([cfg.update_frq, cfg.count_min, cfg.trun_min, cfg.conv_step, cfg.count_max])

    col = dict(Tco=atm.Tco, pco=atm.pco, pico=atm.pico, M=atm.M, Kzz=atm.Kzz, Dzz=atm.Dzz,
               alpha=atm.alpha, ms=atm.ms, gs=np.array(atm.gs), Rp=np.array(cfg.Rp),
               pref_indx=np.array(atm.pref_indx))
    np.savez_compressed(
        os.path.join(OUT, 'ref_chemfuns_nz8.npz'),
        species=np.array(species), elements=np.array(elements), atom_list=np.array(cfg.atom_list),
        reac=reac, prod=prod, nasa9=nasa9, compo=compo, mass=mass, Tg=Tg, Kg=Kg,
        y0=y0, k=karr, **col, **geom0, **gold)

    # ---- ReadRate on a real network file ---------------------------------------
    var2 = store.Variables()
    rr = op.ReadRate()
    with contextlib.redirect_stdout(io.StringIO()):
        var2 = rr.read_rate(var2, atm)
    idx = sorted(i for i in var2.Rf if i % 2 == 1)
    kf = np.stack([var2.k[i] for i in idx], axis=1)
    np.savez_compressed(
        os.path.join(OUT, 'ref_readrate_CHO_thermo.npz'),
        Tco=atm.Tco, M=atm.M, index=np.array(idx), eq=np.array([var2.Rf[i] for i in idx]), kf=kf)
    print('wrote', OUT)


if __name__ == '__main__':
    main()
