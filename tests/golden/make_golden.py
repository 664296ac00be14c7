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
* ``ReadRate.read_rate`` forward rates of ``thermo/CHO_thermo_network.txt``
  and ``thermo/SNCHO_photo_network_2024.txt``,
* two ``Integration`` runs to steady state and ``conv()`` on crafted
  histories (``ref_conv_nz8.npz``).

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


def setup():
    """Imports the reference and builds the seeded column of the generated
    network (rates consistent with its K_eq); returns the objects main()
    and the other generators work on."""
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
    return dict(cfg=cfg, chem_funs=chem_funs, store=store, build_atm=build_atm, op=op, rng=rng, species=species,
                ni=ni, nr=nr, R=R, sp_idx=sp_idx, reac=reac, prod=prod, nasa9=nasa9, elements=elements, compo=compo,
                mass=mass, Tg=Tg, Kg=Kg, var=var, atm=atm, para=para, make_atm=make_atm, geom0=geom0, ini=ini,
                karr=karr, solver=solver)


def main():
    ctx = setup()
    cfg, chem_funs, store, build_atm, op = ctx['cfg'], ctx['chem_funs'], ctx['store'], ctx['build_atm'], ctx['op']
    rng, species, ni, nr, R, sp_idx = ctx['rng'], ctx['species'], ctx['ni'], ctx['nr'], ctx['R'], ctx['sp_idx']
    reac, prod, nasa9, elements, compo, mass = (ctx[k] for k in ('reac', 'prod', 'nasa9', 'elements', 'compo',
                                                                 'mass'))
    Tg, Kg, var, atm, para, make_atm = ctx['Tg'], ctx['Kg'], ctx['var'], ctx['atm'], ctx['para'], ctx['make_atm']
    geom0, ini, karr, solver = ctx['geom0'], ctx['ini'], ctx['karr'], ctx['solver']
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

    # the same for the 2024 S-N-C-H-O photochemical network (ReadRate reads
    # whichever file vulcan_cfg.network names); a wider temperature range
    # reaches both NASA9 ranges and the falloff region of its three-body rows
    atm3 = store.AtmData()
    atm3.Tco = np.linspace(2500., 300., NZ)
    atm3.M = 10.0 ** np.linspace(20., 12., NZ)
    cfg.network = 'thermo/SNCHO_photo_network_2024.txt'
    var3 = store.Variables()
    rr3 = op.ReadRate()
    with contextlib.redirect_stdout(io.StringIO()):
        var3 = rr3.read_rate(var3, atm3)
    idx3 = sorted(i for i in var3.Rf if i % 2 == 1)
    np.savez_compressed(
        os.path.join(OUT, 'ref_readrate_SNCHO_2024.npz'),
        Tco=atm3.Tco, M=atm3.M, index=np.array(idx3), eq=np.array([var3.Rf[i] for i in idx3]),
        kf=np.stack([var3.k[i] for i in idx3], axis=1))
    cfg.network = 'thermo/CHO_thermo_network.txt'

    conv_golden(cfg, op, store, make_atm, ini, solver, var, atm, para, y0, species, geom0)
    print('wrote', OUT)


def _run_to_convergence(cfg, op, store, make_atm, ini, solver, var, atm, para, y0, geom0):
    """Integration.__call__ from y0 (restored geometry) until stop() ends it."""
    for key, v in geom0.items():
        setattr(atm, key[:-1], v.copy())
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
    integ = op.Integration(solver, op.Output())
    with contextlib.redirect_stdout(io.StringIO()):
        integ(var, atm, para, make_atm)
    return var, para, integ


def conv_golden(cfg, op, store, make_atm, ini, solver, var, atm, para, y0, species, geom0):
    """Reference runs that end in steady state (end_case 1), and conv() on
    crafted histories that reach each branch of its decision (op.py:1018-1085)."""
    gold = {}
    cfg.update_frq = 100
    cfg.count_min = 120
    cfg.trun_min = 1e2
    cfg.conv_step = 500
    cfg.count_max = 4000
    runs = {}
    # run a: nothing ignored.  Its conv() tests are recorded (longdy/ymix per
    # entry, longdy and the time span), and run b ignores the smallest set of
    # species (largest longdy/ymix first, at most three) that would have let
    # an earlier test of run a pass; conver_ignore changes nothing else, so
    # run b follows run a's trajectory and stops at that earlier step.
    conv_impl = op.Integration.conv
    record_a = None
    for tag in ('a', 'b'):
        record = []

        def conv_rec(self, var_, para_, atm_, *a, **kw):
            snap = dict(count=para_.count, t=var_.t, ymix=var_.ymix.copy(), y_time=var_.y_time,
                        t_time=np.array(var_.t_time), Hp=atm_.Hp.copy())
            res = conv_impl(self, var_, para_, atm_, *a, **kw)
            snap.update(res=res, longdy=var_.longdy, longdydt=var_.longdydt,
                        wv=np.nan_to_num(np.asarray(para_.where_varies_most), nan=0.0, posinf=0.0).copy(),
                        span=var_.longdy / var_.longdydt if var_.longdydt > 0 else np.inf)
            record.append(snap)
            return res
        if tag == 'a':
            cfg.conver_ignore = []
        else:
            slope_min = max(min(np.amin(atm.Kzz / (0.1 * atm.Hp[:-1]) ** 2), 1.e-8), 1.e-10)
            pick = None
            for sn in record_a:
                cnt, t, wv, span, res = sn['count'], sn['t'], sn['wv'], sn['span'], sn['res']
                if res or not (t > cfg.trun_min and cnt > cfg.count_min):
                    continue
                per_sp = wv.max(axis=0)
                order = np.argsort(-per_sp)
                for nign in (1, 2, 3):
                    rest = np.delete(per_sp, order[:nign])
                    ld = rest.max() if rest.size else 0.0
                    if (ld < cfg.yconv_cri and ld / span < cfg.slope_cri or
                            ld < cfg.yconv_min and ld / span < slope_min):
                        pick = [species[int(q)] for q in order[:nign]]
                        break
                if pick:
                    break
            assert pick, 'no earlier test of run a passes with up to three species ignored'
            cfg.conver_ignore = pick
        op.Integration.conv = conv_rec
        var, para, integ = _run_to_convergence(cfg, op, store, make_atm, ini, solver, var, atm, para, y0, geom0)
        op.Integration.conv = conv_impl
        if tag == 'a':
            record_a = record
        count = para.count
        keep = sorted(set([100, 200, 300, count // 2, count - 1]) & set(range(count)))
        runs[tag] = count
        gold['run%s_ignore' % tag] = np.array(cfg.conver_ignore, dtype='U16')
        gold['run%s_final' % tag] = np.array([var.dt, para.delta, var.longdy, var.longdydt, para.end_case,
                                               para.count, para.delta_count, para.nega_count, para.loss_count,
                                               var.t])
        gold['run%s_t' % tag] = np.array(var.t_time)
        gold['run%s_keep' % tag] = np.array(keep)
        gold['run%s_y' % tag] = np.stack([var.y_time[q] for q in keep])
        gold['run%s_ymix' % tag] = var.ymix.copy()
        gold['run%s_geom' % tag] = np.stack([atm.g, atm.Hp])
        # the last three conv() tests with the states they read: y_time[count-1]
        # and the states around the window start (argmin |t_time - t/2|,
        # stepped back from count-1, cut at count - conv_step)
        for j, sn in enumerate(record[-3:]):
            cnt, tt_ = sn['count'], sn['t_time']
            ix = int(np.abs(tt_ - sn['t'] * cfg.st_factor).argmin())
            if ix == cnt - 1:
                ix -= 1
            ix = max(cnt - cfg.conv_step, ix)
            qs = sorted(set(q for q in (ix - 1, ix, ix + 1, cnt - 1) if 0 <= q < cnt))
            pre = 'run%s_ct%d_' % (tag, j)
            gold[pre + 'meta'] = np.array([cnt, sn['t'], float(sn['res']), sn['longdy'], sn['longdydt']])
            gold[pre + 'ymix'] = sn['ymix']
            gold[pre + 'q'] = np.array(qs)
            gold[pre + 'ys'] = np.stack([sn['y_time'][q] for q in qs])
            gold[pre + 't_time'] = tt_[:cnt]
            gold[pre + 'Hp'] = sn['Hp']
    gold['run_cfg'] = np.array([cfg.update_frq, cfg.count_min, cfg.trun_min, cfg.conv_step, cfg.count_max])

    # ---- conv() on crafted histories -----------------------------------------------
    # nz = 4 layers, ni = 6 species (the first six of the network, so that
    # conver_ignore names resolve through species.index)
    rng = np.random.default_rng(77)
    nzc, nic = 4, 6
    catm = store.AtmData()
    catm.n_0 = 10.0 ** rng.uniform(14, 18, nzc)
    catm.Kzz = np.array([450., 1e3, 1e4])   # slope_min = 5e-9, inside its clip range
    catm.Hp = np.array([3e6, 3.2e6, 3.5e6, 3.9e6])
    integ = op.Integration(solver, op.Output())
    base = 10.0 ** rng.uniform(-8, -1, (nzc, nic))
    base /= base.sum(axis=1)[:, None]
    cases = []

    def case(name, count, t_time, drift, conv_step=500, ignore=(), aflux=0.0, tiny=None):
        """y_time[q] = n_0 * base * (1 + drift[q]); tiny marks entries set below atol."""
        yt = []
        for q in range(count):
            ym = base * (1.0 + drift[q])
            y = np.vstack(catm.n_0) * ym
            if tiny is not None:
                y[tiny] = 0.01 * (1.0 + drift[q][tiny])
            yt.append(y)
        cases.append(dict(name=name, count=count, t_time=np.asarray(t_time, float), y_time=np.stack(yt),
                          conv_step=conv_step, ignore=list(ignore), aflux=aflux))

    n = 60
    tt = np.cumsum(np.geomspace(1.0, 1e10, n))      # geometric steps: argmin near the end
    tlin = np.arange(1.0, n + 1.0)                   # uniform steps: argmin mid-history

    def jump(a):
        """no change until the last state, which moves every entry by a factor 1 + a"""
        d = np.zeros((n, nzc, nic))
        d[-1] = a
        return d

    ramp = lambda a: a * np.linspace(0.0, 1.0, n)[:, None, None] * np.ones((n, nzc, nic))
    # branch 1: longdy < yconv_cri and longdydt < slope_cri
    case('branch1', n, tt, jump(1e-3))
    # branch 2: yconv_cri <= longdy < yconv_min and longdydt < slope_min
    case('branch2', n, tt * 1e8, jump(0.05))
    # longdy between the two bounds but longdydt between slope_min and slope_cri
    case('between_slopes', n, tt * 1e-5, jump(0.05))
    # neither: longdy small but changing too fast
    case('too_fast', n, np.cumsum(np.geomspace(1e-6, 1.0, n)), jump(5e-3))
    # neither: longdy above yconv_min
    case('too_large', n, tt * 1e8, jump(0.5))
    # branch 1 met but the actinic-flux change is too large
    case('aflux', n, tt, jump(1e-3), aflux=0.5)
    # one species moves: not converged, converged once it is ignored
    d1 = jump(1e-3)
    d1[-1, :, 2] = 0.8
    case('one_species', n, tt, d1)
    case('one_species_ignored', n, tt, d1, ignore=(species[2],))
    # entries below atol are masked out whatever they do
    d2 = jump(1e-3)
    d2[-1, 1, 4] = 5.0
    mask = np.zeros((nzc, nic), bool)
    mask[1, 4] = True
    case('below_atol', n, tt, d2, tiny=mask)
    # the last step dominates t: argmin lands on count-1 and steps back one
    tl = np.append(np.arange(1.0, n), 1e12)
    case('last_step_dominates', n, tl, ramp(1e-3))
    # the window is cut at count - conv_step (argmin mid-history)
    case('window_cut', n, tlin, ramp(2e-3), conv_step=10)
    case('window_full', n, tlin, ramp(2e-3))
    out = []
    for c in cases:
        cv = store.Variables()
        cp = store.Parameters()
        cv.y_time = list(c['y_time'])
        cv.t_time = list(c['t_time'])
        cv.y = c['y_time'][-1]
        cv.ymix = cv.y / np.vstack(np.sum(cv.y, axis=1))
        cv.t = c['t_time'][-1]
        cv.aflux_change = c['aflux']
        cp.count = c['count']
        cfg.conv_step = c['conv_step']
        cfg.conver_ignore = c['ignore']
        with contextlib.redirect_stdout(io.StringIO()):
            res = integ.conv(cv, cp, catm)
        out.append((c, bool(res), cv.longdy, cv.longdydt))
    for q, (c, res, ld, ldt) in enumerate(out):
        pre = 'case%d_' % q
        gold[pre + 'name'] = np.array(c['name'])
        gold[pre + 'y_time'] = c['y_time']
        gold[pre + 't_time'] = c['t_time']
        gold[pre + 'meta'] = np.array([c['count'], c['conv_step'], c['aflux']])
        gold[pre + 'ignore'] = np.array(c['ignore'], dtype='U16')
        gold[pre + 'out'] = np.array([float(res), ld, ldt])
    gold['ncases'] = np.array(len(out))
    gold['case_n0'] = catm.n_0
    gold['case_Kzz'] = catm.Kzz
    gold['case_Hp'] = catm.Hp
    gold['case_species'] = np.array(species[:nic])
    np.savez_compressed(os.path.join(OUT, 'ref_conv_nz8.npz'), **gold)
    print('conv runs: a %d steps, b %d steps (ignore %s); cases %s' % (
        runs['a'], runs['b'], list(gold['runb_ignore']), [(c['name'], r) for c, r, _, _ in out]))


if __name__ == '__main__':
    main()
