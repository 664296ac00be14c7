"""Write golden vectors of the reference's photochemistry radiative transfer.

Run in the development container only (the reference is not available
elsewhere):

    python tests/golden/make_golden_photo.py

Imports the reference's modules read-only from /root/reference (bytecode
caching off, plotting off; nothing there is written) and records, for the
C-H-O photochemical network (thermo/CHO_photo_network.txt: 17 photolysed
species, 36 photolysis branches) on an 8-layer column, with the
temperature-dependent cross sections of H2O and CO2 switched on:

* the wavelength bins, absorption / photodissociation-branch / Rayleigh
  cross sections (``make_bins_read_cross``, op.py:495-781) and the stellar
  flux at the top (``read_sflux``, build_atm.py:617-657), which the device
  kernels take as inputs;
* two consecutive calls of ``compute_tau`` / ``compute_flux`` / ``compute_J``
  (op.py:2580-2787) on two states: optical depth, direct beam, diffuse
  fluxes, actinic flux, its relative change and the photolysis rates.

Only data (arrays, strings) is written, as compressed ``.npz``.
"""
import contextlib
import io
import os
import sys

import numpy as np

REF = '/root/reference'
OUT = os.path.dirname(os.path.abspath(__file__))
NZ = 8


def setup_photo():
    """Imports the reference and builds the seeded photochemical column with
    its radiative-transfer tables; returns the objects main() and the other
    generators work on."""
    sys.dont_write_bytecode = True
    os.environ.setdefault('MPLBACKEND', 'Agg')
    os.chdir(REF)
    sys.path.insert(0, REF)
    import vulcan_cfg as cfg
    cfg.nz = NZ
    cfg.use_adapt_rtol = False
    cfg.use_photo = True
    cfg.use_ion = False
    cfg.use_live_plot = False
    cfg.use_live_flux = False
    cfg.use_print_prog = False
    cfg.plot_TP = False
    cfg.network = 'thermo/CHO_photo_network.txt'
    cfg.T_cross_sp = ['H2O', 'CO2']
    import chem_funs
    import store
    import build_atm
    import op

    species = list(chem_funs.spec_list)
    ni = chem_funs.ni
    rng = np.random.default_rng(20250202)
    var = store.Variables()
    atm = store.AtmData()
    make_atm = build_atm.Atm()
    atm = make_atm.f_pico(atm)
    atm.Tco = np.linspace(1500., 250., NZ)        # crosses the 300 K room-temperature table
    atm.Kzz = 10.0 ** np.linspace(6., 10., NZ - 1)
    atm.vz = np.zeros(NZ - 1)
    atm.M = atm.pco / (op.kb * atm.Tco)
    atm.n_0 = atm.M.copy()
    x = 10.0 ** rng.uniform(-10, -3, ni)
    for sp, v in (('H2', 1.0), ('He', 0.16), ('H2O', 1e-3), ('CO', 8e-4), ('CH4', 3e-4), ('CO2', 1e-6)):
        x[species.index(sp)] = v
    xl = np.tile(x, (NZ, 1)) * (1.0 + 0.5 * rng.random((NZ, ni)))
    var.ymix = xl / np.vstack(np.sum(xl, axis=1))
    var.y = atm.n_0[:, None] * var.ymix
    with contextlib.redirect_stdout(io.StringIO()):
        rr = op.ReadRate()
        var = rr.read_rate(var, atm)
        make_atm.mol_diff(atm)
        atm = make_atm.f_mu_dz(var, atm, None)
        solver = op.Ros2()
        rr.make_bins_read_cross(var, atm)
        make_atm.read_sflux(var, atm)

    return dict(cfg=cfg, chem_funs=chem_funs, store=store, build_atm=build_atm, op=op, rng=rng, species=species,
                ni=ni, var=var, atm=atm, make_atm=make_atm, rr=rr, solver=solver)


def main():
    ctx = setup_photo()
    cfg, op, rng, species, ni = ctx['cfg'], ctx['op'], ctx['rng'], ctx['species'], ctx['ni']
    var, atm, solver = ctx['var'], ctx['atm'], ctx['solver']
    photo_sp = sorted(var.photo_sp)
    branches = sorted(var.pho_rate_index)          # (species, branch)
    g = dict(species=np.array(species), Tco=atm.Tco, M=atm.M, dz=atm.dz.copy(),
             bins=var.bins, dbin1=np.array(var.dbin1), dbin2=np.array(var.dbin2),
             din12=np.array(var.sflux_din12_indx), sflux_top=var.sflux_top,
             photo_sp=np.array(photo_sp), scat_sp=np.array(cfg.scat_sp), T_cross_sp=np.array(cfg.T_cross_sp),
             cross=np.stack([var.cross[sp] for sp in photo_sp]),
             cross_scat=np.stack([var.cross_scat[sp] for sp in cfg.scat_sp]),
             cross_T=np.stack([var.cross_T[sp] for sp in cfg.T_cross_sp]),
             branch_sp=np.array([b[0] for b in branches]), branch_no=np.array([b[1] for b in branches]),
             branch_index=np.array([var.pho_rate_index[b] for b in branches]),
             cross_J=np.stack([var.cross_J[b] for b in branches]),
             cross_J_T=np.stack([var.cross_J_T[b] if b[0] in cfg.T_cross_sp else np.zeros((NZ, len(var.bins)))
                                 for b in branches]),
             knobs=np.array([cfg.sl_angle, cfg.edd, cfg.f_diurnal, cfg.flux_atol, op.hc, op.ag0]))
    for call in range(2):
        if call == 1:
            # a second state: the actinic-flux change is now measured against
            # the first call's flux, and the diffuse upward flux of the first
            # call enters the downward sweep (op.py:2705-2708)
            f = 1.0 + 0.3 * rng.random((NZ, ni))
            var.y = var.y * f
            var.ymix = var.y / np.vstack(np.sum(var.y, axis=1))
        solver.compute_tau(var, atm)
        solver.compute_flux(var, atm)
        solver.compute_J(var, atm)
        p = 'c%d_' % call
        g[p + 'y'] = var.y.copy()
        g[p + 'ymix'] = var.ymix.copy()
        g[p + 'tau'] = var.tau.copy()
        g[p + 'sflux'] = var.sflux.copy()
        g[p + 'dflux_u'] = var.dflux_u.copy()
        g[p + 'dflux_d'] = var.dflux_d.copy()
        g[p + 'aflux'] = var.aflux.copy()
        g[p + 'aflux_change'] = np.array(var.aflux_change)
        g[p + 'J'] = np.stack([var.J_sp[b] for b in branches])
        g[p + 'k'] = np.stack([var.k[var.pho_rate_index[b]] for b in branches])
    out = os.path.join(OUT, 'ref_photo_CHO_nz8.npz')
    np.savez_compressed(out, **g)
    print('wrote %s: %d bins, %d photolysed species, %d branches, aflux_change %.3e'
          % (out, len(var.bins), len(photo_sp), len(branches), float(g['c1_aflux_change'])))


if __name__ == '__main__':
    main()
