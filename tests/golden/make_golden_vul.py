"""Write the structure of the reference's `.vul` output (Output.save_out,
op.py:3216-3256) as golden data.

Run in the development container only (the reference is not available
elsewhere):

    python tests/golden/make_golden_vul.py      (runs itself again with --photo)

Runs the reference's own ``Output.save_out`` twice, on the seeded 8-layer
column of ``make_golden.py``'s ``setup()`` after a short ``Integration``
(C-H-O thermochemical network, ``save_evolution`` on) and on the
photochemical C-H-O network of ``make_golden_photo.py`` after its set-up
radiative transfer (``use_photo``), with ``pickle.dump`` intercepted, and
records the structure of the dictionary it would pickle: for every key of
'variable', 'atm' and 'parameter', the type and, for arrays, shape and
dtype (dictionaries: their key count and value type; lists: their length).
Nothing is pickled or unpickled; the output is JSON.
"""
import contextlib
import io
import json
import os
import sys
import tempfile

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_golden import NZ, OUT, setup  # noqa: E402


def describe(v):
    if isinstance(v, np.ndarray):
        return {'type': 'ndarray', 'shape': list(v.shape), 'dtype': str(v.dtype)}
    if isinstance(v, dict):
        vals = list(v.values())
        vt = sorted({type(x).__name__ for x in vals})
        return {'type': 'dict', 'len': len(v), 'value_types': vt}
    if isinstance(v, (list, tuple)):
        return {'type': type(v).__name__, 'len': len(v)}
    return {'type': type(v).__name__}


def capture(op, var, atm, para):
    """The dictionary save_out pickles, without pickling it."""
    got = {}

    def fake_dump(obj, f, protocol=None):
        got['d'] = obj
    real = op.pickle.dump
    op.pickle.dump = fake_dump
    try:
        with tempfile.TemporaryDirectory() as tmp:
            op.vulcan_cfg.output_dir = tmp + '/'
            with contextlib.redirect_stdout(io.StringIO()):
                op.Output().save_out(var, atm, para, '')
    finally:
        op.pickle.dump = real
    d = got['d']
    return {sec: {k: describe(v) for k, v in d[sec].items()} for sec in ('variable', 'atm', 'parameter')}


def thermo_run():
    ctx = setup()
    cfg, op, var, atm, para = ctx['cfg'], ctx['op'], ctx['var'], ctx['atm'], ctx['para']
    cfg.use_plot_end = cfg.use_plot_evo = False
    cfg.output_humanread = False
    cfg.save_evolution = True
    cfg.save_evo_frq = 5
    cfg.update_frq = 10
    cfg.count_min = 20
    cfg.trun_min = 1e-8
    cfg.conv_step = 10
    cfg.count_max = 30
    var.y_ini = var.y.copy()
    var.dt = cfg.dttry
    var.t = 0
    var.y_time, var.t_time, var.atom_loss_time = [], [], []
    para.count = 0
    integ = op.Integration(ctx['solver'], op.Output())
    with contextlib.redirect_stdout(io.StringIO()):
        integ(var, atm, para, ctx['make_atm'])
    return capture(op, var, atm, para), dict(nz=NZ, ni=ctx['ni'], nr=ctx['nr'], count=para.count,
                                             save_evo_frq=cfg.save_evo_frq)


def photo_run():
    from make_golden_photo import setup_photo
    ctx = setup_photo()
    cfg, op, var, atm, solver = ctx['cfg'], ctx['op'], ctx['var'], ctx['atm'], ctx['solver']
    cfg.use_plot_end = cfg.use_plot_evo = False
    cfg.output_humanread = False
    cfg.save_evolution = False
    ini = ctx['build_atm'].InitialAbun()
    var = ini.ele_sum(var)
    var.y_ini = var.y.copy()
    solver.compute_tau(var, atm)
    solver.compute_flux(var, atm)
    solver.compute_J(var, atm)
    para = ctx['store'].Parameters()
    return capture(op, var, atm, para), dict(nz=len(atm.Tco), ni=ctx['ni'], nbin=len(var.bins),
                                             T_cross_sp=list(cfg.T_cross_sp))


def main():
    path = os.path.join(OUT, 'ref_vul_structure.json')
    if len(sys.argv) > 1 and sys.argv[1] == '--photo':
        # a process of its own: the photochemical set-up imports the
        # reference's modules with another network
        with open(path) as f:
            out = json.load(f)
        out['photo'], out['photo_meta'] = photo_run()
    else:
        out = {}
        out['thermo'], out['thermo_meta'] = thermo_run()
    with open(path, 'w') as f:
        json.dump(out, f, indent=1, sort_keys=True)
    for run in ('thermo', 'photo'):
        if run in out:
            print('%s: variable %d keys, atm %d keys, parameter %d keys' % (
                run, len(out[run]['variable']), len(out[run]['atm']), len(out[run]['parameter'])))
    if len(sys.argv) == 1:
        import subprocess
        subprocess.run([sys.executable, os.path.abspath(__file__), '--photo'], check=True)


if __name__ == '__main__':
    main()
