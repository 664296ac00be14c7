"""The steady-state decision (Integration.conv / stop, op.py:1018-1085) of the
oracle against the reference itself.

Golden vectors (``tests/golden/ref_conv_nz8.npz``, written by
``tests/golden/make_golden.py`` from the reference):

* two ``Integration`` runs of the 65-species generated network on 8 layers
  that end in steady state (end_case 1) under the default criteria
  (count_min 120, conv_step 500, trun_min 1e2): run a with nothing ignored,
  run b with a ``conver_ignore`` list that lets it stop earlier;
* ``conv()`` on crafted histories that reach each branch of its decision:
  the yconv_cri/slope_cri branch, the yconv_min/slope_min branch, both
  failing, the actinic-flux gate, a species left out by ``conver_ignore``,
  the atol mask, the ``indx == count - 1`` step back and the conv_step
  window.
"""
import numpy as np
import pytest

from _golden_util import load, network_from, column_from, rel_err
from oracle import ros2_oracle as O
from vulcan_amd import network as nw
from vulcan_amd.config import Config


@pytest.fixture(scope='module')
def cg():
    return load('ref_conv_nz8.npz')


def crafted_cases(g):
    out = []
    for q in range(int(g['ncases'])):
        p = 'case%d_' % q
        count, conv_step, aflux = g[p + 'meta']
        res, ld, ldt = g[p + 'out']
        out.append(dict(name=str(g[p + 'name']), y_time=g[p + 'y_time'], t_time=g[p + 't_time'],
                        count=int(count), conv_step=int(conv_step), aflux=float(aflux),
                        ignore=[str(s) for s in g[p + 'ignore']], res=bool(res), longdy=float(ld),
                        longdydt=float(ldt)))
    return out


class _Col:
    def __init__(self, g):
        self.M = g['case_n0']
        self.Kzz = g['case_Kzz']
        self.Hp = g['case_Hp']


def test_crafted_cases_cover_every_branch(cg):
    names = {c['name']: c['res'] for c in crafted_cases(cg)}
    assert names['branch1'] and names['branch2'] and names['one_species_ignored']
    assert not names['between_slopes'] and not names['too_fast'] and not names['too_large']
    assert not names['aflux'] and not names['one_species']


def test_oracle_conv_matches_reference_on_crafted_histories(cg):
    col = _Col(cg)
    species = [str(s) for s in cg['case_species']]
    for c in crafted_cases(cg):
        yt = c['y_time']
        y = yt[-1].copy()
        st = O.State(y=y, ymix=y / y.sum(axis=1)[:, None], k=None, dt=0.0, atom_ini=None,
                     t=float(c['t_time'][-1]), count=c['count'], aflux_change=c['aflux'],
                     y_time=list(yt), t_time=list(c['t_time']))
        cfg = Config(conv_step=c['conv_step'], conver_ignore=c['ignore'])
        res = O.conv(col, st, cfg, species)
        assert res == c['res'], c['name']
        assert st.longdy == c['longdy'], c['name']
        assert st.longdydt == c['longdydt'], c['name']


def run_config(g, ignore):
    upd, cmin, tmin, cstep, cmax = g['run_cfg']
    return Config(update_frq=int(upd), count_min=int(cmin), trun_min=float(tmin), conv_step=int(cstep),
                  count_max=int(cmax), conver_ignore=list(ignore))


def reference_conv_tests(g):
    """The last three conv() tests of each reference run with the states
    they read: (tag, j, dict(count, t, res, longdy, longdydt, y, ymix,
    states {q: y_time[q]}, t_time, Hp, ignore))."""
    out = []
    for tag in ('a', 'b'):
        ign = [str(s) for s in g['run%s_ignore' % tag]]
        for j in range(3):
            pre = 'run%s_ct%d_' % (tag, j)
            cnt, t, res, ld, ldt = g[pre + 'meta']
            cnt = int(cnt)
            qs = [int(q) for q in g[pre + 'q']]
            states = dict(zip(qs, g[pre + 'ys']))
            out.append((tag, j, dict(count=cnt, t=float(t), res=bool(res), longdy=float(ld), longdydt=float(ldt),
                                     y=states[cnt - 1], ymix=g[pre + 'ymix'], states=states,
                                     t_time=g[pre + 't_time'], Hp=g[pre + 'Hp'], ignore=ign)))
    return out


def test_oracle_conv_on_reference_states(cg):
    """conv() of the oracle on the states the reference's last three conv()
    tests of each run read: bit-for-bit longdy, longdydt and verdict (the
    last test of each run converged, the two before did not)."""
    d = load()
    species = [str(s) for s in d['species']]
    tests = reference_conv_tests(cg)
    for tag, j, ct in tests:
        class Col:
            M = d['M']
            Kzz = d['Kzz']
            Hp = ct['Hp']
        yt = [ct['states'].get(q) for q in range(ct['count'])]
        st = O.State(y=ct['y'].copy(), ymix=ct['ymix'].copy(), k=None, dt=0.0, atom_ini=None, t=ct['t'],
                     count=ct['count'], y_time=yt, t_time=list(ct['t_time']))
        cfg = run_config(cg, ct['ignore'])
        assert O.conv(Col, st, cfg, species) == ct['res'], (tag, j)
        assert st.longdy == ct['longdy'] and st.longdydt == ct['longdydt'], (tag, j)
    verdicts = {(tag, j): ct['res'] for tag, j, ct in tests}
    for tag in ('a', 'b'):
        assert verdicts[(tag, 2)] and not verdicts[(tag, 1)] and not verdicts[(tag, 0)]


@pytest.fixture(scope='module')
def oracle_runs(cg):
    """The oracle integrated from the reference's initial state, once with
    run b's conver_ignore and once without (run a)."""
    d = load()
    out = {}
    for tag in ('a', 'b'):
        net, col = network_from(d), column_from(d)
        tb = nw.build_tables(net)
        cfg = run_config(cg, [str(s) for s in cg['run%s_ignore' % tag]])
        st = O.new_state(net, d['y0'].copy(), d['k'], cfg)
        O.integrate(net, col, st, cfg, tables=tb)
        out[tag] = st
    return out


def check_steady_state(cg, tag, count, end_case, t_time, y):
    """end_case 1; step count within 5 % of the reference's; the first 600
    step times to 1e-8; converged mixing ratios above 1e-10 within yconv_cri
    (0.01 relative) of the reference's steady state."""
    fin = cg['run%s_final' % tag]
    tref = cg['run%s_t' % tag]
    yref = cg['run%s_y' % tag][-1]
    assert end_case == 1 and int(fin[4]) == 1
    assert abs(count - int(fin[5])) <= 0.05 * int(fin[5]), (count, int(fin[5]))
    n = 600
    assert rel_err(np.asarray(t_time[:n]), tref[:n]) < 1e-8
    mix, mref = y / y.sum(axis=1)[:, None], yref / yref.sum(axis=1)[:, None]
    big = mref > 1e-10
    rel = float(np.max(np.abs(mix - mref)[big] / mref[big]))
    print('run %s: %d steps (reference %d), steady state rel %.3e' % (tag, count, int(fin[5]), rel))
    assert rel < 0.01
    return rel


def test_oracle_reaches_the_reference_steady_state(cg, oracle_runs):
    assert int(cg['runb_final'][5]) < int(cg['runa_final'][5]), 'conver_ignore should let run b stop earlier'
    for tag in ('a', 'b'):
        st = oracle_runs[tag]
        check_steady_state(cg, tag, st.count, st.end_case, st.t_time, st.y)


def test_stop_gates_and_limits():
    """stop(): conv is only consulted past trun_min and count_min; runtime
    and count_max end the run with end_case 2 and 3 (op.py:1065-1085)."""
    class C_:
        M = np.ones(2) * 1e18
        Kzz = np.ones(1) * 1e6
        Hp = np.ones(2) * 1e6
    y = np.full((2, 3), 3e17)
    st = O.State(y=y, ymix=y / y.sum(axis=1)[:, None], k=None, dt=1.0, atom_ini=None, t=50.0, count=200,
                 y_time=[y] * 200, t_time=list(np.arange(1.0, 201.0)))
    cfg = Config()
    assert not O.stop(C_, st, cfg) and st.end_case == 0          # t below trun_min
    st.t = 1e3
    assert O.stop(C_, st, cfg) and st.end_case == 1             # no change at all
    st.end_case, st.count = 0, 100
    st.y_time, st.t_time = st.y_time[:100], st.t_time[:100]
    assert not O.stop(C_, st, cfg)                               # count below count_min
    st.t = 2e22
    assert O.stop(C_, st, cfg) and st.end_case == 2
    st.t, st.count = 1.0, cfg.count_max + 1
    assert O.stop(C_, st, cfg) and st.end_case == 3


# --------------------------------------------------------------------------
# the host boundary (no device needed to construct it)
# --------------------------------------------------------------------------

def test_config_accepts_vulcan_cfg_copy_with_registered_solver_name():
    """The binding copies every matching vulcan_cfg attribute into Config,
    including ode_solver = 'Ros2_amd' and conver_ignore."""
    vals = dict(ode_solver='Ros2_amd', use_moldiff=True, use_vz=False, conver_ignore=['HC3N'], rtol=0.25,
                count_max=30000, atm_base='H2')
    cfg = Config(**{k: v for k, v in vals.items() if k in Config.__dataclass_fields__})
    cfg.check_supported()
    assert cfg.ode_solver == 'Ros2_amd'


def test_ros2_rejects_unknown_conver_ignore_species():
    from _golden_util import load as L
    from vulcan_amd.op import Ros2
    net = network_from(L())
    Ros2(net, Config(conver_ignore=[net.species[3]]))
    with pytest.raises(ValueError):
        Ros2(net, Config(conver_ignore=['HC3N_not_here']))


def test_conv_ignore_mask_bits():
    from vulcan_amd.device import conv_ignore_mask
    species = ['s%d' % i for i in range(150)]
    m = conv_ignore_mask(Config(conver_ignore=['s0', 's31', 's32', 's149']), species)
    assert m.dtype == np.uint32 and m.shape == (5,)
    assert m[0] == (1 | (1 << 31)) and m[1] == 1 and m[4] == (1 << (149 - 128))
    with pytest.raises(ValueError):
        conv_ignore_mask(Config(conver_ignore=['x']), species)
