"""Photochemistry radiative transfer against the reference
(tests/golden/ref_photo_CHO_nz8.npz, written by
tests/golden/make_golden_photo.py from the reference: the C-H-O
photochemical network on 8 layers with temperature-dependent H2O and CO2
cross sections).

* host set-up (vulcan_amd.photo.build_tables) against the reference's
  make_bins_read_cross / read_sflux tables: bins, stellar flux, absorption,
  branch and Rayleigh cross sections, the per-layer temperature-dependent
  tables (needs the reference's data files: development container only);
* device update (vk_photo_update) against compute_tau / compute_flux /
  compute_J, on two consecutive states (the second one measures the
  actinic-flux change against the first and starts its downward sweep from
  the first one's upward flux).  The reference sums the optical depth over
  a Python set of species, whose order changes from process to process, so
  the bars are relative (optical depth 1e-13; fluxes, which exponentiate
  it, 1e-10 of each bin's largest value; rates 1e-10).
"""
import os

import numpy as np
import pytest

from _golden_util import load, network_from

REF = '/root/reference'
gpu = pytest.mark.gpu


@pytest.fixture(scope='module')
def pg():
    return load('ref_photo_CHO_nz8.npz')


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, 'thermo', 'photo_cross')),
                    reason='reference data files not present')
def test_host_tables_match_make_bins_read_cross(pg):
    from vulcan_amd import network_text as NT
    from vulcan_amd import photo as P
    net = NT.load_network(os.path.join(REF, 'thermo', 'CHO_photo_network.txt'), os.path.join(REF, 'thermo'))
    pcfg = P.PhotoConfig(T_cross_sp=[str(s) for s in pg['T_cross_sp']])
    t = P.build_tables(net, 'thermo/CHO_photo_network.txt', pg['Tco'], pcfg, root=REF)
    assert np.array_equal(t.bins, pg['bins'])
    assert t.din12 == int(pg['din12'])
    assert np.array_equal(t.sflux_top, pg['sflux_top'])
    assert [net.species[i] for i in t.abs_sp] == [str(s) for s in pg['photo_sp']]
    assert np.array_equal(t.cross, pg['cross'])
    assert np.array_equal(t.cross_scat, pg['cross_scat'])
    assert np.array_equal(t.cross_J, pg['cross_J'])
    Tsp = [str(s) for s in pg['T_cross_sp']]
    for a, i in enumerate(t.abs_sp):
        if t.abs_T[a] >= 0:
            assert np.array_equal(t.cross_T[t.abs_T[a]], pg['cross_T'][Tsp.index(net.species[i])]), net.species[i]
    T_rows = [q for q in range(len(t.br_T)) if t.br_T[q] >= 0]
    assert np.array_equal(t.cross_J_T, pg['cross_J_T'][T_rows])
    # each branch sets the rate the reference's pho_rate_index names: the id
    # printed in the network file (in this file the ids fall two behind the
    # line positions from its 152nd reaction on, so each J lands on the
    # rate of the line before its own, as in the reference)
    assert [c + 1 for c in t.br_col] == [int(x) for x in pg['branch_index']]


def _tables_from_golden(pg, species):
    from vulcan_amd import photo as P
    photo_sp = [str(s) for s in pg['photo_sp']]
    Tsp = [str(s) for s in pg['T_cross_sp']]
    br_sp = [str(s) for s in pg['branch_sp']]
    TJ = [q for q, s in enumerate(br_sp) if s in Tsp]
    sl, edd, fdi, fat, hc, ag0 = pg['knobs']
    assert ag0 == 0
    return P.PhotoTables(bins=pg['bins'], din12=int(pg['din12']), sflux_top=pg['sflux_top'],
                         abs_sp=[species.index(s) for s in photo_sp],
                         abs_T=[Tsp.index(s) if s in Tsp else -1 for s in photo_sp], cross=pg['cross'],
                         cross_T=pg['cross_T'], scat_sp=[species.index(str(s)) for s in pg['scat_sp']],
                         cross_scat=pg['cross_scat'], br_col=[2 * q for q in range(len(br_sp))],
                         br_T=[TJ.index(q) if q in TJ else -1 for q in range(len(br_sp))], cross_J=pg['cross_J'],
                         cross_J_T=pg['cross_J_T'][TJ], dbin1=float(pg['dbin1']), dbin2=float(pg['dbin2']),
                         cos_sl=float(np.cos(sl)), edd=float(edd), f_diurnal=float(fdi), flux_atol=float(fat),
                         hc=float(hc))


def _rel_rows(a, b):
    """max |a - b| / max over levels of |b|, per bin (0/0 -> 0)."""
    scale = np.abs(b).max(axis=0, keepdims=True)
    with np.errstate(invalid='ignore', divide='ignore'):
        r = np.where(scale > 0, np.abs(a - b) / scale, np.abs(a - b))
    return float(np.nanmax(r))


@gpu
def test_device_radiative_transfer_against_reference(pg):
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from vulcan_amd import photo as P
    from vulcan_amd.device import DeviceColumns, DevicePlan
    from vulcan_amd import _lib
    d = load()
    net = network_from(d)
    species = list(net.species)
    assert species == [str(s) for s in pg['species']]
    t = _tables_from_golden(pg, species)
    nz = pg['Tco'].shape[0]
    plan, cols = DevicePlan(net), DeviceColumns([nz])
    ph = P.DevicePhoto(plan, cols, t)
    dev = 'cuda'
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
    dz = T(pg['dz'])
    k = torch.zeros((nz, net.nr), dtype=torch.float64, device=dev)
    st = torch.zeros(_lib.COLSTATE_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    for call in range(2):
        p = 'c%d_' % call
        ph.update(T(pg[p + 'y']), T(pg[p + 'ymix']), dz, k, state=st, force=True)
        torch.cuda.synchronize()
        tau = ph.read(P.VK_PHOTO_TAU)
        r_tau = _rel_rows(tau, pg[p + 'tau'])
        assert r_tau < 1e-13, (call, r_tau)
        out = {}
        for name, what in (('sflux', P.VK_PHOTO_SFLUX), ('dflux_u', P.VK_PHOTO_DFLUX_U),
                           ('dflux_d', P.VK_PHOTO_DFLUX_D), ('aflux', P.VK_PHOTO_AFLUX)):
            out[name] = _rel_rows(ph.read(what), pg[p + name])
        kk = k.cpu().numpy()[:, t.br_col].T                      # [branch][layer]
        ref = pg[p + 'k']
        scale = np.abs(ref).max(axis=1, keepdims=True)
        r_k = float(np.max(np.where(scale > 0, np.abs(kk - ref) / np.where(scale > 0, scale, 1), np.abs(kk - ref))))
        ch = float(st.cpu().numpy().view(_lib.COLSTATE_DTYPE)['aflux_change'][0])
        print('call %d: tau %.1e, %s, k %.1e, aflux_change %.6e vs %.6e'
              % (call, r_tau, ', '.join('%s %.1e' % kv for kv in out.items()), r_k, ch,
                 float(pg[p + 'aflux_change'])))
        for name, v in out.items():
            assert v < 1e-10, (call, name, v)
        assert r_k < 1e-10, (call, r_k)
        assert abs(ch - float(pg[p + 'aflux_change'])) <= 1e-9 * float(pg[p + 'aflux_change'])
        # no other rate is touched
        other = np.ones(net.nr, bool)
        other[t.br_col] = False
        assert not k.cpu().numpy()[:, other].any()
    ph.close()


@gpu
def test_photo_schedule_in_the_device_loop(pg):
    """Attached to the device loop, the update runs on the first attempt of
    every step whose count is a multiple of ini_update_photo_frq (count 0
    included), never again within that step, and every final frequency's
    steps once longdy < 10 yconv_min and longdydt < 1e-6 (op.py:818-830)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from _golden_util import column_from
    from vulcan_amd import photo as P
    from vulcan_amd.config import Config
    from vulcan_amd.device import ColumnBatch, DevicePlan
    d = load()
    net, col = network_from(d), column_from(d)
    t = _tables_from_golden(pg, list(net.species))
    cfg = Config(dttry=1e-9)
    b = ColumnBatch(DevicePlan(net, alpha=d['alpha']), [col], [d['y0']], cfg, k=d['k'], tune=False)
    ph = P.DevicePhoto(b.plan, b.layout, t)
    b.attach_photo(ph, 3, 2)
    cols = list(t.br_col)
    seen, first_visit = [], {}
    prev = b.k.cpu().numpy()[:, cols].copy()
    for attempt in range(40):
        count0 = int(b.host_state()['count'][0])
        first_visit.setdefault(count0, attempt)
        if attempt == 25:
            st = b.host_state()
            st['longdy'], st['longdydt'] = 0.5, 1e-7    # below 10 yconv_min = 1 and 1e-6
            b.set_host_state(st)
        b.integrate(1, use_conv=False)
        kk = b.k.cpu().numpy()[:, cols]
        if not np.array_equal(kk, prev):
            seen.append((attempt, count0))
        prev = kk.copy()
    print('photo updates at (attempt, count):', seen)
    counts = [c for _, c in seen]
    assert len(set(counts)) == len(counts)                   # once per step at most
    assert all(first_visit[c] == a for a, c in seen)         # on the step's first attempt
    before = [c for c, a in first_visit.items() if a < 25]
    after = [c for c, a in first_visit.items() if a > 25]
    assert sorted(c for a, c in seen if a < 25) == sorted(c for c in before if c % 3 == 0)
    assert sorted(c for a, c in seen if a > 25) == sorted(c for c in after if c % 2 == 0)
    assert counts[0] == 0 and any(c % 2 == 0 and c % 3 != 0 for c in counts)
    b.close()
    ph.close()


@gpu
def test_ros2_compute_tau_flux_J_on_vulcan_objects(pg):
    """vulcan_amd.op.Ros2.compute_tau / compute_flux / compute_J on the
    tables VULCAN's set-up leaves in var (make_bins_read_cross, read_sflux):
    var.tau, the fluxes, var.aflux_change and var.k at pho_rate_index as the
    reference's calls produce them (twice: the second from the first's
    state)."""
    import torch
    import types
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    from vulcan_amd.config import Config
    from vulcan_amd.op import Ros2
    d = load()
    net = network_from(d)
    species = list(net.species)
    photo_sp = [str(s) for s in pg['photo_sp']]
    Tsp = [str(s) for s in pg['T_cross_sp']]
    br = [(str(s), int(n)) for s, n in zip(pg['branch_sp'], pg['branch_no'])]
    nz = pg['Tco'].shape[0]
    var = types.SimpleNamespace(
        bins=pg['bins'], sflux_top=pg['sflux_top'], sflux_din12_indx=int(pg['din12']), dbin1=float(pg['dbin1']),
        dbin2=float(pg['dbin2']), photo_sp=set(photo_sp),
        cross={sp: pg['cross'][q] for q, sp in enumerate(photo_sp)},
        cross_T={sp: pg['cross_T'][q] for q, sp in enumerate(Tsp)},
        cross_scat={str(sp): pg['cross_scat'][q] for q, sp in enumerate(pg['scat_sp'])},
        cross_J={b: pg['cross_J'][q] for q, b in enumerate(br)},
        cross_J_T={b: pg['cross_J_T'][q] for q, b in enumerate(br) if b[0] in Tsp},
        pho_rate_index={b: int(i) for b, i in zip(br, pg['branch_index'])},
        k={j + 1: d['k'][:, j].copy() for j in range(net.nr)}, dt=1e-10)
    atm = types.SimpleNamespace(Tco=pg['Tco'], M=pg['M'], dz=pg['dz'], Kzz=d['Kzz'], Dzz=d['Dzz'], ms=d['ms'],
                                g=d['g0'], Hp=d['Hp0'], dzi=d['dzi0'], Ti=d['Ti0'], Hpi=d['Hpi0'], alpha=d['alpha'])
    cfg = Config(use_photo=True, T_cross_sp=Tsp, scat_sp=[str(s) for s in pg['scat_sp']])
    solver = Ros2(net, cfg)
    for call in range(2):
        p = 'c%d_' % call
        var.y, var.ymix = pg[p + 'y'].copy(), pg[p + 'ymix'].copy()
        solver.compute_tau(var, atm)
        solver.compute_flux(var, atm)
        solver.compute_J(var, atm)
        assert _rel_rows(var.tau, pg[p + 'tau']) < 1e-13
        assert _rel_rows(var.aflux, pg[p + 'aflux']) < 1e-10
        assert abs(var.aflux_change - float(pg[p + 'aflux_change'])) <= 1e-9 * float(pg[p + 'aflux_change'])
        for q, b in enumerate(br):
            ref = pg[p + 'k'][q]
            got = var.k[var.pho_rate_index[b]]
            assert np.max(np.abs(got - ref)) <= 1e-10 * np.abs(ref).max(), b
            assert np.allclose(var.J_sp[b], pg[p + 'J'][q], rtol=1e-10, atol=0)
