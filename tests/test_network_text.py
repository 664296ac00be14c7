"""The VULCAN network-text reader against ReadRate's forward rates.

The reference's ReadRate.read_rate output for thermo/CHO_thermo_network.txt
and thermo/SNCHO_photo_network_2024.txt is held in
tests/golden/ref_readrate_*.npz; the network files themselves are read from
/root/reference when present (development container), and their parsed
device tables are committed (tests/golden/net_*.npz) for runs without it.
"""
import os

import numpy as np
import pytest

from _golden_util import load, network_fixture
from oracle import ros2_oracle as O
from vulcan_amd import network as nw
from vulcan_amd import network_text as NT

REF = '/root/reference'
NET = os.path.join(REF, 'thermo', 'CHO_thermo_network.txt')


@pytest.mark.skipif(not os.path.exists(NET), reason='reference network file not present')
def test_forward_rates_match_readrate():
    d = load('ref_readrate_CHO_thermo.npz')
    net = NT.load_network(NET, os.path.join(REF, 'thermo'))
    assert [int(x) for x in d['index']] == [int(x) for x in net.vulcan_index]
    assert [str(e) for e in d['eq']] == net.labels
    k = O.rates(net, d['Tco'], d['M'])
    kf = k[:, 0::2]
    ref = d['kf']
    nzm = ref != 0
    assert np.array_equal(kf == 0, ref == 0)
    assert np.max(np.abs(kf[nzm] - ref[nzm]) / np.abs(ref[nzm])) < 1e-14
    # the special CH3OH form is present and evaluated
    assert (net.kind == nw.KIND_CH3OH).sum() == 1


def test_parser_sections_and_numbering():
    text = """# Two-body Reactions
1  [ OH + H2 -> H2O + H ]  3.57E-16  1.520  1740.0  ref
3  [ 2*O -> O2 ]  1e-12  0  0
# 3-body and Disscoiation Reactions
5  [ H + H + M -> H2 + M ]  1e-33 -1 0  1e-10 0 0
# 3-body reactions without high-pressure rates
7  [ O + O + M -> O2 + M ]  1e-34 0 0
# reverse stops
# photo disscoiation
9  [ H2O -> OH + H ]  H2O  1
"""
    eqs, kinds, pars, rev, index = NT.parse_network(text)
    assert index == [1, 3, 5, 7, 9]
    assert kinds == [nw.KIND_ARRH, nw.KIND_ARRH, nw.KIND_LIND, nw.KIND_K0, nw.KIND_PHOTO]
    assert rev == [1, 1, 1, 1, 0]
    assert pars[2][3] == 1e-10
    assert NT._parse_equation('2*O -> O2') == (['O', 'O'], ['O2'])
    eqs2, _, _, _, idx2 = NT.parse_network(text, use_photo=False)
    assert idx2 == [1, 3, 5, 7] and len(eqs2) == 4


@pytest.mark.parametrize('fixture,golden', [('net_SNCHO_photo_network_2024.npz', 'ref_readrate_SNCHO_2024.npz'),
                                            ('net_CHO_thermo_network.npz', 'ref_readrate_CHO_thermo.npz')])
def test_network_tables_give_readrate_forward_rates(fixture, golden):
    """The committed tables of the real networks (Arrhenius, Lindemann,
    k0-only, the special CH3OH form, photolysis rows at zero) give the
    reference's forward rates."""
    d = load(golden)
    net = network_fixture(fixture)
    assert [int(x) for x in d['index']] == [int(x) for x in net.vulcan_index]
    assert [str(e) for e in d['eq']] == net.labels
    k = O.rates(net, d['Tco'], d['M'])
    kf = k[:, 0::2]
    ref = d['kf']
    nzm = ref != 0
    assert np.array_equal(kf == 0, ref == 0)
    assert np.max(np.abs(kf[nzm] - ref[nzm]) / np.abs(ref[nzm])) < 1e-14
    for kind in (nw.KIND_ARRH, nw.KIND_LIND, nw.KIND_K0, nw.KIND_CH3OH):
        assert (net.kind == kind).any(), kind


@pytest.mark.skipif(not os.path.exists(NET), reason='reference network files not present')
@pytest.mark.parametrize('name', ['SNCHO_photo_network_2024', 'CHO_thermo_network'])
def test_committed_tables_are_the_parse_of_the_network_file(name):
    net = NT.load_network(os.path.join(REF, 'thermo', name + '.txt'), os.path.join(REF, 'thermo'))
    fx = network_fixture('net_%s.npz' % name)
    assert net.species == fx.species and net.labels == fx.labels and net.atom_list == fx.atom_list
    for key in ('compo', 'mass', 'nasa9', 'reac', 'prod', 'kind', 'par', 'rev'):
        assert np.array_equal(getattr(net, key), getattr(fx, key)), key
