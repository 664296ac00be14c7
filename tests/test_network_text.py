"""The VULCAN network-text reader against ReadRate's forward rates.

The reference's ReadRate.read_rate output for thermo/CHO_thermo_network.txt
is held in tests/golden/ref_readrate_CHO_thermo.npz; the network file
itself is read from /root/reference when present (development container).
"""
import os

import numpy as np
import pytest

from _golden_util import load
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
