"""Device tables of two real VULCAN networks, for runs where the network text
files are not available (the GPU box).

    python tests/golden/make_network_tables.py

Reads, in the development container, the network text, the NASA9 files and
all_compose.txt from /root/reference/thermo with this repository's own reader
(vulcan_amd.network_text.load_network; nothing of the reference is imported
or run) and writes the resulting tables as data-only .npz files:

* ``net_SNCHO_photo_network_2024.npz``: the 2024 S-N-C-H-O photochemical
  network (Arrhenius, Lindemann three-body, k0-only three-body, the special
  CH3OH form and photolysis rows; 89 species, 575 pairs).  Its 2025
  successor names CH3CCH, which has neither a row in all_compose.txt nor a
  NASA9 file in the reference, so the reference cannot set it up either;
* ``net_CHO_thermo_network.npz``: the C-H-O thermochemical network.

The forward rates these tables give are pinned against the reference's
ReadRate output (tests/golden/make_golden.py writes it,
tests/test_network_text.py compares).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from vulcan_amd import network_text as NT  # noqa: E402

REF_THERMO = '/root/reference/thermo'
OUT = os.path.dirname(os.path.abspath(__file__))
NETWORKS = ['SNCHO_photo_network_2024.txt', 'CHO_thermo_network.txt']


def save(net, path):
    np.savez_compressed(path, species=np.array(net.species), elements=np.array(net.elements),
                        compo=net.compo, mass=net.mass, nasa9=net.nasa9, reac=net.reac, prod=net.prod,
                        kind=net.kind, par=net.par, rev=net.rev, labels=np.array(net.labels),
                        atom_list=np.array(net.atom_list), vulcan_index=np.asarray(net.vulcan_index))


def main():
    for name in NETWORKS:
        net = NT.load_network(os.path.join(REF_THERMO, name), REF_THERMO)
        out = os.path.join(OUT, 'net_' + name.replace('.txt', '.npz'))
        save(net, out)
        kinds = np.bincount(net.kind, minlength=6)
        print('%s: ni %d, %d pairs, kinds (arrh, lind, k0, troe, photo, ch3oh) %s -> %s'
              % (name, net.ni, net.npairs, kinds.tolist(), out))


if __name__ == '__main__':
    main()
