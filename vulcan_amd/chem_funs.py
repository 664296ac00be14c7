"""``chem_funs``-compatible chemistry functions backed by the HIP kernels.

The reference generates a ``chem_funs`` module per network
(make_chem_funs.py) and calls ``chemdf(y, M, k)``, ``symjac(y, M, k)``,
``neg_symjac(y, M, k)`` and ``Gibbs(i, T)`` on host numpy arrays, with ``k`` a
dict of 1-based rate rows (chem_funs.py:931, 3222, 7456, 2810).  ``ChemFuns``
offers the same names and argument meaning for a ``Network``; the arithmetic
runs on the device and the results come back as numpy arrays.
``as_module(net)`` returns an object that can stand in for the generated
module (``spec_list``, ``ni``, ``nr``, ``re_dict``, ``re_wM_dict`` and the
functions).

``symjac`` / ``neg_symjac`` return the dense (ni*nz)^2 matrix of the
reference only on request (``dense=True``, small sizes); by default they
return the nz diagonal ni x ni blocks, the only nonzero part.
"""
from __future__ import annotations

import types
from typing import Dict, Union

import numpy as np

from . import network as nw

KType = Union[np.ndarray, Dict[int, np.ndarray]]


def k_dict_to_array(k: KType, nz: int, nr: int) -> np.ndarray:
    """VULCAN's {1..nr: (nz,)} rate dict -> [nz][nr] (column j-1 holds k[j])."""
    if isinstance(k, np.ndarray):
        return np.ascontiguousarray(k, dtype=np.float64)
    out = np.zeros((nz, nr))
    for j in range(1, nr + 1):
        if j in k:
            out[:, j - 1] = np.broadcast_to(np.asarray(k[j], dtype=np.float64), (nz,))
    return out


def k_array_to_dict(k: np.ndarray) -> Dict[int, np.ndarray]:
    return {j + 1: k[:, j].copy() for j in range(k.shape[1])}


class ChemFuns:
    def __init__(self, net: nw.Network):
        self.net = net
        self.spec_list = list(net.species)
        self.ni = net.ni
        self.nr = net.nr
        self.re_dict, self.re_wM_dict = {}, {}
        for p in range(net.npairs):
            j = 2 * p + 1
            r_wm = ['M' if t == nw.TERM_M else net.species[t] for t in net.reac[p] if t != nw.TERM_NONE]
            p_wm = ['M' if t == nw.TERM_M else net.species[t] for t in net.prod[p] if t != nw.TERM_NONE]
            r_no = [s for s in r_wm if s != 'M']
            p_no = [s for s in p_wm if s != 'M']
            self.re_dict[j], self.re_dict[j + 1] = [r_no, p_no], [p_no, r_no]
            self.re_wM_dict[j], self.re_wM_dict[j + 1] = [r_wm, p_wm], [p_wm, r_wm]
        self._plan = None
        self._cols = {}

    # -- device objects, created on first use ---------------------------------
    def _device(self, nz: int):
        from .device import DeviceColumns, DevicePlan
        if self._plan is None:
            self._plan = DevicePlan(self.net)
        if nz not in self._cols:
            self._cols[nz] = DeviceColumns([nz])
        return self._plan, self._cols[nz]

    def _call(self, fn_name, y, M, k, out_shape):
        import torch
        from . import _lib
        from .device import dptr, stream_handle
        y = np.ascontiguousarray(y, dtype=np.float64)
        nz = y.shape[0]
        plan, cols = self._device(nz)
        dev = torch.device('cuda')
        yd = torch.as_tensor(y, device=dev)
        Md = torch.as_tensor(np.array(np.broadcast_to(M, (nz,)), dtype=np.float64), device=dev)
        kd = torch.as_tensor(k_dict_to_array(k, nz, self.nr), device=dev)
        out = torch.empty(out_shape, dtype=torch.float64, device=dev)
        lib = _lib.load()
        _lib.check(getattr(lib, fn_name)(plan.handle, cols.handle, dptr(yd), dptr(Md), dptr(kd), dptr(out),
                                         stream_handle()), fn_name)
        return out.cpu().numpy()

    # -- the generated module's functions ----------------------------------------
    def chemdf(self, y, M, k) -> np.ndarray:
        """dydt (nz, ni) (chem_funs.py:931-1000)."""
        return self._call('vk_chemdf', y, M, k, np.shape(y))

    def symjac(self, y, M, k, dense: bool = False) -> np.ndarray:
        """d chemdf / dy (chem_funs.py:3222-7453): (nz, ni, ni) blocks, or the
        dense (ni*nz)^2 matrix with index l*ni+i when ``dense``."""
        nz, ni = np.shape(y)
        J = self._call('vk_jac', y, M, k, (nz, ni, ni))
        return _dense(J) if dense else J

    def neg_symjac(self, y, M, k, dense: bool = False) -> np.ndarray:
        """-symjac (chem_funs.py:7456-11688)."""
        nz, ni = np.shape(y)
        J = -self._call('vk_jac', y, M, k, (nz, ni, ni))
        return _dense(J) if dense else J

    def Gibbs(self, i: int, T) -> np.ndarray:
        """K_eq of forward row i (odd, 1-based) at temperature(s) T (chem_funs.py:2810)."""
        import torch
        from . import _lib
        from .device import DeviceColumns, DevicePlan, dptr, stream_handle
        T = np.atleast_1d(np.asarray(T, dtype=np.float64))
        if i % 2 != 1:
            raise ValueError('Gibbs(i, T) takes a forward (odd) reaction index')
        p = (i - 1) // 2
        # K = k_f / k_r for a unit forward rate: evaluate one pair at every T
        sub = nw.Network(species=self.net.species, elements=self.net.elements, compo=self.net.compo,
                         mass=self.net.mass, nasa9=self.net.nasa9, reac=self.net.reac[p:p + 1],
                         prod=self.net.prod[p:p + 1], kind=np.array([nw.KIND_ARRH], np.int32),
                         par=np.array([[1.0, 0, 0, 0, 0, 0, 0, 0]]), rev=np.array([1], np.int32),
                         atom_list=self.net.atom_list)
        plan = DevicePlan(sub)
        nz = max(len(T), 2)
        Tp = np.resize(T, nz)
        cols = DeviceColumns([nz])
        dev = torch.device('cuda')
        Td = torch.as_tensor(Tp, device=dev)
        Md = torch.ones(nz, dtype=torch.float64, device=dev)
        k = torch.empty((nz, 2), dtype=torch.float64, device=dev)
        atm = _lib.Atm(Tco=dptr(Td), M=dptr(Md))
        import ctypes as C
        _lib.check(_lib.load().vk_rates(plan.handle, cols.handle, C.byref(atm), dptr(k), stream_handle()),
                   'vk_rates')
        kh = k.cpu().numpy()
        return (kh[:, 0] / kh[:, 1])[:len(T)]


def _dense(J: np.ndarray) -> np.ndarray:
    nz, ni, _ = J.shape
    out = np.zeros((ni * nz, ni * nz))
    for l in range(nz):
        out[l * ni:(l + 1) * ni, l * ni:(l + 1) * ni] = J[l]
    return out


def as_module(net: nw.Network) -> types.SimpleNamespace:
    """An object with the generated module's public names, for code written
    against ``import chem_funs``."""
    cf = ChemFuns(net)
    return types.SimpleNamespace(spec_list=cf.spec_list, ni=cf.ni, nr=cf.nr, re_dict=cf.re_dict,
                                 re_wM_dict=cf.re_wM_dict, chemdf=cf.chemdf, symjac=cf.symjac,
                                 neg_symjac=cf.neg_symjac, Gibbs=cf.Gibbs, impl=cf)
