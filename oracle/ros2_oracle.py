"""CPU oracle for VULCAN's Ros2 kinetics-plus-diffusion step.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``vulcan_amd`` package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, as the checker and as the timed CPU baseline.

It restates the reference algorithm in numpy, keeping the reference's
expression trees (operation order) wherever that is cheap, so that the device
kernels can be compared nearly bit-for-bit:

* rate constants                      op.py:145-207, 287-309; chem_funs.py:2769-2796
* ``chemdf``                          make_chem_funs.py:199-316 (generated chem_funs.py:931-1000)
* analytic Jacobian (``symjac``)      make_chem_funs.py:653-717
* ``diffdf`` / ``diffdf_vm`` / ``diffdf_no_mol``   op.py:1496-1597 / 1599-1694 / 1438-1494,
  with upwind vertical advection and the top/bottom flux terms
* ``lhs_jac_tot`` / ``lhs_jac_tot_vm`` / ``lhs_jac_no_mol``   op.py:1973-2042 / 2044-2120 / 2122-2166,
  with the deposition velocity and the diffusion-limited escape term
* fixed bottom mixing ratios, delta without layer 0   op.py:1435-1436, 2945-2953
* ``update_phi_esc``                  op.py:986-999
* ``store_bandM`` + two ``solve_banded`` calls   op.py:2833-2858, 2914, 2929
* ``Ros2.solver``                     op.py:2860-3007
* ``clip``/``loss``/``step_ok``/``step_reject``/``reset_y``  op.py:2447-2534
* ``one_step`` / ``step_size``        op.py:3091-3125
* ``Integration`` loop, ``f_dy``, ``conv``, ``stop``, ``save_step``,
  ``update_mu_dz``                    op.py:808-1105, 944-984

Parity pins for each part are listed in tests/test_oracle.py.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional

import numpy as np
import scipy.linalg

from vulcan_amd.config import Config, kb, Navo
from vulcan_amd import network as nw

R_ROS = 1. + 1. / 2. ** 0.5       # op.py:2890
CORR = kb / 1.e6                   # gibbs_text.txt:1 (P0 = 1 bar)


# ---------------------------------------------------------------------------
# A1: rate constants
# ---------------------------------------------------------------------------

def h_RT(T, a):
    """thermo/gibbs_text.txt:13-14 (chem_funs.py:2781-2782)."""
    return -a[0] / T ** 2 + a[1] * np.log(T) / T + a[2] + a[3] * T / 2. + a[4] * T ** 2 / 3. \
        + a[5] * T ** 3 / 4. + a[6] * T ** 4 / 5. + a[8] / T


def s_R(T, a):
    """thermo/gibbs_text.txt:17-18."""
    return -a[0] / T ** 2 / 2. - a[1] / T + a[2] * np.log(T) + a[3] * T + a[4] * T ** 2 / 2. \
        + a[5] * T ** 3 / 3. + a[6] * T ** 4 / 4. + a[9]


def g_RT(T, nasa9_row):
    """Low set below 1000 K, high set at and above (chem_funs.py:2789-2792)."""
    lo, hi = nasa9_row[:10], nasa9_row[10:20]
    return (T < 1000) * (h_RT(T, lo) - s_R(T, lo)) + (T >= 1000) * (h_RT(T, hi) - s_R(T, hi))


def gibbs_K(net: nw.Network, p: int, T):
    """K_eq of pair p (make_chem_funs.py:568-580 as emitted at chem_funs.py:2810ff).

    exp(-(sum_prod g - sum_reac g)) * (corr*T)**(n_reac - n_prod), summed
    left to right as the generated string does: reactants with '-1*', then
    products with '+1*'; M is excluded from the counts.
    """
    acc = None
    nre = npr = 0
    for t in net.reac[p]:
        if t >= 0:
            term = -1 * g_RT(T, net.nasa9[t])
            acc = term if acc is None else acc + term
            nre += 1
    for t in net.prod[p]:
        if t >= 0:
            term = 1 * g_RT(T, net.nasa9[t])
            acc = term if acc is None else acc + term
            npr += 1
    K = np.exp(-(acc))
    if npr - nre != 0:
        K = K * (CORR * T) ** (nre - npr)
    return K


def forward_rate(net: nw.Network, p: int, T, M, tau=None):
    kind = int(net.kind[p])
    par = net.par[p]
    A, n, E = par[nw.P_A], par[nw.P_N], par[nw.P_E]
    if kind in (nw.KIND_ARRH, nw.KIND_K0):
        return A * T ** n * np.exp(-E / T)                                   # op.py:166
    if kind in (nw.KIND_LIND, nw.KIND_TROE):
        k0 = A * T ** n * np.exp(-E / T)
        kinf = par[nw.P_AINF] * T ** par[nw.P_NINF] * np.exp(-par[nw.P_EINF] / T)   # op.py:179
        k = k0 / (1 + k0 * M / kinf)                                         # op.py:181
        if kind == nw.KIND_LIND:
            return k
        Fc = par[nw.P_FC]
        nn = 0.75 - 1.27 * np.log(Fc)                                        # op.py:204
        ff = np.exp(np.log(Fc) / (1. + (np.log(k0 * M / kinf) / nn) ** 2))   # op.py:205
        return k0 / (1 + k0 * M / kinf) * ff                                 # op.py:207
    if kind == nw.KIND_PHOTO:
        return par[nw.P_J0] * np.exp(-(0.0 if tau is None else tau)) * np.ones_like(T)
    if kind == nw.KIND_CH3OH:                                                # op.py:197-207
        k0 = 1.932E3 * T ** -9.88 * np.exp(-7544. / T) + 5.109E-11 * T ** -6.25 * np.exp(-1433. / T)
        kinf = 1.031E-10 * T ** -0.018 * np.exp(16.74 / T)
        Fc = 0.1855 * np.exp(-T / 155.8) + 0.8145 * np.exp(-T / 1675.) + np.exp(-4531. / T)
        nn = 0.75 - 1.27 * np.log(Fc)
        ff = np.exp(np.log(Fc) / (1. + (np.log(k0 * M / kinf) / nn) ** 2))
        return k0 / (1 + k0 * M / kinf) * ff
    raise ValueError(kind)


def rates(net: nw.Network, T, M, tau=None) -> np.ndarray:
    """k[l, 2p] forward, k[l, 2p+1] reverse = k_f / K_eq or 0 (op.py:287-309)."""
    nz = T.shape[0]
    k = np.zeros((nz, net.nr))
    for p in range(net.npairs):
        kf = forward_rate(net, p, T, M, tau)
        k[:, 2 * p] = kf
        if net.rev[p]:
            k[:, 2 * p + 1] = kf / gibbs_K(net, p, T)
    return k


# ---------------------------------------------------------------------------
# A2/A3: chemistry right-hand side and Jacobian
# ---------------------------------------------------------------------------

def _side_product(terms, kcol, y, M):
    acc = kcol
    for t in terms:
        if t == nw.TERM_NONE:
            continue
        acc = acc * (M if t == nw.TERM_M else y[:, t])
    return acc


def pair_rates(net, y, M, k):
    """v_p per layer, evaluated left to right (make_chem_funs.py:293-315)."""
    v = np.empty((y.shape[0], net.npairs))
    for p in range(net.npairs):
        v[:, p] = _side_product(net.reac[p], k[:, 2 * p], y, M) \
            - _side_product(net.prod[p], k[:, 2 * p + 1], y, M)
    return v


def chemdf(net, y, M, k, tables=None) -> np.ndarray:
    """dydt (nz, ni): -v per reactant occurrence, +v per product occurrence,
    summed in reaction order (chem_funs.py:931-1000)."""
    tables = tables or nw.build_tables(net)
    v = pair_rates(net, y, M, k)
    out = np.zeros_like(y)
    for i in range(net.ni):
        lo, hi = tables.chem_ptr[i], tables.chem_ptr[i + 1]
        if lo == hi:
            continue
        acc = tables.chem_sign[lo] * v[:, tables.chem_pair[lo]]
        for e in range(lo + 1, hi):
            acc = acc + tables.chem_sign[e] * v[:, tables.chem_pair[e]]
        out[:, i] = acc
    return out


def _dside(terms, kcol, y, M, m):
    """d/dy_m of kcol * prod(terms)."""
    tot = None
    for q, t in enumerate(terms):
        if t != m:
            continue
        acc = kcol
        for q2, t2 in enumerate(terms):
            if q2 == q or t2 == nw.TERM_NONE:
                continue
            acc = acc * (M if t2 == nw.TERM_M else y[:, t2])
        tot = acc if tot is None else tot + acc
    return 0.0 if tot is None else tot


def jac_blocks(net, y, M, k, tables=None) -> np.ndarray:
    """J[l, i, m] = d chemdf[l, i] / d y[l, m] (make_chem_funs.py:653-684)."""
    tables = tables or nw.build_tables(net)
    nz, ni = y.shape
    dv = np.empty((nz, tables.nslot))
    for s in range(tables.nslot):
        p, m = tables.slot_pair[s], tables.slot_sp[s]
        dv[:, s] = _dside(net.reac[p], k[:, 2 * p], y, M, m) - _dside(net.prod[p], k[:, 2 * p + 1], y, M, m)
    J = np.zeros((nz, ni, ni))
    for e_dense in np.nonzero(tables.jac_map >= 0)[0]:
        e = tables.jac_map[e_dense]
        i, m = divmod(int(e_dense), ni)
        lo, hi = tables.jac_ptr[e], tables.jac_ptr[e + 1]
        acc = tables.jac_coef[lo] * dv[:, tables.jac_slot[lo]]
        for q in range(lo + 1, hi):
            acc = acc + tables.jac_coef[q] * dv[:, tables.jac_slot[q]]
        J[:, i, m] = acc
    return J


def neg_symjac_dense(J: np.ndarray) -> np.ndarray:
    """Dense (ni*nz)^2 -J with index l*ni+i, one fancy-index write per (i, m)
    as the generated neg_symjac does (chem_funs.py:7456-7461)."""
    nz, ni, _ = J.shape
    dfdy = np.zeros((ni * nz, ni * nz))
    indx = [np.arange(j, j + ni * nz, ni) for j in range(ni)]
    for i in range(ni):
        for m in range(ni):
            dfdy[indx[i], indx[m]] = -(J[:, i, m])
    return dfdy


# ---------------------------------------------------------------------------
# A4: diffusion right-hand side
# ---------------------------------------------------------------------------

def _drift(col, j_int, gl):
    """-1/Hpi + ms g/(Navo kb Ti) + alpha/Ti (T[j+1]-T[j])/dzi at interface j_int with gravity gl."""
    Ti, Hpi, Tco, dzi = col.Ti, col.Hpi, col.Tco, col.dzi
    return -1. / Hpi[j_int] + col.ms * gl / (Navo * kb * Ti[j_int]) \
        + col.alpha / Ti[j_int] * (Tco[j_int + 1] - Tco[j_int]) / dzi[j_int]


def _vp(v):
    """(v > 0) * v: the upwind split of an advection velocity (op.py:1545-1548)."""
    return (v > 0) * v


def _vn(v):
    """(v < 0) * v."""
    return (v < 0) * v


def _bc(cfg):
    """(use_vm_mol, use_topflux, use_botflux) of a configuration (None: all off)."""
    if cfg is None:
        return False, False, False
    return bool(cfg.use_vm_mol), bool(cfg.use_topflux), bool(cfg.use_botflux)


def diffdf(col, y, use_moldiff=True, cfg=None) -> np.ndarray:
    """op.py:1496-1597 (use_moldiff), op.py:1599-1694 (use_moldiff and
    cfg.use_vm_mol) or op.py:1438-1494.  Vertical advection uses col.vz
    (zeros unless use_vz); cfg switches the constant top flux and the bottom
    flux with deposition velocity (col.top_flux, col.bot_flux, col.bot_vdep)."""
    use_vm, use_top, use_bot = _bc(cfg)
    nz, ni = y.shape
    ysum = np.sum(y, axis=1)
    dzi, Kzz, Dzz, g, vz = col.dzi, col.Kzz, col.Dzz, col.g, col.vz
    A, B, C = np.zeros(nz), np.zeros(nz), np.zeros(nz)
    A[0] = -1. / (dzi[0]) * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[0]
    B[0] = 1. / (dzi[0]) * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[1]
    A[nz - 1] = -1. / (dzi[nz - 2]) * (Kzz[nz - 2] / dzi[nz - 2]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 1]
    C[nz - 1] = 1. / (dzi[nz - 2]) * (Kzz[nz - 2] / dzi[nz - 2]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 2]
    A[0] += -_vp(vz[0]) / dzi[0]
    B[0] += -_vn(vz[0]) / dzi[0]
    A[-1] += _vn(vz[-1]) / dzi[-1]
    C[-1] += _vp(vz[-1]) / dzi[-1]
    j = np.arange(1, nz - 1)
    dz_ave = 0.5 * (dzi[j - 1] + dzi[j])
    A[j] = -1. / dz_ave * (Kzz[j] / dzi[j] * (ysum[j + 1] + ysum[j]) / 2. + Kzz[j - 1] / dzi[j - 1] * (ysum[j] + ysum[j - 1]) / 2.) / ysum[j]
    B[j] = 1. / dz_ave * Kzz[j] / dzi[j] * (ysum[j + 1] + ysum[j]) / 2. / ysum[j + 1]
    C[j] = 1. / dz_ave * Kzz[j - 1] / dzi[j - 1] * (ysum[j] + ysum[j - 1]) / 2. / ysum[j - 1]
    A[j] += -(_vp(vz[j]) - _vn(vz[j - 1])) / dz_ave
    B[j] += -_vn(vz[j]) / dz_ave
    C[j] += _vp(vz[j - 1]) / dz_ave
    diff = np.empty_like(y)
    if not use_moldiff:
        diff[0] = A[0] * y[0] + B[0] * y[1]
        diff[1:nz - 1] = A[j, None] * y[j] + B[j, None] * y[j + 1] + C[j, None] * y[j - 1]
        diff[nz - 1] = A[nz - 1] * y[nz - 1] + C[nz - 1] * y[nz - 2]
        return _flux_bc(col, y, diff, use_top, use_bot)
    Ai = np.zeros((nz, ni)); Bi = np.zeros((nz, ni)); Ci = np.zeros((nz, ni))
    if use_vm:
        vm = col.vm
        Ai[0] = -1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[0] - _vp(vm[0]) / dzi[0]
        Bi[0] = 1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[1] - _vn(vm[0]) / dzi[0]
        Ai[nz - 1] = -1. / (dzi[-1]) * (Dzz[nz - 2] / dzi[-1]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 1] \
            + _vn(vm[-1]) / dzi[-1]
        Ci[nz - 1] = 1. / (dzi[-1]) * (Dzz[nz - 2] / dzi[-1]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 2] \
            + _vp(vm[-1]) / dzi[-1]
    else:
        d0 = _drift(col, 0, g[0])
        Ai[0] = -1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[0] + 1. / (dzi[0]) * Dzz[0] / 2. * d0
        Bi[0] = 1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / 2. / ysum[1] + 1. / (dzi[0]) * Dzz[0] / 2. * d0
        dt_ = _drift(col, nz - 2, g[nz - 1])
        Ai[nz - 1] = -1. / (dzi[-1]) * (Dzz[nz - 2] / dzi[-1]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 1] \
            - 1. / (dzi[-1]) * Dzz[-1] / 2. * dt_
        Ci[nz - 1] = 1. / (dzi[-1]) * (Dzz[nz - 2] / dzi[-1]) * (ysum[nz - 1] + ysum[nz - 2]) / 2. / ysum[nz - 2] \
            - 1. / (dzi[-1]) * Dzz[-1] / 2. * dt_
    for jj in range(1, nz - 1):
        da = 0.5 * (dzi[jj - 1] + dzi[jj])
        Ai[jj] = -1. / da * (Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / 2. + Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj] + ysum[jj - 1]) / 2.) / ysum[jj]
        Bi[jj] = 1. / da * Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / 2. / ysum[jj + 1]
        Ci[jj] = 1. / da * Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj] + ysum[jj - 1]) / 2. / ysum[jj - 1]
        if use_vm:
            # upwind molecular velocity on the layer rows of vm (op.py:1666-1670)
            Ai[jj] += -(_vp(vm[jj]) - _vn(vm[jj - 1])) / da
            Bi[jj] += -_vn(vm[jj]) / da
            Ci[jj] += +_vp(vm[jj - 1]) / da
        else:
            Ai[jj] += 1. / (2. * da) * (Dzz[jj] * _drift(col, jj, g[jj]) - Dzz[jj - 1] * _drift(col, jj - 1, g[jj]))
            Bi[jj] += 1. / (2. * da) * Dzz[jj] * _drift(col, jj, g[jj + 1])
            Ci[jj] += -1. / (2. * da) * Dzz[jj - 1] * _drift(col, jj - 1, g[jj - 1])
    diff[0] = (A[0] + Ai[0]) * y[0] + (B[0] + Bi[0]) * y[1]
    mid = A[j, None] * y[j] + B[j, None] * y[j + 1] + C[j, None] * y[j - 1]
    mid += Ai[j] * y[j] + Bi[j] * y[j + 1] + Ci[j] * y[j - 1]
    diff[1:nz - 1] = mid
    diff[nz - 1] = (A[nz - 1] + Ai[nz - 1]) * y[nz - 1] + (C[nz - 1] + Ci[nz - 1]) * y[nz - 2]
    return _flux_bc(col, y, diff, use_top, use_bot)


def _flux_bc(col, y, diff, use_top, use_bot):
    """Constant top flux and bottom flux minus deposition (op.py:1589-1596)."""
    dzi = col.dzi
    if use_top:
        diff[-1] += col.top_flux / dzi[-1]
    if use_bot:
        diff[0] += (col.bot_flux - y[0] * col.bot_vdep) / dzi[0]
    return diff


# ---------------------------------------------------------------------------
# A5: Ros2 left-hand side as blocks
# ---------------------------------------------------------------------------

def lhs_blocks(col, y, dt, J, use_moldiff=True, cfg=None, esc=None):
    """lhs = c0 I - J_chem - J_diff (op.py:1973-2042, 2044-2120 with
    cfg.use_vm_mol, 2122-2166 without molecular diffusion).

    Returns D (nz, ni, ni) diagonal blocks, Lsub (nz, ni) with Lsub[j] the
    diagonal of block (j, j-1) and Usup (nz, ni) with Usup[j] the diagonal
    of block (j, j+1); Lsub[0] = Usup[nz-1] = 0.  cfg.use_botflux adds the
    deposition velocity col.bot_vdep to the bottom diagonal; with the _vm
    operators, the species indices ``esc`` (diff_esc) take
    top_flux / y on the top diagonal.
    """
    use_vm, _, use_bot = _bc(cfg)
    use_vm = use_vm and use_moldiff
    nz, ni = y.shape
    ysum = np.sum(y, axis=1)
    dzi, Kzz, Dzz, g, vz = col.dzi, col.Kzz, col.Dzz, col.g, col.vz
    vm = col.vm if use_vm else None
    c0 = 1. / (R_ROS * dt)
    D = -J.copy()
    idx = np.arange(ni)
    D[:, idx, idx] = c0 + D[:, idx, idx]                  # np.fill_diagonal(dfdy, c0 + diag)
    L = np.zeros((nz, ni)); U = np.zeros((nz, ni))
    for jj in range(1, nz - 1):
        da = 0.5 * (dzi[jj - 1] + dzi[jj])
        D[jj, idx, idx] -= -1. / da * (Kzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / 2. + Kzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / 2.) / ysum[jj] \
            - (_vp(vz[jj]) - _vn(vz[jj - 1])) / da
        U[jj] -= 1. / da * (Kzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / (2. * ysum[jj + 1])) - _vn(vz[jj]) / da
        L[jj] -= 1. / da * (Kzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / (2. * ysum[jj - 1])) + _vp(vz[jj - 1]) / da
        if use_vm:
            D[jj, idx, idx] -= -1. / da * (Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / 2. + Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / 2.) / ysum[jj] \
                - (_vp(vm[jj]) - _vn(vm[jj - 1])) / da
            U[jj] -= 1. / da * (Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / (2. * ysum[jj + 1])) - _vn(vm[jj]) / da
            L[jj] -= 1. / da * (Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / (2. * ysum[jj - 1])) + _vp(vm[jj - 1]) / da
        elif use_moldiff:
            D[jj, idx, idx] -= -1. / da * (Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / 2. + Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / 2.) / ysum[jj] \
                + 1. / (2. * da) * (Dzz[jj] * _drift(col, jj, g[jj]) - Dzz[jj - 1] * _drift(col, jj - 1, g[jj]))
            U[jj] -= 1. / da * (Dzz[jj] / dzi[jj] * (ysum[jj + 1] + ysum[jj]) / (2. * ysum[jj + 1])) \
                + 1. / (2. * da) * Dzz[jj] * _drift(col, jj, g[jj + 1])
            L[jj] -= 1. / da * (Dzz[jj - 1] / dzi[jj - 1] * (ysum[jj - 1] + ysum[jj]) / (2. * ysum[jj - 1])) \
                - 1. / (2. * da) * Dzz[jj - 1] * _drift(col, jj - 1, g[jj - 1])
    D[0, idx, idx] -= -1. / (dzi[0]) * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[0]) - _vp(vz[0]) / dzi[0]
    if use_vm:
        D[0, idx, idx] -= -1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[0]) - _vp(vm[0]) / dzi[0]
    elif use_moldiff:
        d0 = _drift(col, 0, g[0])
        D[0, idx, idx] -= -1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[0]) \
            + 1. / (dzi[0]) * Dzz[0] / 2. * d0
    if use_bot:
        D[0, idx, idx] -= -1. * col.bot_vdep / dzi[0]
    if use_vm and esc is not None and len(esc):
        # diffusion-limited escape in the Jacobian (op.py:2101-2107)
        diff_lim = np.zeros(ni)
        for i in esc:
            if y[-1, i] > 0:
                diff_lim[i] += col.top_flux[i] / y[-1, i]
        D[nz - 1, idx, idx] -= diff_lim
    U[0] -= 1. / (dzi[0]) * (Kzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[1]) - _vn(vz[0]) / dzi[0]
    if use_vm:
        U[0] -= 1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[1]) - _vn(vm[0]) / dzi[0]
    elif use_moldiff:
        U[0] -= 1. / (dzi[0]) * (Dzz[0] / dzi[0]) * (ysum[1] + ysum[0]) / (2. * ysum[1]) \
            + 1. / (dzi[0]) * Dzz[0] / 2. * d0
    t = nz - 1
    D[t, idx, idx] -= -1. / (dzi[nz - 2]) * (Kzz[nz - 2] / dzi[nz - 2]) * (ysum[t - 1] + ysum[t]) / (2. * ysum[t]) \
        + _vn(vz[-1]) / dzi[-1]
    if use_vm:
        D[t, idx, idx] -= -1. / (dzi[nz - 2]) * (Dzz[nz - 2] / dzi[nz - 2]) * (ysum[t] + ysum[nz - 2]) / (2. * ysum[t]) \
            + _vn(vm[-1]) / dzi[-1]
    elif use_moldiff:
        dtop = _drift(col, nz - 2, g[t])
        D[t, idx, idx] -= -1. / (dzi[nz - 2]) * (Dzz[nz - 2] / dzi[nz - 2]) * (ysum[t] + ysum[nz - 2]) / (2. * ysum[t]) \
            - 1. / (dzi[-1]) * Dzz[-1] / 2. * dtop
    L[t] -= 1. / (dzi[nz - 2]) * (Kzz[nz - 2] / dzi[nz - 2]) * (ysum[t - 1] + ysum[t]) / (2. * ysum[t - 1]) \
        + _vp(vz[-1]) / dzi[-1]
    if use_vm:
        L[t] -= 1. / (dzi[nz - 2]) * (Dzz[nz - 2] / dzi[nz - 2]) * (ysum[t] + ysum[nz - 2]) / (2. * ysum[t - 1]) \
            + _vp(vm[-1]) / dzi[-1]
    elif use_moldiff:
        L[t] -= 1. / (dzi[nz - 2]) * (Dzz[nz - 2] / dzi[nz - 2]) * (ysum[t] + ysum[nz - 2]) / (2. * ysum[t - 1]) \
            - 1. / (dzi[-1]) * Dzz[-1] / 2. * dtop
    return D, L, U


def lhs_dense(D, L, U):
    """Dense (ni*nz)^2 matrix from blocks (small sizes and the CPU baseline)."""
    nz, ni, _ = D.shape
    N = ni * nz
    a = np.zeros((N, N))
    for j in range(nz):
        s = slice(j * ni, (j + 1) * ni)
        a[s, s] = D[j]
        if j > 0:
            a[np.arange(j * ni, (j + 1) * ni), np.arange((j - 1) * ni, j * ni)] = L[j]
        if j < nz - 1:
            a[np.arange(j * ni, (j + 1) * ni), np.arange((j + 1) * ni, (j + 2) * ni)] = U[j]
    return a


# ---------------------------------------------------------------------------
# A6/A7: band storage and the banded solve
# ---------------------------------------------------------------------------

def store_bandM(a, nb, nn):
    """op.py:2833-2858: dense -> LAPACK diagonal-ordered band storage."""
    bw = 2 * nb - 1
    ab = np.zeros((2 * bw + 1, nb * nn))
    for i in range(0, 2 * nb):
        ab[-(2 * nb + i):, i] = a[0:2 * nb + i, i]
    for i in range(2 * nb, nn * nb - 2 * nb):
        ab[:, i] = a[(i - 2 * nb + 1):(i - 2 * nb + 1) + (2 * bw + 1), i]
    for ne, i in enumerate(range(nn * nb - 2 * nb, nn * nb)):
        ab[:(2 * bw + 1 - ne), i] = a[-(2 * bw + 1 - ne):, i]
    return ab, bw


def band_from_blocks(D, L, U):
    """The same band storage filled straight from the blocks: ab[bw+i-j, j] = a[i, j]."""
    nz, ni, _ = D.shape
    bw = 2 * ni - 1
    N = ni * nz
    ab = np.zeros((2 * bw + 1, N))
    rows = np.arange(ni)
    for j in range(nz):
        c = j * ni + rows                      # global column index
        for i in range(ni):
            r = j * ni + i
            ab[bw + r - c, c] = D[j, i, :]
        if j > 0:
            r = j * ni + rows
            cc = (j - 1) * ni + rows
            ab[bw + r - cc, cc] = L[j]
        if j < nz - 1:
            r = j * ni + rows
            cc = (j + 1) * ni + rows
            ab[bw + r - cc, cc] = U[j]
    return ab, bw


# ---------------------------------------------------------------------------
# A8-A11: solver state machine and integration loop
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class State:
    y: np.ndarray
    ymix: np.ndarray
    k: np.ndarray
    dt: float
    atom_ini: np.ndarray
    t: float = 0.0
    count: int = 0
    delta: float = 0.0
    y_prev: Optional[np.ndarray] = None
    atom_sum: Optional[np.ndarray] = None
    atom_loss: Optional[np.ndarray] = None
    atom_loss_prev: Optional[np.ndarray] = None
    small_y: float = 0.0
    nega_y: float = 0.0
    nega_count: int = 0
    loss_count: int = 0
    delta_count: int = 0
    dy: float = 1.0
    dydt: float = 1.0
    longdy: float = 1.0
    longdydt: float = 1.0
    aflux_change: float = 0.0
    end_case: int = 0
    y_time: List[np.ndarray] = dataclasses.field(default_factory=list)
    t_time: List[float] = dataclasses.field(default_factory=list)
    k1: Optional[np.ndarray] = None
    k2: Optional[np.ndarray] = None
    attempts: int = 0
    exact_berr: float = 0.0     # backward error of the refined solves (exact=True)


def atom_sums(y, atom_mat):
    """loss(): sum_i compo[i, a] * y[:, i] over all layers (op.py:2472-2487)."""
    ni = y.shape[1]
    return np.array([np.sum([atom_mat[i, a] * y[:, i] for i in range(ni)])
                     for a in range(atom_mat.shape[1])])


def new_state(net, y, k, cfg: Config) -> State:
    """Initial state: ymix from y, element baseline atom_ini (build_atm.py:283-291)."""
    atom_mat = net.atom_matrix()
    st = State(y=y.copy(), ymix=y / np.sum(y, axis=1)[:, None], k=k, dt=cfg.dttry,
               atom_ini=atom_sums(y, atom_mat))
    st.atom_sum = st.atom_ini.copy()
    st.atom_loss = np.zeros_like(st.atom_ini)
    st.atom_loss_prev = np.zeros_like(st.atom_ini)
    st.y_prev = st.y.copy()
    return st


def _residual_ld(Dl, Ll, Ul, bl, x):
    """b - A x of the block-tridiagonal system, in extended precision."""
    r = bl - np.einsum('lij,lj->li', Dl, x)
    r[1:] -= Ll[1:] * x[:-1]
    r[:-1] -= Ul[:-1] * x[1:]
    return r


def backward_error(D, L, U, rhs, x) -> float:
    """Componentwise backward error max_i |b - A x|_i / (|A| |x| + |b|)_i of
    x for the block-tridiagonal system (D, L, U), in extended precision."""
    nz, ni = D.shape[:2]
    Dl, Ll, Ul = D.astype(np.longdouble), L.astype(np.longdouble), U.astype(np.longdouble)
    bl = np.asarray(rhs).reshape(nz, ni).astype(np.longdouble)
    xl = np.asarray(x).reshape(nz, ni).astype(np.longdouble)
    r = np.abs(_residual_ld(Dl, Ll, Ul, bl, xl))
    ax = np.einsum('lij,lj->li', np.abs(Dl), np.abs(xl)) + np.abs(bl)
    ax[1:] += np.abs(Ll[1:] * xl[:-1])
    ax[:-1] += np.abs(Ul[:-1] * xl[1:])
    with np.errstate(invalid='ignore', divide='ignore'):
        q = np.where(ax > 0, r / ax, 0.0)
    return float(np.max(q))


def solve_refined(D, L, U, lhs_b, bw, rhs, iters=3):
    """The banded LAPACK solve followed by `iters` steps of iterative
    refinement with the residual b - A x formed blockwise in extended
    precision: the linear solve to (nearly) full FP64 accuracy.  Not part of
    the reference; it gives the exact Ros2 step against which the
    reference's own banded solve and the device solve are both measured.
    Returns (x, componentwise backward error of the returned FP64 x)."""
    nz, ni = D.shape[:2]
    Dl, Ll, Ul = D.astype(np.longdouble), L.astype(np.longdouble), U.astype(np.longdouble)
    bl = rhs.reshape(nz, ni).astype(np.longdouble)
    x = scipy.linalg.solve_banded((bw, bw), lhs_b, rhs).reshape(nz, ni).astype(np.longdouble)
    for _ in range(iters):
        r = _residual_ld(Dl, Ll, Ul, bl, x)
        dx = scipy.linalg.solve_banded((bw, bw), lhs_b, r.astype(np.float64).ravel())
        x = x + dx.reshape(nz, ni).astype(np.longdouble)
    x64 = x.astype(np.float64).ravel()
    return x64, backward_error(D, L, U, rhs, x64)


def ros2_solver(net, col, st: State, cfg: Config, tables=None, dense_lhs=False, exact=False):
    """One Ros2 attempt (op.py:2860-3007).  Mutates st.y, st.ymix, st.delta.
    exact=True replaces the two banded solves by solve_refined."""
    tables = tables or nw.build_tables(net)
    y, ymix, h, k = st.y, st.ymix, st.dt, st.k
    M = col.M
    r = R_ROS
    esc = esc_indices(net.species, cfg)
    df = chemdf(net, y, M, k, tables).flatten() + diffdf(col, y, cfg.use_moldiff, cfg).flatten()
    J = jac_blocks(net, y, M, k, tables)
    D, L, U = lhs_blocks(col, y, h, J, cfg.use_moldiff, cfg, esc)
    if dense_lhs:
        # the reference's cost profile: a dense (ni*nz)^2 neg_symjac, then the
        # diagonal/diffusion updates and the column-loop band repack; the
        # values written are the blocks above, so both paths give equal bits
        lhs = neg_symjac_dense(J)
        nz_, ni_ = y.shape
        rows = np.arange(ni_)
        for j in range(nz_):
            s = slice(j * ni_, (j + 1) * ni_)
            lhs[s, s] = D[j]
            if j > 0:
                lhs[j * ni_ + rows, (j - 1) * ni_ + rows] = L[j]
            if j < nz_ - 1:
                lhs[j * ni_ + rows, (j + 1) * ni_ + rows] = U[j]
        lhs_b, bw = store_bandM(lhs, net.ni, y.shape[0])
        del lhs
    else:
        lhs_b, bw = band_from_blocks(D, L, U)
    st.exact_berr = 0.0

    def solve(rhs_):
        if not exact:
            return scipy.linalg.solve_banded((bw, bw), lhs_b, rhs_)
        x_, berr = solve_refined(D, L, U, lhs_b, bw, rhs_)
        st.exact_berr = max(st.exact_berr, berr)
        return x_
    k1_flat = solve(df)
    k1 = k1_flat.reshape(y.shape)
    yk2 = y + k1 / r
    df = chemdf(net, yk2, M, k, tables).flatten() + diffdf(col, yk2, cfg.use_moldiff, cfg).flatten()
    rhs = df - 2. / (r * h) * k1_flat
    k2 = solve(rhs).reshape(y.shape)
    sol = y + 3. / (2. * r) * k1 + 1 / (2. * r) * k2
    if cfg.use_fix_sp_bot:
        # fixed bottom mixing ratios (op.py:1435-1436, 2945-2946)
        fix_idx = [net.species.index(sp) for sp in cfg.use_fix_sp_bot.keys()]
        fix_mix = np.array([cfg.use_fix_sp_bot[sp] for sp in cfg.use_fix_sp_bot.keys()])
        sol[0, fix_idx] = fix_mix * col.M[0]
    delta = np.abs(sol - yk2)
    delta[ymix < cfg.mtol] = 0
    delta[sol < cfg.atol] = 0
    if cfg.use_botflux or cfg.use_fix_sp_bot:
        delta[0] = 0                     # op.py:2953
    pos = sol > 0
    delta = np.amax(delta[pos] / sol[pos])
    st.y = sol
    st.ymix = st.y / np.vstack(np.sum(st.y, axis=1))
    st.delta = delta
    st.k1, st.k2 = k1, k2
    st.attempts += 1
    return st


def clip(st: State, atom_mat, cfg: Config):
    """op.py:2447-2470 (loss at 2472-2487)."""
    y, ymix = st.y, st.ymix.copy()
    st.small_y += np.abs(np.sum(y[np.logical_and(y < cfg.pos_cut, y >= 0)]))
    st.nega_y += np.abs(np.sum(y[np.logical_and(y > cfg.nega_cut, y <= 0)]))
    y[np.logical_and(y < cfg.pos_cut, y >= cfg.nega_cut)] = 0.
    y[np.logical_and(ymix < cfg.mtol, y < 0)] = 0.
    st.atom_sum = atom_sums(y, atom_mat)
    st.atom_loss = (st.atom_sum - st.atom_ini) / st.atom_ini
    st.y, st.ymix = y, y / np.vstack(np.sum(y, axis=1))
    return st


def step_ok(st: State, cfg: Config) -> bool:
    """op.py:2489-2493 (rtol and loss_eps bound at import: cfg values)."""
    return bool(np.all(st.y >= 0) and np.amax(np.abs(st.atom_loss - st.atom_loss_prev)) < cfg.loss_eps
                and st.delta <= cfg.rtol)


def step_reject(st: State, cfg: Config) -> bool:
    """op.py:2495-2534.  Returns True when the solver gives up at dt < dt_min."""
    if st.delta > cfg.rtol:
        st.delta_count += 1
    elif np.any(st.y < 0):
        st.nega_count += 1
    else:
        st.loss_count += 1
    st.y = st.y_prev                     # reset_y: ymix stays stale (op.py:2524-2534)
    st.dt *= cfg.dt_var_min
    if st.dt < cfg.dt_min:
        st.dt = cfg.dt_min
        st.y[st.y < 0] = 0.
        return True
    return False


def one_step(net, col, st: State, cfg: Config, tables=None, max_attempts=10 ** 6, dense_lhs=False, exact=False):
    """op.py:3091-3103."""
    atom_mat = net.atom_matrix()
    for _ in range(max_attempts):
        ros2_solver(net, col, st, cfg, tables, dense_lhs=dense_lhs, exact=exact)
        clip(st, atom_mat, cfg)
        if step_ok(st, cfg):
            break
        if step_reject(st, cfg):
            break
    return st


def step_size(st: State, cfg: Config):
    """op.py:3105-3125."""
    h = st.dt
    delta = st.delta
    rtol = cfg.rtol
    if delta == 0:
        delta = 0.01 * rtol
    h_factor = 0.9 * (rtol / delta) ** 0.5
    h_factor = np.maximum(h_factor, cfg.dt_var_min)
    h_factor = np.minimum(h_factor, cfg.dt_var_max)
    h *= h_factor
    h = np.maximum(h, cfg.dt_min)
    h = np.minimum(h, cfg.dt_max)
    st.dt = float(h)
    return st


def update_mu_dz(col, st: State):
    """op.py:944-984 (mean_mass build_atm.py:515-520); mutates col."""
    nz = col.nz
    mu = np.zeros(nz)
    for i in range(st.ymix.shape[1]):
        mu += col.ms[i] * st.ymix[:, i]
    col.mu = mu
    Tco, pico = col.Tco, col.pico
    pref = col.pref_indx
    for i in range(pref, nz):
        if i == pref:
            col.g[i] = col.gs
            col.Hp[i] = kb * Tco[i] / (col.mu[i] / Navo * col.gs)
        else:
            col.g[i] = col.gs * (col.Rp / (col.Rp + col.zco[i])) ** 2
            col.Hp[i] = kb * Tco[i] / (col.mu[i] / Navo * col.g[i])
        col.dz[i] = col.Hp[i] * np.log(pico[i] / pico[i + 1])
        col.zco[i + 1] = col.zco[i] + col.dz[i]
    if pref != 0:
        for i in range(pref - 1, -1, -1):
            col.g[i] = col.gs * (col.Rp / (col.Rp + col.zco[i + 1])) ** 2
            col.Hp[i] = kb * Tco[i] / (col.mu[i] / Navo * col.g[i])
            col.dz[i] = col.Hp[i] * np.log(pico[i] / pico[i + 1])
            col.zco[i] = col.zco[i + 1] - col.dz[i]
    col.dzi = (0.5 * (col.dz + np.roll(col.dz, 1)))[1:]
    col.Ti = (0.5 * (Tco + np.roll(Tco, -1)))[:-1]
    col.Hpi = (0.5 * (col.Hp + np.roll(col.Hp, -1)))[:-1]
    return col


def esc_indices(species, cfg: Config):
    """Species of diff_esc whose escape changes the step: through the top
    flux of diffdf (use_topflux) or the _vm Jacobian (op.py:2101-2107)."""
    if not cfg.diff_esc or not (cfg.use_topflux or (cfg.use_vm_mol and cfg.use_moldiff)):
        return []
    return [list(species).index(sp) for sp in cfg.diff_esc]


def update_phi_esc(col, st: State, cfg: Config, species):
    """Diffusion-limited escape flux at the top (op.py:986-999); mutates
    col.top_flux."""
    for i in esc_indices(species, cfg):
        f = -col.Dzz[-1, i] * st.y[-1, i] * (1. / col.Hp[-1] - col.ms[i] * col.g[-1] / (Navo * kb * col.Tco[-1]))
        col.top_flux[i] = max(f, cfg.max_flux * (-1))
    return col


def f_dy(st: State, cfg: Config):
    """op.py:1003-1015."""
    if st.count == 0:
        st.dy, st.dydt = 1., 1.
        return st
    y, ymix, y_prev = st.y, st.ymix, st.y_prev
    dy = np.abs(y - y_prev)
    dy[ymix < cfg.mtol] = 0
    dy[y < cfg.atol] = 0
    dy = np.amax(dy[y > 0] / y[y > 0])
    st.dy, st.dydt = dy, dy / st.dt
    return st


def conv(col, st: State, cfg: Config, species=None) -> bool:
    """op.py:1018-1063 (aflux_change stays 0 without radiative transfer)."""
    y, ymix, y_time, t_time = st.y.copy(), st.ymix.copy(), st.y_time, np.asarray(st.t_time)
    count = st.count
    slope_min = min(np.amin(col.Kzz / (0.1 * col.Hp[:-1]) ** 2), 1.e-8)
    slope_min = max(slope_min, 1.e-10)
    indx = np.abs(t_time - st.t * cfg.st_factor).argmin()
    if indx == count - 1:
        indx -= 1
    indx = max(count - cfg.conv_step, indx)
    longdy = np.abs((y_time[count - 1] - y_time[indx]) / np.vstack(col.M))
    longdy[ymix < cfg.mtol_conv] = 0
    longdy[y < cfg.atol] = 0
    for sp in cfg.conver_ignore:
        longdy[:, species.index(sp)] = 0
    longdy = np.amax(longdy[ymix > 0] / ymix[ymix > 0])
    longdydt = longdy / (t_time[-1] - t_time[indx])
    st.longdy, st.longdydt = longdy, longdydt
    return bool((longdy < cfg.yconv_cri and longdydt < cfg.slope_cri or
                 longdy < cfg.yconv_min and longdydt < slope_min) and st.aflux_change < cfg.flux_cri)


def stop(col, st: State, cfg: Config, species=None) -> bool:
    """op.py:1065-1085."""
    if st.t > cfg.trun_min and st.count > cfg.count_min and conv(col, st, cfg, species):
        st.end_case = 1
        return True
    if st.t > cfg.runtime:
        st.end_case = 2
        return True
    if st.count > cfg.count_max:
        st.end_case = 3
        return True
    return False


def integrate(net, col, st: State, cfg: Config, max_steps=None, tables=None, update_mu=True,
              keep_history=True):
    """Integration.__call__ (op.py:808-936) without photo RT, condensation or plots.

    ``max_steps`` bounds the number of accepted steps taken in this call (the
    reference has no such knob; it stops only through ``stop``).
    """
    tables = tables or nw.build_tables(net)
    taken = 0
    while not stop(col, st, cfg, net.species):
        if max_steps is not None and taken >= max_steps:
            break
        st.y_prev = np.copy(st.y)                      # backup, op.py:938-942
        st.atom_loss_prev = st.atom_loss.copy()
        one_step(net, col, st, cfg, tables)
        if st.count % cfg.update_frq == 0:
            if update_mu:
                update_mu_dz(col, st)
            update_phi_esc(col, st, cfg, net.species)
        st.y = np.vstack(col.M) * st.ymix              # op.py:909-914
        f_dy(st, cfg)
        st.t += st.dt                                  # save_step, op.py:1087-1105
        st.count += 1
        if keep_history:
            st.y_time.append(st.y)
        st.t_time.append(st.t)
        step_size(st, cfg)
        taken += 1
    return st
