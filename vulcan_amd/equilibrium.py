"""Chemical-equilibrium initial state for a network (host-side setup).

The reference starts from FastChem equilibrium abundances (build_atm.py:78-164,
ini_mix = 'EQ'); FastChem does not know synthetic species, so this module
computes the ideal-gas equilibrium of any network from its own NASA9 data:
minimise G/RT = sum_j n_j (g_j/RT + ln(P/P0) + ln(n_j/n)) subject to the
element totals, with the Gordon-McBride iteration of NASA CEA
(NASA RP-1311, chapter 2: element potentials pi_i and Delta ln n), damped
as there (|Delta ln n_j| <= 2 per iteration).  Vectorised over layers.
"""
from __future__ import annotations

import numpy as np

from .network import Network


def g_RT(T, nasa9):
    """(nz, ni) g/RT from NASA9 (low set below 1000 K, high at and above)."""
    T = np.asarray(T, dtype=np.float64)[:, None]

    def hs(a):
        a = [a[:, q][None, :] for q in range(10)]
        h = -a[0] / T ** 2 + a[1] * np.log(T) / T + a[2] + a[3] * T / 2. + a[4] * T ** 2 / 3. \
            + a[5] * T ** 3 / 4. + a[6] * T ** 4 / 5. + a[8] / T
        s = -a[0] / T ** 2 / 2. - a[1] / T + a[2] * np.log(T) + a[3] * T + a[4] * T ** 2 / 2. \
            + a[5] * T ** 3 / 3. + a[6] * T ** 4 / 4. + a[9]
        return h - s
    lo, hi = hs(nasa9[:, :10]), hs(nasa9[:, 10:20])
    return np.where(T < 1000, lo, hi)


def equilibrium(net: Network, T, P, b, iters: int = 200, tol: float = 1e-11) -> np.ndarray:
    """Equilibrium mole fractions (nz, ni).

    T (nz,) K; P (nz,) dyn/cm2; b (E,) or (nz, E) element amounts for the
    network's ``elements`` (relative scale is irrelevant)."""
    A = net.compo.astype(np.float64)          # (ni, E)
    ni, E = A.shape
    T = np.asarray(T, dtype=np.float64)
    nz = T.shape[0]
    b = np.broadcast_to(np.asarray(b, dtype=np.float64), (nz, E)).copy()
    present = b.max(axis=0) > 0
    # species made of absent elements are excluded
    ok_sp = np.all((A[:, ~present] == 0), axis=1)
    A = A[:, present]
    b = b[:, present]
    b = b / b.sum(axis=1, keepdims=True)
    E = A.shape[1]
    mu0 = g_RT(T, net.nasa9)[:, ok_sp] + np.log(np.asarray(P, dtype=np.float64) / 1e6)[:, None]
    As = A[ok_sp]
    m = As.shape[0]
    lnn = np.full((nz, m), np.log(0.1 / m))
    lnN = np.full(nz, np.log(0.1))
    for it in range(iters):
        n = np.exp(lnn)
        N = np.exp(lnN)
        mu = mu0 + lnn - lnN[:, None]                              # mu_j/RT
        # matrix (E+1)x(E+1) per layer
        G = np.zeros((nz, E + 1, E + 1))
        G[:, :E, :E] = np.einsum('zj,jk,ji->zki', n, As, As)
        G[:, :E, E] = n @ As
        G[:, E, :E] = n @ As
        G[:, E, E] = n.sum(axis=1) - N
        rhs = np.zeros((nz, E + 1))
        rhs[:, :E] = b - n @ As + np.einsum('zj,jk,zj->zk', n, As, mu)
        rhs[:, E] = N - n.sum(axis=1) + (n * mu).sum(axis=1)
        sol = np.linalg.solve(G, rhs[..., None])[..., 0]
        pi, dlnN = sol[:, :E], sol[:, E]
        dlnn = -mu + pi @ As.T + dlnN[:, None]
        # damping (CEA: limit the largest change of the major species)
        big = np.maximum(np.abs(dlnn).max(axis=1), np.abs(dlnN) * 5)
        lam = np.minimum(1.0, 2.0 / np.maximum(big, 1e-300))
        lnn = lnn + lam[:, None] * dlnn
        lnN = lnN + lam * dlnN
        lnn = np.maximum(lnn, -745.0)
        if np.max(np.abs(lam[:, None] * dlnn) * (n > 1e-12 * N[:, None])) < tol and it > 3:
            break
    x = np.zeros((nz, ni))
    x[:, ok_sp] = np.exp(lnn - lnN[:, None])
    return x / x.sum(axis=1, keepdims=True)
