// Device helpers shared by the kernels.  Expression trees follow the
// reference's numpy expressions operation by operation (the library is
// compiled with -ffp-contract=off), so that the results agree with the
// reference to the last bit wherever the inputs do.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace vk {

constexpr double KB = 1.38064852E-16;        // phy_const.py:3
constexpr double NAVO = 6.02214086E+23;      // phy_const.py:4
constexpr double NAVO_KB = NAVO * KB;        // evaluated left to right as Navo*kb*Ti
constexpr double CORR = KB / 1.e6;           // thermo/gibbs_text.txt:1
constexpr double R_ROS = 1.7071067811865475; // 1. + 1./2.**0.5, op.py:2890

// numpy's pairwise summation of a contiguous row (loops_utils.h.src):
// np.sum(y, axis=1) == 0.0 + pairwise(row)
__device__ inline double pw_block(const double *a, int n) {
    if (n < 8) {
        double r = 0.0;
        for (int i = 0; i < n; i++) r += a[i];
        return r;
    }
    double r0 = a[0], r1 = a[1], r2 = a[2], r3 = a[3], r4 = a[4], r5 = a[5], r6 = a[6], r7 = a[7];
    int i;
    for (i = 8; i < n - (n % 8); i += 8) {
        r0 += a[i + 0]; r1 += a[i + 1]; r2 += a[i + 2]; r3 += a[i + 3];
        r4 += a[i + 4]; r5 += a[i + 5]; r6 += a[i + 6]; r7 += a[i + 7];
    }
    double res = ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
    for (; i < n; i++) res += a[i];
    return res;
}

__device__ inline double pw_sum(const double *a, int n) {
    if (n <= 128) return 0.0 + pw_block(a, n);
    // one or two levels of halving cover every n <= 512
    int n2 = n / 2;
    n2 -= n2 % 8;
    double lo, hi;
    if (n2 <= 128) lo = pw_block(a, n2);
    else { int m = n2 / 2; m -= m % 8; lo = pw_block(a, m) + pw_block(a + m, n2 - m); }
    int r = n - n2;
    if (r <= 128) hi = pw_block(a + n2, r);
    else { int m = r / 2; m -= m % 8; hi = pw_block(a + n2, m) + pw_block(a + n2 + m, r - m); }
    return 0.0 + (lo + hi);
}

// NASA9 h/RT and s/R (thermo/gibbs_text.txt:13-18)
__device__ inline double h_RT(double T, const double *a) {
    return -a[0] / (T * T) + a[1] * log(T) / T + a[2] + a[3] * T / 2. + a[4] * (T * T) / 3.
         + a[5] * pow(T, 3.0) / 4. + a[6] * pow(T, 4.0) / 5. + a[8] / T;
}
__device__ inline double s_R(double T, const double *a) {
    return -a[0] / (T * T) / 2. - a[1] / T + a[2] * log(T) + a[3] * T + a[4] * (T * T) / 2.
         + a[5] * pow(T, 3.0) / 3. + a[6] * pow(T, 4.0) / 4. + a[9];
}
// g/RT: low set below 1000 K, high set at and above (chem_funs.py:2789-2792)
__device__ inline double g_RT(double T, const double *a20) {
    double glo = h_RT(T, a20) - s_R(T, a20);
    double ghi = h_RT(T, a20 + 10) - s_R(T, a20 + 10);
    double mlo = (T < 1000.0) ? 1.0 : 0.0;
    double mhi = (T >= 1000.0) ? 1.0 : 0.0;
    return mlo * glo + mhi * ghi;
}

// Column-local view of the atmosphere arrays.
struct ColAtm {
    const double *Kzz, *dzi, *Ti, *Hpi, *Tco, *g, *Dzz;
    const double *vz;                        // interfaces, NULL: no advection
    const double *vm;                        // [nz][ni] layers, NULL: drift form of the molecular terms
    const double *top_flux, *bot_flux, *bot_vdep;   // [ni] of the column (when switched on)
    int nz, ni;
    bool use_top, use_bot;
};

__device__ inline ColAtm col_atm(const vk_atm &a, int c, int off, int nz, int ni) {
    ColAtm r;
    r.Kzz = a.Kzz + off; r.dzi = a.dzi + off; r.Ti = a.Ti + off; r.Hpi = a.Hpi + off;
    r.Tco = a.Tco + off; r.g = a.g + off; r.Dzz = a.Dzz + (size_t)off * ni;
    r.vz = a.vz ? a.vz + off : nullptr;
    r.vm = a.vm ? a.vm + (size_t)off * ni : nullptr;
    r.use_top = a.use_topflux != 0 && a.top_flux;
    r.use_bot = a.use_botflux != 0 && a.bot_flux && a.bot_vdep;
    r.top_flux = r.use_top ? a.top_flux + (size_t)c * ni : nullptr;
    r.bot_flux = r.use_bot ? a.bot_flux + (size_t)c * ni : nullptr;
    r.bot_vdep = r.use_bot ? a.bot_vdep + (size_t)c * ni : nullptr;
    r.nz = nz; r.ni = ni;
    return r;
}

// any species set in the diff_esc mask
__device__ inline bool esc_any(const vk_atm &a) {
    unsigned m = 0;
    for (int w = 0; w < VK_SPECIES_WORDS; w++) m |= a.esc[w];
    return m != 0;
}

// (v > 0) * v and (v < 0) * v: the upwind split of a velocity as the
// reference writes it (op.py:1545-1548), signed zeros included
__device__ inline double vpos(double v) { return (v > 0.0 ? 1.0 : 0.0) * v; }
__device__ inline double vneg(double v) { return (v < 0.0 ? 1.0 : 0.0) * v; }

// -1./Hpi[j] + ms*g/(Navo*kb*Ti[j]) + alpha/Ti[j]*(Tco[j+1]-Tco[j])/dzi[j]  (op.py:1549-1580)
__device__ inline double drift(const ColAtm &a, int j, double gl, double ms, double alpha) {
    return -1. / a.Hpi[j] + ms * gl / (NAVO_KB * a.Ti[j]) + alpha / a.Ti[j] * (a.Tco[j + 1] - a.Tco[j]) / a.dzi[j];
}

// diffdf (op.py:1496-1597), diffdf_vm (op.py:1599-1694) or diffdf_no_mol
// (op.py:1438-1494) for species i of column-local layer j, without the flux
// boundary terms.  ysm/ys0/ysp are the row sums of layers j-1, j, j+1 and
// ym/y0/yp the species values there (unused ones may be anything).
__device__ inline double diff_core(const ColAtm &a, int j, int i, double ysm, double ys0, double ysp,
                                   double ym, double y0, double yp, double ms, double alpha, bool mol) {
    const int nz = a.nz, ni = a.ni;
    if (j == 0) {
        double dzi0 = a.dzi[0], K0 = a.Kzz[0];
        double A = -1. / (dzi0) * (K0 / dzi0) * (ysp + ys0) / 2. / ys0;
        double B = 1. / (dzi0) * (K0 / dzi0) * (ysp + ys0) / 2. / ysp;
        if (a.vz) {
            A = A + (-vpos(a.vz[0])) / dzi0;
            B = B + (-vneg(a.vz[0])) / dzi0;
        }
        if (!mol) return A * y0 + B * yp;
        double Dz = a.Dzz[i];
        double Ai, Bi;
        if (a.vm) {
            Ai = -1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / 2. / ys0 - vpos(a.vm[i]) / dzi0;
            Bi = 1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / 2. / ysp - vneg(a.vm[i]) / dzi0;
        } else {
            double d0 = drift(a, 0, a.g[0], ms, alpha);
            Ai = -1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / 2. / ys0 + 1. / (dzi0) * Dz / 2. * d0;
            Bi = 1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / 2. / ysp + 1. / (dzi0) * Dz / 2. * d0;
        }
        return (A + Ai) * y0 + (B + Bi) * yp;
    }
    if (j == nz - 1) {
        double dzt = a.dzi[nz - 2], Kt = a.Kzz[nz - 2];
        double A = -1. / (dzt) * (Kt / dzt) * (ys0 + ysm) / 2. / ys0;
        double C = 1. / (dzt) * (Kt / dzt) * (ys0 + ysm) / 2. / ysm;
        if (a.vz) {
            A = A + vneg(a.vz[nz - 2]) / dzt;
            C = C + vpos(a.vz[nz - 2]) / dzt;
        }
        if (!mol) return A * y0 + C * ym;
        double Dz = a.Dzz[(size_t)(nz - 2) * ni + i];
        double Ai, Ci;
        if (a.vm) {
            // the top layer's own row vm[nz-1] (op.py:1649-1653)
            const double v = a.vm[(size_t)(nz - 1) * ni + i];
            Ai = -1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / 2. / ys0 + vneg(v) / dzt;
            Ci = 1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / 2. / ysm + vpos(v) / dzt;
        } else {
            double dT = drift(a, nz - 2, a.g[nz - 1], ms, alpha);
            Ai = -1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / 2. / ys0 - 1. / (dzt) * Dz / 2. * dT;
            Ci = 1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / 2. / ysm - 1. / (dzt) * Dz / 2. * dT;
        }
        return (A + Ai) * y0 + (C + Ci) * ym;
    }
    double dzm = a.dzi[j - 1], dz0 = a.dzi[j];
    double Km = a.Kzz[j - 1], K0 = a.Kzz[j];
    double da = 0.5 * (dzm + dz0);
    double A = -1. / da * (K0 / dz0 * (ysp + ys0) / 2. + Km / dzm * (ys0 + ysm) / 2.) / ys0;
    double B = 1. / da * K0 / dz0 * (ysp + ys0) / 2. / ysp;
    double C = 1. / da * Km / dzm * (ys0 + ysm) / 2. / ysm;
    if (a.vz) {
        const double v0 = a.vz[j], vm1 = a.vz[j - 1];
        A = A + (-(vpos(v0) - vneg(vm1))) / da;
        B = B + (-vneg(v0)) / da;
        C = C + vpos(vm1) / da;
    }
    double t = A * y0 + B * yp + C * ym;
    if (!mol) return t;
    double Dm = a.Dzz[(size_t)(j - 1) * ni + i], D0 = a.Dzz[(size_t)j * ni + i];
    double Ai = -1. / da * (D0 / dz0 * (ysp + ys0) / 2. + Dm / dzm * (ys0 + ysm) / 2.) / ys0;
    double Bi = 1. / da * D0 / dz0 * (ysp + ys0) / 2. / ysp;
    double Ci = 1. / da * Dm / dzm * (ys0 + ysm) / 2. / ysm;
    if (a.vm) {
        // layer rows j and j-1 of vm at interfaces j and j-1 (op.py:1666-1670)
        const double w0 = a.vm[(size_t)j * ni + i], wm = a.vm[(size_t)(j - 1) * ni + i];
        Ai = Ai + (-(vpos(w0) - vneg(wm))) / da;
        Bi = Bi + (-vneg(w0)) / da;
        Ci = Ci + (+vpos(wm)) / da;
    } else {
        Ai = Ai + 1. / (2. * da) * (D0 * drift(a, j, a.g[j], ms, alpha) - Dm * drift(a, j - 1, a.g[j], ms, alpha));
        Bi = Bi + 1. / (2. * da) * D0 * drift(a, j, a.g[j + 1], ms, alpha);
        Ci = Ci + -1. / (2. * da) * Dm * drift(a, j - 1, a.g[j - 1], ms, alpha);
    }
    return t + (Ai * y0 + Bi * yp + Ci * ym);
}

// the diffusion tendency with the flux boundary terms added last
// (op.py:1589-1596): top_flux / dzi[-1] on the top layer and
// (bot_flux - y bot_vdep) / dzi[0] on the bottom one
__device__ inline double diff_value(const ColAtm &a, int j, int i, double ysm, double ys0, double ysp,
                                    double ym, double y0, double yp, double ms, double alpha, bool mol) {
    double d = diff_core(a, j, i, ysm, ys0, ysp, ym, y0, yp, ms, alpha, mol);
    if (j == a.nz - 1 && a.use_top) d = d + a.top_flux[i] / a.dzi[a.nz - 2];
    if (j == 0 && a.use_bot) d = d + (a.bot_flux[i] - y0 * a.bot_vdep[i]) / a.dzi[0];
    return d;
}

// lhs_jac_tot / lhs_jac_tot_vm / lhs_jac_no_mol diffusion terms for species
// i of layer j (op.py:1973-2166): diag -= d0, d1; U = (0 - u0) - u1;
// L = (0 - l0) - l1.  Terms that do not exist are returned as 0 with has_* =
// false so that the caller applies exactly the reference's operations.  The
// deposition velocity (dep, bottom layer) and the escape term are applied by
// the caller in the reference's order.
struct LhsDiff {
    double d0, d1, u0, u1, l0, l1;
    bool has_u, has_l, has_mol;
};

__device__ inline LhsDiff lhs_diff(const ColAtm &a, int j, int i, double ysm, double ys0, double ysp,
                                   double ms, double alpha, bool mol) {
    const int nz = a.nz, ni = a.ni;
    LhsDiff r;
    r.d1 = r.u1 = r.l1 = 0.0;
    r.u0 = r.l0 = 0.0;
    r.has_mol = mol;
    if (j == 0) {
        double dzi0 = a.dzi[0], K0 = a.Kzz[0];
        r.has_u = true; r.has_l = false;
        r.d0 = -1. / (dzi0) * (K0 / dzi0) * (ysp + ys0) / (2. * ys0);
        r.u0 = 1. / (dzi0) * (K0 / dzi0) * (ysp + ys0) / (2. * ysp);
        if (a.vz) {
            r.d0 = r.d0 - vpos(a.vz[0]) / dzi0;
            r.u0 = r.u0 - vneg(a.vz[0]) / dzi0;
        }
        if (mol) {
            double Dz = a.Dzz[i];
            if (a.vm) {
                r.d1 = -1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / (2. * ys0) - vpos(a.vm[i]) / dzi0;
                r.u1 = 1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / (2. * ysp) - vneg(a.vm[i]) / dzi0;
            } else {
                double d0 = drift(a, 0, a.g[0], ms, alpha);
                r.d1 = -1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / (2. * ys0) + 1. / (dzi0) * Dz / 2. * d0;
                r.u1 = 1. / (dzi0) * (Dz / dzi0) * (ysp + ys0) / (2. * ysp) + 1. / (dzi0) * Dz / 2. * d0;
            }
        }
        return r;
    }
    if (j == nz - 1) {
        double dzt = a.dzi[nz - 2], Kt = a.Kzz[nz - 2];
        r.has_u = false; r.has_l = true;
        r.d0 = -1. / (dzt) * (Kt / dzt) * (ysm + ys0) / (2. * ys0);
        r.l0 = 1. / (dzt) * (Kt / dzt) * (ysm + ys0) / (2. * ysm);
        if (a.vz) {
            r.d0 = r.d0 + vneg(a.vz[nz - 2]) / dzt;
            r.l0 = r.l0 + vpos(a.vz[nz - 2]) / dzt;
        }
        if (mol) {
            double Dz = a.Dzz[(size_t)(nz - 2) * ni + i];
            if (a.vm) {
                const double v = a.vm[(size_t)(nz - 1) * ni + i];
                r.d1 = -1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / (2. * ys0) + vneg(v) / dzt;
                r.l1 = 1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / (2. * ysm) + vpos(v) / dzt;
            } else {
                double dT = drift(a, nz - 2, a.g[nz - 1], ms, alpha);
                r.d1 = -1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / (2. * ys0) - 1. / (dzt) * Dz / 2. * dT;
                r.l1 = 1. / (dzt) * (Dz / dzt) * (ys0 + ysm) / (2. * ysm) - 1. / (dzt) * Dz / 2. * dT;
            }
        }
        return r;
    }
    r.has_u = true; r.has_l = true;
    double dzm = a.dzi[j - 1], dz0 = a.dzi[j];
    double Km = a.Kzz[j - 1], K0 = a.Kzz[j];
    double da = 0.5 * (dzm + dz0);
    r.d0 = -1. / da * (K0 / dz0 * (ysp + ys0) / 2. + Km / dzm * (ysm + ys0) / 2.) / ys0;
    r.u0 = 1. / da * (K0 / dz0 * (ysp + ys0) / (2. * ysp));
    r.l0 = 1. / da * (Km / dzm * (ysm + ys0) / (2. * ysm));
    if (a.vz) {
        const double v0 = a.vz[j], vm1 = a.vz[j - 1];
        r.d0 = r.d0 - (vpos(v0) - vneg(vm1)) / da;
        r.u0 = r.u0 - vneg(v0) / da;
        r.l0 = r.l0 + vpos(vm1) / da;
    }
    if (mol) {
        double Dm = a.Dzz[(size_t)(j - 1) * ni + i], D0 = a.Dzz[(size_t)j * ni + i];
        if (a.vm) {
            const double w0 = a.vm[(size_t)j * ni + i], wm = a.vm[(size_t)(j - 1) * ni + i];
            r.d1 = -1. / da * (D0 / dz0 * (ysp + ys0) / 2. + Dm / dzm * (ysm + ys0) / 2.) / ys0
                 - (vpos(w0) - vneg(wm)) / da;
            r.u1 = 1. / da * (D0 / dz0 * (ysp + ys0) / (2. * ysp)) - vneg(w0) / da;
            r.l1 = 1. / da * (Dm / dzm * (ysm + ys0) / (2. * ysm)) + vpos(wm) / da;
        } else {
            r.d1 = -1. / da * (D0 / dz0 * (ysp + ys0) / 2. + Dm / dzm * (ysm + ys0) / 2.) / ys0
                 + 1. / (2. * da) * (D0 * drift(a, j, a.g[j], ms, alpha) - Dm * drift(a, j - 1, a.g[j], ms, alpha));
            r.u1 = 1. / da * (D0 / dz0 * (ysp + ys0) / (2. * ysp)) + 1. / (2. * da) * D0 * drift(a, j, a.g[j + 1], ms, alpha);
            r.l1 = 1. / da * (Dm / dzm * (ysm + ys0) / (2. * ysm)) - 1. / (2. * da) * Dm * drift(a, j - 1, a.g[j - 1], ms, alpha);
        }
    }
    return r;
}

}  // namespace vk
