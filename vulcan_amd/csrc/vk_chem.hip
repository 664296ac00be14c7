// Rate constants, chemistry + diffusion right-hand side, Jacobian and Ros2
// left-hand-side assembly.  One workgroup per layer for the per-layer
// kernels: the layer's y row and its reaction rates live in LDS, species
// rows are gathered from per-species CSR lists (no atomics, fixed order).
#include <algorithm>

#include "vk_internal.h"
#include "vk_device.h"

using namespace vk;

// Bounds-checked build (compiled with VK_CHECKED, libvulcan_amd_checked.so):
// every index the gather/scatter kernels read from their tables is checked
// against the length of the array it indexes; the first index out of range
// records its site in vk_check_word and the access uses index 0 instead,
// so nothing is read or written outside its array.  The normal build
// compiles the checks away.
enum VkCheckSite {
    VK_SITE_SPECIES = 1,   // species of a reaction side (rates of pairs, slot derivatives)
    VK_SITE_PAIR = 2,      // pair of a production/loss entry of a species
    VK_SITE_CHEM = 3,      // production/loss entry of a species
    VK_SITE_SLOTPAIR = 4,  // pair of a derivative slot
    VK_SITE_TERM = 5,      // Jacobian term of an entry
    VK_SITE_SLOT = 6,      // derivative slot of a Jacobian term
    VK_SITE_DENSE = 7,     // dense index of a long Jacobian entry
};
#ifdef VK_CHECKED
static __device__ int vk_check_word;
__device__ __forceinline__ int vk_ix(int i, int n, int site) {
    if ((unsigned)i < (unsigned)n) return i;
    atomicCAS(&vk_check_word, 0, site);
    return 0;
}
#define VK_IX(i, n, site) vk_ix((i), (n), (site))
#else
#define VK_IX(i, n, site) (i)
#endif

namespace {

constexpr int BLK = 256;

// ---------------------------------------------------------------------------
// rates (op.py:145-207, 287-309; photolysis op.py:2786)
// ---------------------------------------------------------------------------

__global__ void k_grt(int L, int ni, const double *__restrict__ Tco, const double *__restrict__ nasa9,
                      double *__restrict__ grt) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= L * ni) return;
    int l = idx / ni, s = idx - l * ni;
    grt[idx] = g_RT(Tco[l], nasa9 + 20 * s);
}

__device__ inline double arrh(double A, double n, double E, double T) {
    return A * pow(T, n) * exp(-E / T);
}

__global__ void k_rates(int L, int npairs, int ni, const int *__restrict__ kind, const double *__restrict__ par,
                        const int *__restrict__ rev, const int *__restrict__ reac, const int *__restrict__ prod,
                        const double *__restrict__ Tco, const double *__restrict__ M, const double *__restrict__ tau,
                        const double *__restrict__ grt, double *__restrict__ k) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= L * npairs) return;
    int l = idx / npairs, p = idx - l * npairs;
    double T = Tco[l], m = M[l];
    const double *pa = par + 8 * p;
    double kf;
    switch (kind[p]) {
    case VK_KIND_ARRH:
    case VK_KIND_K0:
        kf = arrh(pa[0], pa[1], pa[2], T);
        break;
    case VK_KIND_LIND: {
        double k0 = arrh(pa[0], pa[1], pa[2], T);
        double kinf = arrh(pa[3], pa[4], pa[5], T);
        kf = k0 / (1 + k0 * m / kinf);
        break;
    }
    case VK_KIND_TROE: {
        double k0 = arrh(pa[0], pa[1], pa[2], T);
        double kinf = arrh(pa[3], pa[4], pa[5], T);
        double Fc = pa[6];
        double nn = 0.75 - 1.27 * log(Fc);
        double q = log(k0 * m / kinf) / nn;
        double ff = exp(log(Fc) / (1. + q * q));
        kf = k0 / (1 + k0 * m / kinf) * ff;
        break;
    }
    case VK_KIND_PHOTO:
        kf = pa[7] * exp(-(tau ? tau[l] : 0.0));
        break;
    case VK_KIND_CH3OH: {
        double k0 = 1.932E3 * pow(T, -9.88) * exp(-7544. / T) + 5.109E-11 * pow(T, -6.25) * exp(-1433. / T);
        double kinf = 1.031E-10 * pow(T, -0.018) * exp(16.74 / T);
        double Fc = 0.1855 * exp(-T / 155.8) + 0.8145 * exp(-T / 1675.) + exp(-4531. / T);
        double nn = 0.75 - 1.27 * log(Fc);
        double q = log(k0 * m / kinf) / nn;
        double ff = exp(log(Fc) / (1. + q * q));
        kf = k0 / (1 + k0 * m / kinf) * ff;
        break;
    }
    default:
        kf = 0.0;
    }
    double kr = 0.0;
    if (rev[p]) {
        // K = exp(-(-g_r1 - g_r2 + g_p1 + g_p2)) * (corr T)^(n_r - n_p)  (make_chem_funs.py:568-580)
        const double *g = grt + (size_t)l * ni;
        double acc = 0.0;
        bool first = true;
        int nre = 0, npr = 0;
        for (int q = 0; q < 4; q++) {
            int t = reac[4 * p + q];
            if (t >= 0) {
                t = VK_IX(t, ni, VK_SITE_SPECIES);
                acc = first ? -1 * g[t] : acc - 1 * g[t];
                first = false;
                nre++;
            }
        }
        for (int q = 0; q < 4; q++) {
            int t = prod[4 * p + q];
            if (t >= 0) {
                t = VK_IX(t, ni, VK_SITE_SPECIES);
                acc = first ? 1 * g[t] : acc + 1 * g[t];
                first = false;
                npr++;
            }
        }
        double K = exp(-(acc));
        int dn = nre - npr;
        if (dn != 0) {
            double x = CORR * T;
            double f;
            if (dn == 1) f = x;
            else if (dn == -1) f = 1.0 / x;
            else if (dn == 2) f = x * x;
            else f = pow(x, (double)dn);
            K = K * f;
        }
        kr = kf / K;
    }
    k[(size_t)l * 2 * npairs + 2 * p] = kf;
    k[(size_t)l * 2 * npairs + 2 * p + 1] = kr;
}

// ---------------------------------------------------------------------------
// per-layer helpers
// ---------------------------------------------------------------------------

struct LayerCtx {
    int l, c, off, nz, j;
};

__device__ inline LayerCtx layer_ctx(const int *off, const int *layer_col, int l) {
    LayerCtx x;
    x.l = l;
    x.c = layer_col[l];
    x.off = off[x.c];
    x.nz = off[x.c + 1] - x.off;
    x.j = l - x.off;
    return x;
}

// v_p = k_f * prod(reac) - k_r * prod(prod), left to right (make_chem_funs.py:293-315);
// the four terms of a side are one 16-byte load
__device__ inline double side_product(int4 t4, double kcol, const double *yl, double m, int ni) {
    const int t[4] = {t4.x, t4.y, t4.z, t4.w};
    double a = kcol;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (t[q] == -2) continue;
        a = a * (t[q] == -1 ? m : yl[VK_IX(t[q], ni, VK_SITE_SPECIES)]);
    }
    return a;
}

__device__ inline double pair_rate(const int4 *reac, const int4 *prod, const double *kl, const double *yl,
                                   double m, int p, int ni) {
    const double a = side_product(reac[p], kl[2 * p], yl, m, ni);
    const double b = side_product(prod[p], kl[2 * p + 1], yl, m, ni);
    return a - b;
}

// d/dy_sp of kcol * prod(terms)
__device__ inline double dside(int4 t4, double kcol, const double *yl, double m, int sp, bool &any, int ni) {
    const int t[4] = {t4.x, t4.y, t4.z, t4.w};
    double tot = 0.0;
    any = false;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (t[q] != sp) continue;
        double acc = kcol;
#pragma unroll
        for (int q2 = 0; q2 < 4; q2++) {
            if (q2 == q || t[q2] == -2) continue;
            acc = acc * (t[q2] == -1 ? m : yl[VK_IX(t[q2], ni, VK_SITE_SPECIES)]);
        }
        tot = any ? tot + acc : acc;
        any = true;
    }
    return tot;
}

// ysum[l] = sum_i y[l][i] in numpy's pairwise order, one thread per layer
__global__ void k_ysum(int L, int ni, const double *__restrict__ y, double *__restrict__ ysum) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < L) ysum[l] = pw_sum(y + (size_t)l * ni, ni);
}

// the three layer totals a layer's diffusion terms use (0 outside the column)
__device__ inline void ysum_around(const double *ysum, const LayerCtx &x, double &ym, double &y0, double &yp) {
    const double *ys = ysum + x.off;
    ym = x.j > 0 ? ys[x.j - 1] : 0.0;
    y0 = ys[x.j];
    yp = x.j < x.nz - 1 ? ys[x.j + 1] : 0.0;
}

// ---------------------------------------------------------------------------
// chemdf + diffdf (+ stage-2 term), one workgroup per layer
// mode bit 0: include chemistry, bit 1: include diffusion
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(BLK) k_rhs(
    int ni, int npairs, int nchem, const int *__restrict__ off, const int *__restrict__ layer_col,
    const int *__restrict__ reac, const int *__restrict__ prod, const int *__restrict__ chem_ptr,
    const int2 *__restrict__ chem_ent, const double *__restrict__ ms,
    const double *__restrict__ alpha, vk_atm atm, const double *__restrict__ y, const double *__restrict__ k,
    const double *__restrict__ M, int mode, int mol, const double *__restrict__ k1,
    const double *__restrict__ dtc, const double *__restrict__ ysum, const vk_colstate *__restrict__ st,
    double *__restrict__ out) {
    extern __shared__ double smem[];
    double *yl = smem;              // ni
    double *v = smem + VK_NI_MAX;   // npairs
    const int l = blockIdx.x;
    LayerCtx x = layer_ctx(off, layer_col, l);
    if (st && st[x.c].done) return;
    const double *yrow = y + (size_t)l * ni;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) yl[i] = yrow[i];
    double ysm = 0.0, ys0 = 0.0, ysp = 0.0;
    if (mode & 2) ysum_around(ysum, x, ysm, ys0, ysp);
    __syncthreads();
    if (mode & 1) {
        const double *kl = k + (size_t)l * 2 * npairs;
        double m = M[l];
        for (int p = threadIdx.x; p < npairs; p += blockDim.x)
            v[p] = pair_rate((const int4 *)reac, (const int4 *)prod, kl, yl, m, p, ni);
        __syncthreads();
    }
    ColAtm ca = col_atm(atm, x.c, x.off, x.nz, ni);
    double c2 = 0.0;
    if (k1) c2 = 2. / (R_ROS * dtc[x.c]);
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        double chem = 0.0;
        if (mode & 1) {
            const int lo = chem_ptr[i], hi = chem_ptr[i + 1];
            if (lo < hi) {
                int2 t = chem_ent[VK_IX(lo, nchem, VK_SITE_CHEM)];   // (pair, sign)
                chem = t.y * v[VK_IX(t.x, npairs, VK_SITE_PAIR)];
                // the terms in order, eight entries loaded ahead of their sums
                int e = lo + 1;
                for (; e + 8 <= hi; e += 8) {
                    int2 tt[8];
#pragma unroll
                    for (int u = 0; u < 8; u++) tt[u] = chem_ent[VK_IX(e + u, nchem, VK_SITE_CHEM)];
#pragma unroll
                    for (int u = 0; u < 8; u++) chem = chem + tt[u].y * v[VK_IX(tt[u].x, npairs, VK_SITE_PAIR)];
                }
                for (; e < hi; e++) {
                    t = chem_ent[VK_IX(e, nchem, VK_SITE_CHEM)];
                    chem = chem + t.y * v[VK_IX(t.x, npairs, VK_SITE_PAIR)];
                }
            }
        }
        double val = chem;
        if (mode & 2) {
            double ym = x.j > 0 ? y[(size_t)(l - 1) * ni + i] : 0.0;
            double yp = x.j < x.nz - 1 ? y[(size_t)(l + 1) * ni + i] : 0.0;
            double d = diff_value(ca, x.j, i, ysm, ys0, ysp, ym, yl[i], yp, ms[i], alpha[i], mol != 0);
            val = (mode & 1) ? chem + d : d;
        }
        if (k1) val = val - c2 * k1[(size_t)l * ni + i];
        out[(size_t)l * ni + i] = val;
    }
}

// ---------------------------------------------------------------------------
// Jacobian blocks (mode 0: J = d chemdf/dy) or Ros2 lhs blocks (mode 1:
// D = c0 I - J - J_diff, Lsub, Usup), one workgroup per layer
// ---------------------------------------------------------------------------

// The diagonal of a layer's lhs block after c0 - J: the diffusion terms in
// the reference's order.  xmode 1: the escape term first (top layer, _vm
// operators, op.py:2101-2107); xmode 2: the deposition velocity last (bottom
// layer, op.py:2029).
__device__ __forceinline__ double lhs_diag(double val, int i, int mol, const double *dg0, const double *dg1,
                                           const double *dgx, int xmode) {
    if (xmode == 1) val = val - dgx[i];
    val = val - dg0[i];
    if (mol) val = val - dg1[i];
    if (xmode == 2) val = val - dgx[i];
    return val;
}

// one entry of the lhs (mode 1) or of the Jacobian (mode 0): its terms in order
__device__ __forceinline__ void jac_entry(int q, const int4 *__restrict__ jac_ent, const int *__restrict__ jac_tp,
                                          const double *dv, int ni, int njac, int nslot, int mode, int mol, double c0,
                                          const double *dg0, const double *dg1, const double *dgx, int xmode,
                                          double *__restrict__ Dl) {
    const int4 en = jac_ent[q];           // (dense index, first term, end of terms, row)
    double J = 0.0;
    if (en.z > en.y) {
        const int t = jac_tp[VK_IX(en.y, njac, VK_SITE_TERM)];
        J = vk_jac_coef(t) * dv[VK_IX(vk_jac_slot(t), nslot, VK_SITE_SLOT)];
        // terms in order; four loads in flight per round
        for (int u = en.y + 1; u < en.z; u += 4) {
            const int t0 = jac_tp[VK_IX(u, njac, VK_SITE_TERM)];
            const int t1 = u + 1 < en.z ? jac_tp[VK_IX(u + 1, njac, VK_SITE_TERM)] : t0;
            const int t2 = u + 2 < en.z ? jac_tp[VK_IX(u + 2, njac, VK_SITE_TERM)] : t0;
            const int t3 = u + 3 < en.z ? jac_tp[VK_IX(u + 3, njac, VK_SITE_TERM)] : t0;
            J = J + vk_jac_coef(t0) * dv[VK_IX(vk_jac_slot(t0), nslot, VK_SITE_SLOT)];
            if (u + 1 < en.z) J = J + vk_jac_coef(t1) * dv[VK_IX(vk_jac_slot(t1), nslot, VK_SITE_SLOT)];
            if (u + 2 < en.z) J = J + vk_jac_coef(t2) * dv[VK_IX(vk_jac_slot(t2), nslot, VK_SITE_SLOT)];
            if (u + 3 < en.z) J = J + vk_jac_coef(t3) * dv[VK_IX(vk_jac_slot(t3), nslot, VK_SITE_SLOT)];
        }
    }
    const int ex = VK_IX(en.x, ni * ni, VK_SITE_DENSE);
    if (mode == 0) {
        Dl[ex] = J;
        return;
    }
    const int i = en.w;
    double val = -J;
    if (ex == i * ni + i) val = lhs_diag(c0 + val, i, mol, dg0, dg1, dgx, xmode);
    Dl[ex] = val;
}

__global__ void __launch_bounds__(BLK) k_jac_lhs(
    int ni, int npairs, int nslot, int nent, int njac, const int *__restrict__ off, const int *__restrict__ layer_col,
    const int *__restrict__ reac, const int *__restrict__ prod, const int *__restrict__ slot_pair,
    const int *__restrict__ slot_sp, const int4 *__restrict__ jac_ent, const int *__restrict__ jac_emap, int nheavy,
    const int *__restrict__ jac_tp,
    const double *__restrict__ ms, const double *__restrict__ alpha, vk_atm atm, const double *__restrict__ y,
    const double *__restrict__ k, const double *__restrict__ M, int mode, int mol, const double *__restrict__ dtc,
    const double *__restrict__ ysum, const vk_colstate *__restrict__ st, double *__restrict__ D,
    double *__restrict__ Lsub, double *__restrict__ Usup, int write_zeros) {
    extern __shared__ double smem[];
    double *yl = smem;                       // ni
    double *dg0 = smem + VK_NI_MAX;          // diffusion terms of the diagonal, per species
    double *dg1 = smem + 2 * VK_NI_MAX;
    double *dgx = smem + 3 * VK_NI_MAX;      // escape (top) or deposition (bottom) term
    double *dv = smem + 4 * VK_NI_MAX;       // nslot
    const int l = blockIdx.x;
    LayerCtx x = layer_ctx(off, layer_col, l);
    if (st && st[x.c].done) return;
    const double *yrow = y + (size_t)l * ni;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) yl[i] = yrow[i];
    __syncthreads();   // the slot derivatives read yl
    const int nn = ni * ni;
    double *Dl = D + (size_t)l * nn;
    const double *kl = k + (size_t)l * 2 * npairs;
    const double m = M[l];
    // slot derivatives: JS slots per thread per round, every load of a round
    // (slot -> pair -> the pair's sides and rates, three loads deep) issued
    // before its arithmetic
    constexpr int JS = 4;
    for (int s0 = threadIdx.x; s0 < nslot; s0 += JS * blockDim.x) {
        int p[JS], sp[JS];
#pragma unroll
        for (int u = 0; u < JS; u++) {
            const int s = s0 + u * blockDim.x;
            p[u] = s < nslot ? VK_IX(slot_pair[s], npairs, VK_SITE_SLOTPAIR) : 0;
            sp[u] = s < nslot ? slot_sp[s] : 0;
        }
        int4 rt[JS], pt[JS];
        double kf[JS], kb[JS];
#pragma unroll
        for (int u = 0; u < JS; u++) {
            rt[u] = ((const int4 *)reac)[p[u]];
            pt[u] = ((const int4 *)prod)[p[u]];
            kf[u] = kl[2 * p[u]];
            kb[u] = kl[2 * p[u] + 1];
        }
#pragma unroll
        for (int u = 0; u < JS; u++) {
            const int s = s0 + u * blockDim.x;
            if (s < nslot) {
                bool a1, a2;
                const double dr = dside(rt[u], kf[u], yl, m, sp[u], a1, ni);
                const double dp = dside(pt[u], kb[u], yl, m, sp[u], a2, ni);
                dv[s] = (a1 ? dr : 0.0) - (a2 ? dp : 0.0);
            }
        }
    }
    double c0 = 0.0;
    int xmode = 0;
    if (mode == 1) {
        // diffusion part of the lhs (op.py:2012-2040): once per species
        c0 = 1. / (R_ROS * dtc[x.c]);
        double ysm, ys0, ysp;
        ysum_around(ysum, x, ysm, ys0, ysp);
        ColAtm ca = col_atm(atm, x.c, x.off, x.nz, ni);
        if (x.j == 0 && ca.use_bot) xmode = 2;
        else if (x.j == x.nz - 1 && ca.vm && mol && atm.top_flux && esc_any(atm)) xmode = 1;
        for (int i = threadIdx.x; i < ni; i += blockDim.x) {
            LhsDiff r = lhs_diff(ca, x.j, i, ysm, ys0, ysp, ms[i], alpha[i], mol != 0);
            dg0[i] = r.d0;
            dg1[i] = r.d1;
            if (xmode == 2) {
                dgx[i] = -1. * ca.bot_vdep[i] / ca.dzi[0];
            } else if (xmode == 1) {
                // diff_lim = top_flux / y for escaping species with y > 0 (no 1/dzi, as there)
                double e = 0.0;
                if (((atm.esc[i >> 5] >> (i & 31)) & 1u) && yrow[i] > 0)
                    e = 0.0 + atm.top_flux[(size_t)x.c * ni + i] / yrow[i];
                dgx[i] = e;
            }
            double u = 0.0, lo = 0.0;
            if (r.has_u) { u = 0.0 - r.u0; if (r.has_mol) u = u - r.u1; }
            if (r.has_l) { lo = 0.0 - r.l0; if (r.has_mol) lo = lo - r.l1; }
            if (blockIdx.y == 0) {
                Usup[(size_t)l * ni + i] = u;
                Lsub[(size_t)l * ni + i] = lo;
            }
        }
    }
    __syncthreads();
    // Every dense entry is written once (the structural zeros only with
    // write_zeros: the integrator clears its blocks once and skips them).
    // Entries with at most JAC_LIGHT
    // terms (and the structural zeros) in one coalesced pass over the dense
    // index; the few long term lists (jac_ent[0, nheavy): the list is sorted
    // by term count) by their own threads alongside.  With few layers the
    // entries of a layer are split over the grid's y dimension (each
    // workgroup forms the layer's shared terms itself).
    const int e0 = threadIdx.x + blockIdx.y * blockDim.x, es = blockDim.x * gridDim.y;
    for (int q = e0; q < nheavy; q += es)
        jac_entry(q, jac_ent, jac_tp, dv, ni, njac, nslot, mode, mol, c0, dg0, dg1, dgx, xmode, Dl);
#pragma unroll 4
    for (int e = e0; e < nn; e += es) {
        // (first term << 4 | term count), -1: structural zero, -2: long term list
        const int code = jac_emap[e];
        if (code == -2 || (code == -1 && !write_zeros)) continue;
        double val = 0.0;
        if (code >= 0) {
            const int u0 = code >> 4, nt = code & 15;
            double J = 0.0;
            if (nt > 0) {
                // only the entry's own terms are loaded (all issued before the sums)
                int t[VK_JAC_LIGHT];
#pragma unroll
                for (int m = 0; m < VK_JAC_LIGHT; m++) t[m] = m < nt ? jac_tp[VK_IX(u0 + m, njac, VK_SITE_TERM)] : 0;
                J = vk_jac_coef(t[0]) * dv[VK_IX(vk_jac_slot(t[0]), nslot, VK_SITE_SLOT)];
#pragma unroll
                for (int m = 1; m < VK_JAC_LIGHT; m++)
                    if (m < nt) J = J + vk_jac_coef(t[m]) * dv[VK_IX(vk_jac_slot(t[m]), nslot, VK_SITE_SLOT)];
            }
            if (mode == 0) {
                val = J;
            } else {
                const int i = e / ni;
                val = -J;
                if (e == i * ni + i) val = lhs_diag(c0 + val, i, mol, dg0, dg1, dgx, xmode);
            }
        }
        Dl[e] = val;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host entry points
// ---------------------------------------------------------------------------

static inline hipStream_t S(void *s) { return (hipStream_t)s; }

static size_t rhs_smem(const vk_plan *p) { return sizeof(double) * (VK_NI_MAX + (size_t)p->npairs); }
static size_t jac_smem(const vk_plan *p) { return sizeof(double) * (4 * VK_NI_MAX + (size_t)p->nslot); }

extern "C" int vk_rates(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, double *d_k, void *stream) {
    if (!plan || !cols || !atm || !d_k) return VK_EINVAL;
    int L = cols->L;
    double *grt = nullptr;
    VK_CHECK(hipMallocAsync((void **)&grt, sizeof(double) * (size_t)L * plan->ni, S(stream)));
    int n1 = L * plan->ni;
    hipLaunchKernelGGL(k_grt, dim3((n1 + 255) / 256), dim3(256), 0, S(stream), L, plan->ni, atm->Tco, plan->nasa9, grt);
    int n2 = L * plan->npairs;
    hipLaunchKernelGGL(k_rates, dim3((n2 + 255) / 256), dim3(256), 0, S(stream), L, plan->npairs, plan->ni,
                       plan->kind, plan->par, plan->rev, plan->reac, plan->prod, atm->Tco, atm->M, atm->tau, grt, d_k);
    VK_CHECK(hipGetLastError());
    VK_CHECK(hipFreeAsync(grt, S(stream)));
    return 0;
}

static int launch_ysum(const vk_columns *cols, int ni, const double *y, double *ysum, hipStream_t s) {
    hipLaunchKernelGGL(k_ysum, dim3((cols->L + 127) / 128), dim3(128), 0, s, cols->L, ni, y, ysum);
    VK_CHECK(hipGetLastError());
    return 0;
}

// ysum: scratch of L doubles (needed when the diffusion terms are evaluated)
int vk_launch_rhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *y,
                  const double *k, int mode, int mol, const double *k1, const double *dtc,
                  const vk_colstate *st, double *ysum, double *out, hipStream_t s) {
    vk_atm a = *atm;
    if (mode & 2) {
        if (!ysum) return VK_EINVAL;
        int r = launch_ysum(cols, plan->ni, y, ysum, s);
        if (r) return r;
    }
    hipLaunchKernelGGL(k_rhs, dim3(cols->L), dim3(BLK), rhs_smem(plan), s, plan->ni, plan->npairs, plan->nchem,
                       cols->d_off,
                       cols->d_layer_col, plan->reac, plan->prod, plan->chem_ptr, plan->chem_ent, plan->ms,
                       plan->alpha, a, y, k, atm->M, mode, mol, k1, dtc, ysum, st, out);
    VK_CHECK(hipGetLastError());
    return 0;
}

// ysum_ready: the caller's earlier launch on s has left the layer sums of y in ysum
int vk_launch_jac_lhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *y,
                      const double *k, const double *M, int mode, int mol, const double *dtc,
                      const vk_colstate *st, double *ysum, double *D, double *Lsub, double *Usup, hipStream_t s,
                      bool ysum_ready = false, bool write_zeros = true) {
    vk_atm a{};
    if (atm) a = *atm;
    if (mode == 1) {
        if (!ysum) return VK_EINVAL;
        if (!ysum_ready) {
            int r = launch_ysum(cols, plan->ni, y, ysum, s);
            if (r) return r;
        }
    }
    // fewer layers than twice the CUs: split each layer's entries over up to 4 workgroups
    static const int ncu = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 0;
        return n;
    }();
    const int gy = ncu > 0 ? std::max(1, std::min(4, 2 * ncu / std::max(cols->L, 1))) : 1;
    hipLaunchKernelGGL(k_jac_lhs, dim3(cols->L, gy), dim3(BLK), jac_smem(plan), s, plan->ni, plan->npairs, plan->nslot,
                       plan->njent, plan->njac, cols->d_off, cols->d_layer_col, plan->reac, plan->prod, plan->slot_pair,
                       plan->slot_sp, plan->jac_ent, plan->jac_emap, plan->njheavy, plan->jac_tp, plan->ms, plan->alpha, a, y, k, M, mode, mol,
                       dtc, ysum, st, D, Lsub, Usup, (int)write_zeros);
    VK_CHECK(hipGetLastError());
    return 0;
}

// temporary layer totals for the one-shot entry points
struct YsumTmp {
    double *p = nullptr;
    hipStream_t s;
    explicit YsumTmp(hipStream_t st) : s(st) {}
    int alloc(int L) { return hipMallocAsync((void **)&p, sizeof(double) * (size_t)L, s) == hipSuccess ? 0 : VK_ENOMEM; }
    ~YsumTmp() { if (p) (void)hipFreeAsync(p, s); }
};

extern "C" int vk_chemdf(const vk_plan *plan, const vk_columns *cols, const double *d_y, const double *d_M,
                         const double *d_k, double *d_dydt, void *stream) {
    if (!plan || !cols || !d_y || !d_M || !d_k || !d_dydt) return VK_EINVAL;
    vk_atm a{};
    a.M = d_M;
    return vk_launch_rhs(plan, cols, &a, d_y, d_k, 1, 0, nullptr, nullptr, nullptr, nullptr, d_dydt, S(stream));
}

extern "C" int vk_diffdf(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *d_y,
                         int use_moldiff, double *d_out, void *stream) {
    if (!plan || !cols || !atm || !d_y || !d_out) return VK_EINVAL;
    YsumTmp ys(S(stream));
    if (ys.alloc(cols->L)) return VK_ENOMEM;
    return vk_launch_rhs(plan, cols, atm, d_y, nullptr, 2, use_moldiff, nullptr, nullptr, nullptr, ys.p, d_out,
                         S(stream));
}

extern "C" int vk_rhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *d_y,
                      const double *d_k, int use_moldiff, const double *d_k1, const double *d_dt, double *d_out,
                      void *stream) {
    if (!plan || !cols || !atm || !d_y || !d_k || !d_out || (d_k1 && !d_dt)) return VK_EINVAL;
    YsumTmp ys(S(stream));
    if (ys.alloc(cols->L)) return VK_ENOMEM;
    return vk_launch_rhs(plan, cols, atm, d_y, d_k, 3, use_moldiff, d_k1, d_dt, nullptr, ys.p, d_out, S(stream));
}

extern "C" int vk_jac(const vk_plan *plan, const vk_columns *cols, const double *d_y, const double *d_M,
                      const double *d_k, double *d_J, void *stream) {
    if (!plan || !cols || !d_y || !d_M || !d_k || !d_J) return VK_EINVAL;
    return vk_launch_jac_lhs(plan, cols, nullptr, d_y, d_k, d_M, 0, 0, nullptr, nullptr, nullptr, d_J, nullptr, nullptr,
                             S(stream));
}

extern "C" int vk_lhs(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *d_y,
                      const double *d_k, int use_moldiff, const double *d_dt, double *d_D, double *d_Lsub,
                      double *d_Usup, void *stream) {
    if (!plan || !cols || !atm || !d_y || !d_k || !d_dt || !d_D || !d_Lsub || !d_Usup) return VK_EINVAL;
    YsumTmp ys(S(stream));
    if (ys.alloc(cols->L)) return VK_ENOMEM;
    return vk_launch_jac_lhs(plan, cols, atm, d_y, d_k, atm->M, 1, use_moldiff, d_dt, nullptr, ys.p, d_D, d_Lsub, d_Usup,
                             S(stream));
}

extern "C" int vk_debug_checks(int *site_out, int reset) {
    if (!site_out) return VK_EINVAL;
#ifdef VK_CHECKED
    int v = 0;
    VK_CHECK(hipMemcpyFromSymbol(&v, HIP_SYMBOL(vk_check_word), sizeof(int)));
    if (reset) {
        const int z = 0;
        VK_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(vk_check_word), &z, sizeof(int)));
    }
    *site_out = v;
    return 0;
#else
    (void)reset;
    return VK_ENOTCHECKED;
#endif
}
