// Plan / column layout objects and the device-resident Ros2 state machine:
// one attempt = rhs -> lhs -> factor -> solve -> rhs2 -> solve -> finalize
// (op.py:2860-3007), then per column clip / loss / step_ok / step_reject
// (op.py:2447-2534), update_mu_dz (944-984), renormalisation (909-914),
// save_step (1087-1105), step_size (3105-3125) and stop/conv (1018-1085).
#include <cstring>
#include <algorithm>
#include <type_traits>
#include "vk_internal.h"
#include "vk_device.h"

using namespace vk;

int vk_launch_rhs(const vk_plan *, const vk_columns *, const vk_atm *, const double *, const double *, int, int,
                  const double *, const double *, const vk_colstate *, double *, double *, hipStream_t);
int vk_launch_jac_lhs(const vk_plan *, const vk_columns *, const vk_atm *, const double *, const double *,
                      const double *, int, int, const double *, const vk_colstate *, double *, double *, double *,
                      double *, hipStream_t, bool ysum_ready = false, bool write_zeros = true);
int vk_launch_factor(const vk_plan *, const vk_columns *, const double *, const double *, const double *, double *,
                     int *, const vk_colstate *, const vk_bcr_ws *, hipStream_t, bool info_zeroed = false);
int vk_launch_solve_refined(const vk_plan *, const vk_columns *, const double *, const double *, const double *,
                            const double *, const double *, double *, double *, double *, int, const vk_colstate *, const vk_bcr_ws *,
                            hipStream_t);

namespace {

constexpr int NP = 4 + VK_NATOM_MAX;   // per-layer partials stride
constexpr double C3 = 3. / (2. * R_ROS);
constexpr double C1 = 1 / (2. * R_ROS);

template <typename T>
int dalloc(std::vector<void *> &v, T **p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) return VK_ENOMEM;
    v.push_back(*p);
    return 0;
}

template <typename T>
int upload(std::vector<void *> &v, T **p, const T *src, size_t n) {
    int r = dalloc(v, p, n);
    if (r) return r;
    if (n && src) {
        hipError_t e = hipMemcpy(*p, src, sizeof(T) * n, hipMemcpyHostToDevice);
        if (e != hipSuccess) return -(int)e;
    }
    return 0;
}

__device__ inline void layer_of(const int *off, const int *layer_col, int l, int &c, int &o, int &nz) {
    c = layer_col[l];
    o = off[c];
    nz = off[c + 1] - o;
}

// dt of every column, gathered contiguously; also clears the column's
// zero-pivot record before the attempt's factorisation
__global__ void k_gather_dt(int n, const vk_colstate *__restrict__ st, double *__restrict__ dt,
                            int *__restrict__ info) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) {
        dt[c] = st[c].dt;
        if (info) info[c] = 0;
    }
}

// yk2 = y + k1/r  (op.py:2917)
__global__ void k_yk2(int n, int ni, const int *__restrict__ layer_col, const double *__restrict__ y,
                      const double *__restrict__ k1, const vk_colstate *__restrict__ st, double *__restrict__ yk2) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    if (st[layer_col[idx / ni]].done) return;
    yk2[idx] = y[idx] + k1[idx] / R_ROS;
}

// sol, delta partial, ymix; with clip also the clip/loss partials (op.py:2932-2993, 2447-2487).
// fix_bot (may be NULL): fixed bottom mixing ratios, < 0 free (op.py:2945-2946);
// mask_bot: layer 0 leaves delta (op.py:2953)
__global__ void __launch_bounds__(256) k_finalize(int ni, int natom, const int *__restrict__ off,
                                                  const int *__restrict__ layer_col,
                                                  const double *__restrict__ atom_mat, const double *y,
                                                  const double *ymix_in, const double *__restrict__ k1,
                                                  const double *__restrict__ k2, const double *__restrict__ yk2,
                                                  vk_params prm, int clip, const vk_colstate *__restrict__ st,
                                                  const double *__restrict__ fix_bot, const double *__restrict__ M,
                                                  int mask_bot, double *ynew, double *ymixnew,
                                                  double *__restrict__ part) {
    __shared__ double sol[VK_NI_MAX];
    __shared__ double red[256];
    __shared__ int negf[256];
    __shared__ double sm[256], ng[256];
    __shared__ double ysum;
    const int l = blockIdx.x;
    int c, o, nz;
    layer_of(off, layer_col, l, c, o, nz);
    if (st[c].done) return;
    const size_t base = (size_t)l * ni;
    const bool bottom = l == o;
    const double *fb = fix_bot && bottom ? fix_bot + (size_t)c * ni : nullptr;
    double rmax = -1.0;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        double s = y[base + i] + C3 * k1[base + i] + C1 * k2[base + i];
        if (fb && fb[i] >= 0) s = fb[i] * M[l];
        double d = fabs(s - yk2[base + i]);
        if (ymix_in[base + i] < prm.mtol) d = 0;
        if (s < prm.atol) d = 0;
        if (bottom && mask_bot) d = 0;
        if (s > 0) rmax = fmax(rmax, d / s);
        sol[i] = s;
    }
    // maximum over the workgroup (exact in any order): within each wave by
    // lane exchanges, then over the waves
    for (int o = 32; o > 0; o >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = rmax;
    __syncthreads();
    // the serial sums on threads of different waves, each in its own order
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) r = fmax(r, red[q]);
        part[(size_t)l * NP + VKP_DELTA] = r;
    }
    if (threadIdx.x == 64) ysum = pw_sum(sol, ni);
    __syncthreads();
    const double ys = ysum;
    if (!clip) {
        for (int i = threadIdx.x; i < ni; i += blockDim.x) {
            ynew[base + i] = sol[i];
            ymixnew[base + i] = sol[i] / ys;
        }
        return;
    }
    double small = 0.0, nega = 0.0;
    int anyneg = 0;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        double v = sol[i];
        double mx = v / ys;
        if (v < prm.pos_cut && v >= 0) small += v;
        if (v > prm.nega_cut && v <= 0) nega += v;
        if (v < prm.pos_cut && v >= prm.nega_cut) v = 0.;
        if (mx < prm.mtol && v < 0) v = 0.;
        if (v < 0) anyneg = 1;
        sol[i] = v;
    }
    sm[threadIdx.x] = small;
    ng[threadIdx.x] = nega;
    negf[threadIdx.x] = anyneg;
    __syncthreads();
    // threads from ni on hold zero partials (adding +0.0 to a sum that
    // starts at +0.0 changes nothing); each sum in thread order, the sums on
    // threads of different waves
    const int nt = min((int)blockDim.x, ni);
    if (threadIdx.x == 0) {
        double s1 = 0.0;
        for (int t = 0; t < nt; t++) s1 += sm[t];
        part[(size_t)l * NP + VKP_SMALL] = s1;
    } else if (threadIdx.x == 64) {
        double s2 = 0.0;
        int nf = 0;
        for (int t = 0; t < nt; t++) { s2 += ng[t]; nf |= negf[t]; }
        part[(size_t)l * NP + VKP_NEG] = nf;
        part[(size_t)l * NP + VKP_NEGY] = s2;
    } else if (threadIdx.x == 128) {
        ysum = pw_sum(sol, ni);
    }
    if (threadIdx.x >= 192 && threadIdx.x - 192 < natom) {
        const int a = threadIdx.x - 192;
        double acc = 0.0;
        for (int i = 0; i < ni; i++) acc += atom_mat[(size_t)i * natom + a] * sol[i];
        part[(size_t)l * NP + VKP_ATOM + a] = acc;
    }
    __syncthreads();
    const double ys2 = ysum;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        ynew[base + i] = sol[i];
        ymixnew[base + i] = sol[i] / ys2;
    }
}

// element sums per column: same per-layer partial + in-order column sum as k_finalize/k_decide
__global__ void k_atom_layer(int ni, int natom, const double *__restrict__ atom_mat, const double *__restrict__ y,
                             double *__restrict__ part) {
    const int l = blockIdx.x;
    if (threadIdx.x < natom) {
        const int a = threadIdx.x;
        double acc = 0.0;
        for (int i = 0; i < ni; i++) acc += atom_mat[(size_t)i * natom + a] * y[(size_t)l * ni + i];
        part[(size_t)l * NP + VKP_ATOM + a] = acc;
    }
}

__global__ void k_atom_col(int natom, const int *__restrict__ off, const double *__restrict__ part,
                           double *__restrict__ out) {
    const int c = blockIdx.x;
    const int a = threadIdx.x;
    if (a >= natom) return;
    double acc = 0.0;
    for (int l = off[c]; l < off[c + 1]; l++) acc += part[(size_t)l * NP + VKP_ATOM + a];
    out[(size_t)c * natom + a] = acc;
}

// per-column decision.  mode 0: only set delta (plain solver); mode 1: one_step logic.
// flags: 0 retry, 1 accepted, 2 accepted after giving up, 3 inactive; bit 4: run update_mu_dz
// One workgroup per column: the column's per-layer partials are staged in
// LDS with coalesced loads (when they fit), then one thread reduces them in
// layer order (the order of the reference's sums) and takes the decision.
constexpr int DECIDE_NT = 256;
constexpr size_t DECIDE_STAGE_MAX = 48 * 1024;

__global__ void __launch_bounds__(DECIDE_NT) k_decide(int natom, const int *__restrict__ off,
                                                      const double *__restrict__ part, vk_params prm, int mode,
                                                      const double *__restrict__ atom_ini,
                                                      double *__restrict__ atom_loss,
                                                      double *__restrict__ atom_loss_prev,
                                                      vk_colstate *__restrict__ st, int *__restrict__ flags,
                                                      double *__restrict__ thist, int tcap, int stage) {
    extern __shared__ double dsm[];
    const int c = blockIdx.x;
    vk_colstate s = st[c];
    if (s.done) { if (flags && threadIdx.x == 0) flags[c] = 3; return; }
    const int l0 = off[c], l1 = off[c + 1];
    const double *pc = part + (size_t)l0 * NP;
    if (stage) {
        for (int e = threadIdx.x; e < (l1 - l0) * NP; e += DECIDE_NT) dsm[e] = pc[e];
        __syncthreads();
        pc = dsm;
    }
    // the column reductions, one thread each (every one in layer order):
    // 0 the largest delta, 1 any negative entry, 2 / 3 the small and
    // negative sums, 4 + a the total of atom a
    __shared__ double red[4 + VK_NATOM_MAX];
    const int t = threadIdx.x, nq = mode == 0 ? 1 : 4 + natom;
    if (t < nq) {
        const int col = t == 0 ? VKP_DELTA : t == 1 ? VKP_NEG : t == 2 ? VKP_SMALL : t == 3 ? VKP_NEGY
                                                                                      : VKP_ATOM + t - 4;
        const double *pq = pc + col;
        // one loop for all lanes (no divergence): each keeps the reduction it needs
        double sum = 0.0, mx = -1.0;
        int nz = 0;
#pragma unroll 8
        for (int l = 0; l < l1 - l0; l++) {
            const double v = pq[(size_t)l * NP];
            sum += v;
            mx = fmax(mx, v);
            nz |= v != 0.0;
        }
        red[t] = t == 0 ? mx : t == 1 ? (nz ? 1.0 : 0.0) : sum;
    }
    __syncthreads();
    if (t != 0) return;
    double delta = red[0];
    if (delta < 0) delta = 0.0;
    s.delta = delta;
    if (mode == 0) { st[c] = s; return; }
    const bool neg = red[1] != 0.0;
    s.small_y += fabs(red[2]);
    s.nega_y += fabs(red[3]);
    double lossmax = 0.0;
    bool lossnan = false;
    for (int a = 0; a < natom; a++) {
        const double acc = red[4 + a];
        const double ini = atom_ini[(size_t)c * natom + a];
        const double loss = (acc - ini) / ini;
        atom_loss[(size_t)c * natom + a] = loss;
        const double d = fabs(loss - atom_loss_prev[(size_t)c * natom + a]);
        if (d != d) lossnan = true;
        lossmax = fmax(lossmax, d);
    }
    s.attempts += 1;
    const bool ok = !neg && !lossnan && lossmax < prm.loss_eps && delta <= prm.rtol;
    int f;
    if (ok) {
        f = 1;
    } else {
        if (delta > prm.rtol) s.delta_count += 1;
        else if (neg) s.nega_count += 1;
        else s.loss_count += 1;
        s.dt *= prm.dt_var_min;
        if (s.dt < prm.dt_min) { s.dt = prm.dt_min; f = 2; }
        else f = 0;
    }
    if (f) {
        for (int a = 0; a < natom; a++) atom_loss_prev[(size_t)c * natom + a] = atom_loss[(size_t)c * natom + a];
        if (s.count % prm.update_frq == 0) f |= 16;      // update_mu_dz / update_phi_esc due
        // save_step (op.py:1087-1105)
        s.t += s.dt;
        s.count += 1;
        if (thist && s.count - 1 < tcap) thist[(size_t)c * tcap + s.count - 1] = s.t;
        // step_size (op.py:3105-3125)
        double h = s.dt;
        double dl = s.delta;
        if (dl == 0) dl = 0.01 * prm.rtol;
        double hf = 0.9 * sqrt(prm.rtol / dl);
        hf = fmax(hf, prm.dt_var_min);
        hf = fmin(hf, prm.dt_var_max);
        h *= hf;
        h = fmax(h, prm.dt_min);
        h = fmin(h, prm.dt_max);
        s.dt = h;
    }
    st[c] = s;
    flags[c] = f;
}

// update_mu_dz (op.py:944-984) for flagged columns
__global__ void k_update_mu(int ni, const int *__restrict__ off, const double *__restrict__ ms,
                            const double *__restrict__ ymix, vk_atm atm, vk_geom geo, const int *__restrict__ flags,
                            int force, const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x;
    if (force) { if (st && st[c].done) return; }
    else if (!(flags[c] & 16)) return;
    const int o = off[c], nz = off[c + 1] - o;
    for (int j = threadIdx.x; j < nz; j += blockDim.x) {
        const double *ym = ymix + (size_t)(o + j) * ni;
        double mu = 0.0;
        for (int i = 0; i < ni; i++) mu += ms[i] * ym[i];
        geo.mu[o + j] = mu;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double gs = geo.gs[c], Rp = geo.Rp[c];
        const int pref = geo.pref[c];
        double *g = atm.g + o, *Hp = atm.Hp + o, *dz = geo.dz + o, *zco = geo.zco + o + c;
        const double *T = atm.Tco + o, *mu = geo.mu + o, *lnp = geo.lnp + o;
        for (int i = pref; i < nz; i++) {
            if (i == pref) {
                g[i] = gs;
                Hp[i] = KB * T[i] / (mu[i] / NAVO * gs);
            } else {
                double q = Rp / (Rp + zco[i]);
                g[i] = gs * (q * q);
                Hp[i] = KB * T[i] / (mu[i] / NAVO * g[i]);
            }
            dz[i] = Hp[i] * lnp[i];
            zco[i + 1] = zco[i] + dz[i];
        }
        for (int i = pref - 1; i >= 0; i--) {
            double q = Rp / (Rp + zco[i + 1]);
            g[i] = gs * (q * q);
            Hp[i] = KB * T[i] / (mu[i] / NAVO * g[i]);
            dz[i] = Hp[i] * lnp[i];
            zco[i] = zco[i + 1] - dz[i];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < nz - 1; j += blockDim.x) {
        atm.dzi[o + j] = 0.5 * (geo.dz[o + j + 1] + geo.dz[o + j]);
        atm.Hpi[o + j] = 0.5 * (atm.Hp[o + j] + atm.Hp[o + j + 1]);
    }
}

// update_phi_esc (op.py:986-999) for flagged columns: the diffusion-limited
// top flux of each escaping species from the step's y (before the
// renormalisation), capped at max_flux
__global__ void k_phi_esc(int ni, const int *__restrict__ off, const double *__restrict__ ms,
                          const double *__restrict__ y, vk_atm atm, const int *__restrict__ flags, int force,
                          const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x;
    if (force) { if (st && st[c].done) return; }
    else if (!(flags[c] & 16)) return;
    const int t = off[c + 1] - 1;      // top layer; Dzz[-1] is interface t - 1
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        if (!((atm.esc[i >> 5] >> (i & 31)) & 1u)) continue;
        const double f = -atm.Dzz[(size_t)(t - 1) * ni + i] * y[(size_t)t * ni + i] *
                         (1. / atm.Hp[t] - ms[i] * atm.g[t] / (NAVO_KB * atm.Tco[t]));
        const double cap = atm.max_flux * (-1);
        atm.top_flux[(size_t)c * ni + i] = cap > f ? cap : f;      // Python's max(f, cap)
    }
}

// reset_y (op.py:2524-2534) or renormalise + backup + history (op.py:909-914, 938-942, 1096)
__global__ void k_apply(int ni, const int *__restrict__ layer_col, const int *__restrict__ flags,
                        const double *__restrict__ M, const double *__restrict__ ymix, double *y, double *y_prev,
                        const vk_colstate *__restrict__ st, double *__restrict__ yhist, int H, size_t Lni) {
    const int l = blockIdx.x;
    const int c = layer_col[l];
    const int f = flags[c] & 15;
    if (f == 3) return;
    const size_t base = (size_t)l * ni;
    if (f == 0) {
        for (int i = threadIdx.x; i < ni; i += blockDim.x) y[base + i] = y_prev[base + i];
        return;
    }
    const double n0 = M[l];
    double *h = nullptr;
    if (yhist && H > 0) h = yhist + (size_t)((st[c].count - 1) % H) * Lni + base;
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        double v = n0 * ymix[base + i];
        y[base + i] = v;
        y_prev[base + i] = v;
        if (h) h[i] = v;
    }
}

// stop() with conv() (op.py:1018-1085)
__global__ void __launch_bounds__(256) k_conv(int ni, const int *__restrict__ off, vk_params prm, vk_atm atm,
                                              const double *__restrict__ y, const double *__restrict__ ymix,
                                              const int *__restrict__ flags, vk_colstate *__restrict__ st,
                                              const double *__restrict__ yhist, int H, size_t Lni,
                                              const double *__restrict__ thist, int tcap) {
    const int c = blockIdx.x;
    if (flags && (flags[c] & 15) != 1 && (flags[c] & 15) != 2) return;
    __shared__ double rv[256];
    __shared__ int ri[256];
    __shared__ vk_colstate s;
    if (threadIdx.x == 0) s = st[c];
    __syncthreads();
    if (s.done) return;
    const int o = off[c], nz = off[c + 1] - o;
    bool conv = false;
    const bool test = prm.use_conv && H > 0 && s.t > prm.trun_min && s.count > prm.count_min && s.count <= tcap;
    if (test) {
        const int count = s.count;
        const double x = s.t * prm.st_factor;
        const int w0 = count - prm.conv_step > 0 ? count - prm.conv_step : 0;
        double bv = INFINITY;
        int bi = 0x7fffffff;
        for (int q = w0 + threadIdx.x; q < count; q += blockDim.x) {
            double v = fabs(thist[(size_t)c * tcap + q] - x);
            if (v < bv || (v == bv && q < bi)) { bv = v; bi = q; }
        }
        rv[threadIdx.x] = bv;
        ri[threadIdx.x] = bi;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) {
                double ov = rv[threadIdx.x + w];
                int oi = ri[threadIdx.x + w];
                if (ov < rv[threadIdx.x] || (ov == rv[threadIdx.x] && oi < ri[threadIdx.x])) {
                    rv[threadIdx.x] = ov;
                    ri[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        int indx = ri[0];
        if (indx == count - 1) indx -= 1;
        if (indx < count - prm.conv_step) indx = count - prm.conv_step;
        if (indx < 0) indx += count;   // Python negative index
        const double *ycur = y;        // y_time[count-1] is the renormalised current y
        const double *yold = yhist + (size_t)(indx % H) * Lni;
        double lmax = -1.0;
        for (int q = threadIdx.x; q < nz * ni; q += blockDim.x) {
            const int j = q / ni;
            const size_t e = (size_t)(o + j) * ni + (q - j * ni);
            const int i = q - j * ni;
            double d = fabs((ycur[e] - yold[e]) / atm.M[o + j]);
            const double mx = ymix[e];
            if (mx < prm.mtol_conv) d = 0;
            if (ycur[e] < prm.atol) d = 0;
            if ((prm.conv_ignore[i >> 5] >> (i & 31)) & 1u) d = 0;   // conver_ignore, op.py:1045-1046
            if (mx > 0) lmax = fmax(lmax, d / mx);
        }
        __syncthreads();
        rv[threadIdx.x] = lmax;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if (threadIdx.x < w) rv[threadIdx.x] = fmax(rv[threadIdx.x], rv[threadIdx.x + w]);
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double longdy = rv[0];
            double longdydt = longdy / (thist[(size_t)c * tcap + count - 1] - thist[(size_t)c * tcap + indx]);
            double smin = 1e300;
            for (int j = 0; j < nz - 1; j++) {
                double hq = 0.1 * atm.Hp[o + j];
                smin = fmin(smin, atm.Kzz[o + j] / (hq * hq));
            }
            smin = fmin(smin, 1.e-8);
            smin = fmax(smin, 1.e-10);
            s.longdy = longdy;
            s.longdydt = longdydt;
            conv = ((longdy < prm.yconv_cri && longdydt < prm.slope_cri) ||
                    (longdy < prm.yconv_min && longdydt < smin)) && s.aflux_change < prm.flux_cri;
        }
    }
    if (threadIdx.x == 0) {
        if (conv) { s.end_case = 1; s.done = 1; }
        else if (s.t > prm.runtime) { s.end_case = 2; s.done = 1; }
        else if (s.count > prm.count_max) { s.end_case = 3; s.done = 1; }
        st[c] = s;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// plan / columns
// ---------------------------------------------------------------------------

extern "C" int vk_plan_create(const vk_network_desc *n, vk_plan **out) {
    if (!n || !out || n->ni <= 0 || n->npairs <= 0) return VK_EINVAL;
    if (n->ni > VK_NI_MAX || n->natom > VK_NATOM_MAX) return VK_ETOOBIG;
    // the per-layer kernels keep a layer's pair rates / Jacobian slots in LDS (64 KiB)
    if (sizeof(double) * ((size_t)VK_NI_MAX + n->npairs) > 65536 ||
        sizeof(double) * (4 * (size_t)VK_NI_MAX + n->nslot) > 65536)
        return VK_ETOOBIG;
    vk_plan *p = new vk_plan();
    p->ni = n->ni; p->npairs = n->npairs; p->nr = 2 * n->npairs; p->natom = n->natom;
    p->factor_variant = VK_FACTOR_REGISTER;
    p->nchem = n->nchem; p->nslot = n->nslot; p->njnz = n->njnz; p->njac = n->njac;
    int r = 0;
    auto &v = p->allocs;
    r = r ? r : upload(v, &p->reac, n->reac, (size_t)4 * n->npairs);
    r = r ? r : upload(v, &p->prod, n->prod, (size_t)4 * n->npairs);
    r = r ? r : upload(v, &p->kind, n->kind, (size_t)n->npairs);
    r = r ? r : upload(v, &p->rev, n->rev, (size_t)n->npairs);
    r = r ? r : upload(v, &p->par, n->par, (size_t)8 * n->npairs);
    r = r ? r : upload(v, &p->nasa9, n->nasa9, (size_t)20 * n->ni);
    r = r ? r : upload(v, &p->atom_mat, n->atom_mat, (size_t)n->ni * n->natom);
    r = r ? r : upload(v, &p->ms, n->ms, (size_t)n->ni);
    r = r ? r : upload(v, &p->alpha, n->alpha, (size_t)n->ni);
    r = r ? r : upload(v, &p->chem_ptr, n->chem_ptr, (size_t)n->ni + 1);
    r = r ? r : upload(v, &p->chem_pair, n->chem_pair, (size_t)n->nchem);
    r = r ? r : upload(v, &p->chem_sign, n->chem_sign, (size_t)n->nchem);
    r = r ? r : upload(v, &p->slot_pair, n->slot_pair, (size_t)n->nslot);
    r = r ? r : upload(v, &p->slot_sp, n->slot_sp, (size_t)n->nslot);
    r = r ? r : upload(v, &p->jac_map, n->jac_map, (size_t)n->ni * n->ni);
    r = r ? r : upload(v, &p->jac_ptr, n->jac_ptr, (size_t)n->njnz + 1);
    r = r ? r : upload(v, &p->jac_slot, n->jac_slot, (size_t)n->njac);
    r = r ? r : upload(v, &p->jac_coef, n->jac_coef, (size_t)n->njac);
    if (!r) {
        std::vector<int4> ent;
        for (int e = 0; e < n->ni * n->ni; e++) {
            const int jm = n->jac_map[e], row = e / n->ni;
            if (jm >= 0) ent.push_back(make_int4(e, n->jac_ptr[jm], n->jac_ptr[jm + 1], row));
            else if (e % n->ni == row) ent.push_back(make_int4(e, 0, 0, row));
        }
        // heaviest term lists first: the lanes of a wave then carry term
        // lists of similar length (the diagonal entries, with the longest
        // lists, share a few waves instead of lengthening every wave)
        std::stable_sort(ent.begin(), ent.end(),
                         [](const int4 &x, const int4 &y) { return (x.z - x.y) > (y.z - y.y); });
        p->njent = (int)ent.size();
        p->njheavy = 0;
        std::vector<int> emap((size_t)n->ni * n->ni, -1);
        for (int q = 0; q < p->njent; q++) {
            const int nt = ent[q].z - ent[q].y;
            if (nt > VK_JAC_LIGHT) p->njheavy = q + 1;
            emap[ent[q].x] = nt > VK_JAC_LIGHT ? -2 : (ent[q].y << 4) | nt;
        }
        r = upload(v, &p->jac_ent, ent.data(), ent.size());
        r = r ? r : upload(v, &p->jac_emap, emap.data(), emap.size());
    }
    if (!r) {
        std::vector<int2> ce((size_t)n->nchem);
        for (size_t q = 0; q < ce.size(); q++) ce[q] = make_int2(n->chem_pair[q], n->chem_sign[q]);
        r = upload(v, &p->chem_ent, ce.data(), ce.size());
    }
    if (!r) {
        std::vector<int> terms((size_t)n->njac);
        for (size_t q = 0; q < terms.size() && !r; q++) {
            const double c = n->jac_coef[q];
            const int sl = n->jac_slot[q];
            // the coefficients are net stoichiometric numbers
            if (c != (double)(int)c || c < -128.0 || c > 127.0 || sl < 0 || sl >= VK_JAC_SLOT_MAX) r = VK_EINVAL;
            else terms[q] = vk_jac_pack(sl, (int)c);
        }
        r = r ? r : upload(v, &p->jac_tp, terms.data(), terms.size());
    }
    if (r) { vk_plan_destroy(p); return r; }
    *out = p;
    return 0;
}

extern "C" int vk_plan_destroy(vk_plan *p) {
    if (!p) return 0;
    for (void *q : p->allocs) (void)hipFree(q);
    delete p;
    return 0;
}

extern "C" int vk_plan_info(const vk_plan *p, int *ni, int *nr, int *max_ni) {
    if (ni) *ni = p ? p->ni : 0;
    if (nr) *nr = p ? p->nr : 0;
    if (max_ni) *max_ni = VK_NI_MAX;
    return 0;
}

// levels of block cyclic reduction for every column (see vk_bcr_sched)
static int bcr_schedule(vk_columns *c) {
    vk_bcr_sched &b = c->bcr;
    std::vector<std::vector<int>> act(c->ncols);
    for (int i = 0; i < c->ncols; i++)
        for (int l = c->off[i]; l < c->off[i + 1]; l++) act[i].push_back(l);
    std::vector<int4> el, sv;
    std::vector<int> fl;
    std::vector<int2> top;
    b.elim_off.push_back(0);
    b.surv_off.push_back(0);
    for (;;) {
        bool any = false;
        for (int i = 0; i < c->ncols; i++) {
            const std::vector<int> &a = act[i];
            const int n = (int)a.size();
            if (n < 2) continue;
            any = true;
            // eliminated positions: the even ones of an odd count (floor(n/2)
            // survivors, one level fewer than keeping the ends), else the odd ones
            const int pe = n & 1 ? 0 : 1;
            for (int q = pe; q < n; q += 2)
                el.push_back(make_int4(a[q], q - 1 >= 0 ? a[q - 1] : -1, q + 1 < n ? a[q + 1] : -1, i));
            std::vector<int> keep;
            for (int q = 1 - pe; q < n; q += 2) {
                const int left = q - 1 >= 0 ? a[q - 1] : -1, right = q + 1 < n ? a[q + 1] : -1;
                sv.push_back(make_int4(a[q], left, right, i));
                // new couplings reach the layers two positions away, where they exist
                fl.push_back((q - 2 >= 0 ? 1 : 0) | (q + 2 < n ? 2 : 0));
                keep.push_back(a[q]);
            }
            act[i] = keep;
        }
        if (!any) break;
        b.elim_off.push_back((int)el.size());
        b.surv_off.push_back((int)sv.size());
        b.nlev++;
    }
    for (int i = 0; i < c->ncols; i++) top.push_back(make_int2(act[i][0], i));
    if (c->ncols == 1 && b.nlev >= 2) {   // level 0 reads the input and places the rest: not fused
        const int e0 = b.elim_off[b.nlev - 1], s0 = b.surv_off[b.nlev - 1];
        b.tail = (int)el.size() - e0 == 1 && (int)sv.size() - s0 == 1 && sv[s0].x == top[0].x && sv[s0].y < 0 &&
                 sv[s0].z == el[e0].x && el[e0].y == top[0].x && el[e0].z < 0;
    }
    b.nsurv = (int)sv.size();
    b.ntop = (int)top.size();
    auto put = [](auto **dst, const auto &vec) -> int {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        const size_t n = vec.empty() ? 1 : vec.size();
        if (hipMalloc((void **)dst, sizeof(T) * n) != hipSuccess) return VK_ENOMEM;
        if (!vec.empty() && hipMemcpy(*dst, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice) != hipSuccess)
            return VK_ENOMEM;
        return 0;
    };
    int r = put(&b.d_elim, el);
    r = r ? r : put(&b.d_surv, sv);
    r = r ? r : put(&b.d_sflag, fl);
    r = r ? r : put(&b.d_top, top);
    return r;
}

// Chain orders of the blocked chain factorisations (vk_columns::d_chain_order).
// Chain 2c runs the top m = nz/2 layers of column c, chain 2c+1 the bottom
// nz-1-m; the ranks sort them longest first (ties by index).  ncu: compute
// units of the device.
static int chain_orders(const std::vector<int> &off, int ncu, std::vector<int> &lpt, std::vector<int> &pair) {
    const int ncols = (int)off.size() - 1, n = 2 * ncols;
    std::vector<int> len(n);
    for (int c = 0; c < ncols; c++) {
        const int nz = off[c + 1] - off[c], m = nz / 2;
        len[2 * c] = m;
        len[2 * c + 1] = nz - 1 - m;
    }
    lpt.resize(n);
    for (int i = 0; i < n; i++) lpt[i] = i;
    std::stable_sort(lpt.begin(), lpt.end(), [&](int a, int b) { return len[a] > len[b]; });
    pair = lpt;
    const int H = ncu;
    if (H > 0 && n > H && n <= 2 * H) {
        // the longest H-k chains alone on their CUs, the next k paired with
        // the k shortest, longest with shortest
        const int k = n - H;
        for (int j = 0; j < k; j++) pair[j] = lpt[H - k + j];
        for (int j = k; j < H; j++) pair[j] = lpt[j - k];
        for (int j = 0; j < k; j++) pair[H + j] = lpt[n - 1 - j];
    } else if (H > 0 && n > 2 * H) {
        // the longest 2H chains paired among themselves; the rest follow, longest first
        for (int j = 0; j < H; j++) {
            pair[j] = lpt[j];
            pair[H + j] = lpt[2 * H - 1 - j];
        }
    }
    // both must be permutations of the chains
    for (const std::vector<int> *o : {&lpt, &pair}) {
        std::vector<char> seen(n, 0);
        for (int b = 0; b < n; b++) {
            const int v = (*o)[b];
            if (v < 0 || v >= n || seen[v]) return VK_EINVAL;
            seen[v] = 1;
        }
    }
    return 0;
}

extern "C" int vk_columns_create(int ncols, const int *nz, vk_columns **out) {
    if (ncols <= 0 || !nz || !out) return VK_EINVAL;
    vk_columns *c = new vk_columns();
    c->ncols = ncols;
    c->off.resize(ncols + 1);
    c->off[0] = 0;
    c->max_nz = 0;
    for (int i = 0; i < ncols; i++) {
        if (nz[i] < 2) { delete c; return VK_EINVAL; }
        c->off[i + 1] = c->off[i] + nz[i];
        c->max_nz = nz[i] > c->max_nz ? nz[i] : c->max_nz;
    }
    c->L = c->off[ncols];
    std::vector<int> lc(c->L);
    for (int i = 0; i < ncols; i++)
        for (int l = c->off[i]; l < c->off[i + 1]; l++) lc[l] = i;
    if (hipMalloc((void **)&c->d_off, sizeof(int) * (ncols + 1)) != hipSuccess) { delete c; return VK_ENOMEM; }
    if (hipMalloc((void **)&c->d_layer_col, sizeof(int) * c->L) != hipSuccess) {
        (void)hipFree(c->d_off); delete c; return VK_ENOMEM;
    }
    VK_CHECK(hipMemcpy(c->d_off, c->off.data(), sizeof(int) * (ncols + 1), hipMemcpyHostToDevice));
    VK_CHECK(hipMemcpy(c->d_layer_col, lc.data(), sizeof(int) * c->L, hipMemcpyHostToDevice));
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                 hipSuccess)
        ncu = 0;
    std::vector<int> ord[2];
    int r = chain_orders(c->off, ncu, ord[0], ord[1]);
    for (int v = 0; v < 2 && !r; v++) {
        if (hipMalloc((void **)&c->d_chain_order[v], sizeof(int) * 2 * ncols) != hipSuccess) r = VK_ENOMEM;
        else if (hipMemcpy(c->d_chain_order[v], ord[v].data(), sizeof(int) * 2 * ncols, hipMemcpyHostToDevice) !=
                 hipSuccess)
            r = VK_ENOMEM;
    }
    if (r) { vk_columns_destroy(c); return r; }
    r = bcr_schedule(c);
    if (r) { vk_columns_destroy(c); return r; }
    *out = c;
    return 0;
}

extern "C" int vk_columns_destroy(vk_columns *c) {
    if (!c) return 0;
    (void)hipFree(c->d_off);
    (void)hipFree(c->d_layer_col);
    for (int *o : c->d_chain_order)
        if (o) (void)hipFree(o);
    if (c->bcr.d_elim) (void)hipFree(c->bcr.d_elim);
    if (c->bcr.d_surv) (void)hipFree(c->bcr.d_surv);
    if (c->bcr.d_sflag) (void)hipFree(c->bcr.d_sflag);
    if (c->bcr.d_top) (void)hipFree(c->bcr.d_top);
    delete c;
    return 0;
}

extern "C" int vk_columns_layers(const vk_columns *c, int *nlayers) {
    if (!c || !nlayers) return VK_EINVAL;
    *nlayers = c->L;
    return 0;
}

// ---------------------------------------------------------------------------
// integrator
// ---------------------------------------------------------------------------

extern "C" int vk_integrator_create(const vk_plan *plan, const vk_columns *cols, int history, vk_integrator **out) {
    if (!plan || !cols || !out || history < 0) return VK_EINVAL;
    vk_integrator *it = new vk_integrator();
    it->plan = plan;
    it->cols = cols;
    it->history = history;
    it->tcap = 0;
    const size_t L = cols->L, ni = plan->ni, nc = cols->ncols;
    auto &v = it->allocs;
    int r = 0;
    r = r ? r : dalloc(v, &it->D, L * ni * ni);
    // the blocks' structural zeros, written once: the lhs assembly of an
    // attempt writes only the entries that can be nonzero (nothing else
    // writes D)
    if (!r && hipMemset(it->D, 0, sizeof(double) * L * ni * ni) != hipSuccess) r = VK_ENOMEM;
    r = r ? r : dalloc(v, &it->Dinv, L * ni * ni);
    r = r ? r : dalloc(v, &it->Lsub, L * ni);
    r = r ? r : dalloc(v, &it->Usup, L * ni);
    r = r ? r : dalloc(v, &it->b1, L * ni);
    r = r ? r : dalloc(v, &it->x1, L * ni);
    r = r ? r : dalloc(v, &it->b2, L * ni);
    r = r ? r : dalloc(v, &it->x2, L * ni);
    r = r ? r : dalloc(v, &it->work, L * ni);
    r = r ? r : dalloc(v, &it->ztmp, L * ni);
    r = r ? r : dalloc(v, &it->yk2, L * ni);
    r = r ? r : dalloc(v, &it->part, L * NP);
    r = r ? r : dalloc(v, &it->ysum, L);
    r = r ? r : dalloc(v, &it->info, nc);
    r = r ? r : dalloc(v, &it->flags, nc);
    r = r ? r : dalloc(v, &it->colscr, nc * 16);
    if (!r && history > 0) {
        r = dalloc(v, &it->yhist, (size_t)history * L * ni);
        it->tcap = 0;
    }
    if (r) { vk_integrator_destroy(it); return r; }
    if (hipEventCreate(&it->ev0) != hipSuccess || hipEventCreate(&it->ev1) != hipSuccess) {
        vk_integrator_destroy(it);
        return VK_ENOMEM;
    }
    it->timing = 0;
    it->use_graph = 0;
    it->gexec = nullptr;
    it->gstream = nullptr;
    it->factor_ms = 0.0;
    it->factor_launches = 0;
    it->photo = nullptr;
    it->photo_ini_frq = it->photo_final_frq = 0;
    *out = it;
    return 0;
}

// the cyclic reduction workspace, allocated the first time the plan selects it
static int ensure_bcr(vk_integrator *it) {
    if (it->plan->factor_variant != VK_FACTOR_BCR || it->bcr.A) return 0;
    const size_t nn = (size_t)it->plan->ni * it->plan->ni, L = (size_t)it->cols->L;
    const size_t ns = (size_t)it->cols->bcr.nsurv + 1;
    auto &v = it->allocs;
    int r = dalloc(v, &it->bcr.A, L * nn);
    r = r ? r : dalloc(v, &it->bcr.B, L * nn);
    r = r ? r : dalloc(v, &it->bcr.C, L * nn);
    r = r ? r : dalloc(v, &it->bcr.P, ns * nn);
    r = r ? r : dalloc(v, &it->bcr.Q, ns * nn);
    r = r ? r : dalloc(v, &it->bcr.XB, L * nn);
    r = r ? r : dalloc(v, &it->bcr.XC, L * nn);
    return r;
}

static int ensure_thist(vk_integrator *it, int count_max) {
    const int need = count_max + 2;
    if (it->history <= 0 || it->tcap >= need) return 0;
    double *t = nullptr;
    if (hipMalloc((void **)&t, sizeof(double) * (size_t)it->cols->ncols * need) != hipSuccess) return VK_ENOMEM;
    VK_CHECK(hipMemset(t, 0, sizeof(double) * (size_t)it->cols->ncols * need));
    if (it->thist) {
        // keep the recorded times
        for (int c = 0; c < it->cols->ncols; c++)
            VK_CHECK(hipMemcpy(t + (size_t)c * need, it->thist + (size_t)c * it->tcap, sizeof(double) * it->tcap,
                               hipMemcpyDeviceToDevice));
        (void)hipFree(it->thist);
    }
    it->thist = t;
    it->tcap = need;
    return 0;
}

extern "C" int vk_integrator_destroy(vk_integrator *it) {
    if (!it) return 0;
    if (it->gexec) (void)hipGraphExecDestroy(it->gexec);
    if (it->gstream) (void)hipStreamDestroy(it->gstream);
    for (void *q : it->allocs) (void)hipFree(q);
    if (it->thist) (void)hipFree(it->thist);
    if (it->ev0) (void)hipEventDestroy(it->ev0);
    if (it->ev1) (void)hipEventDestroy(it->ev1);
    delete it;
    return 0;
}

extern "C" int vk_integrator_set_timing(vk_integrator *it, int on) {
    if (!it) return VK_EINVAL;
    it->timing = on ? 1 : 0;
    return 0;
}

extern "C" int vk_integrator_stats(const vk_integrator *it, double *factor_ms, long long *launches) {
    if (!it) return VK_EINVAL;
    if (factor_ms) *factor_ms = it->factor_ms;
    if (launches) *launches = it->factor_launches;
    return 0;
}

// one Ros2 attempt on all active columns (op.py:2890-2932): k1 in it->x1, k2 in it->x2
// escape is active when a species is marked and there is a top-flux array
static bool esc_on(const vk_atm *a) {
    unsigned m = 0;
    for (int w = 0; w < VK_SPECIES_WORDS; w++) m |= a->esc[w];
    return m != 0 && a->top_flux;
}

static int ros2_attempt(vk_integrator *it, const vk_atm *atm, const vk_params *prm, const double *k, double *y,
                        double *ymix, vk_colstate *st, int clip, bool timed, hipStream_t s) {
    const vk_plan *p = it->plan;
    const vk_columns *cl = it->cols;
    const int ni = p->ni;
    const size_t L = cl->L;
    // dt per column lives in the state struct; gather it contiguously
    double *dt = it->colscr;
    hipLaunchKernelGGL(k_gather_dt, dim3((cl->ncols + 127) / 128), dim3(128), 0, s, cl->ncols, st, dt, it->info);
    int r = vk_launch_rhs(p, cl, atm, y, k, 3, prm->use_moldiff, nullptr, nullptr, st, it->ysum, it->b1, s);
    if (r) return r;
    // the right-hand side above has left the layer sums of y in it->ysum
    r = vk_launch_jac_lhs(p, cl, atm, y, k, atm->M, 1, prm->use_moldiff, dt, st, it->ysum, it->D, it->Lsub, it->Usup,
                          s, true, false);
    if (r) return r;
    if (timed) VK_CHECK(hipEventRecord(it->ev0, s));
    r = vk_launch_factor(p, cl, it->D, it->Lsub, it->Usup, it->Dinv, it->info, st, &it->bcr, s, true);
    if (r) return r;
    if (timed) VK_CHECK(hipEventRecord(it->ev1, s));
    r = vk_launch_solve_refined(p, cl, it->D, it->Dinv, it->Lsub, it->Usup, it->b1, it->x1, it->work, it->ztmp, prm->refine,
                                st, &it->bcr, s);   // k1
    if (r) return r;
    const int n = (int)(L * ni);
    hipLaunchKernelGGL(k_yk2, dim3((n + 255) / 256), dim3(256), 0, s, n, ni, cl->d_layer_col, y, it->x1, st, it->yk2);
    r = vk_launch_rhs(p, cl, atm, it->yk2, k, 3, prm->use_moldiff, it->x1, dt, st, it->ysum, it->b2,
                      s);   // f(yk2) - 2/(r h) k1
    if (r) return r;
    r = vk_launch_solve_refined(p, cl, it->D, it->Dinv, it->Lsub, it->Usup, it->b2, it->x2, it->work, it->ztmp, prm->refine,
                                st, &it->bcr, s);   // k2
    if (r) return r;
    const int mask_bot = (atm->use_botflux && atm->bot_flux && atm->bot_vdep) || atm->fix_bot;
    hipLaunchKernelGGL(k_finalize, dim3(cl->L), dim3(256), 0, s, ni, p->natom, cl->d_off, cl->d_layer_col,
                       p->atom_mat, y, ymix, it->x1, it->x2, it->yk2, *prm, clip, st, atm->fix_bot, atm->M, mask_bot,
                       y, ymix, it->part);
    VK_CHECK(hipGetLastError());
    return 0;
}

static void accumulate_factor_time(vk_integrator *it) {
    float ms = 0.f;
    if (hipEventSynchronize(it->ev1) == hipSuccess && hipEventElapsedTime(&ms, it->ev0, it->ev1) == hipSuccess) {
        it->factor_ms += ms;
        it->factor_launches += 1;
    }
}

extern "C" int vk_ros2_solver(vk_integrator *it, const vk_atm *atm, const vk_params *prm, const double *d_k,
                              double *d_y, double *d_ymix, vk_colstate *d_st, void *stream) {
    if (!it || !atm || !prm || !d_k || !d_y || !d_ymix || !d_st) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    int r = ensure_bcr(it);
    if (r) return r;
    r = ros2_attempt(it, atm, prm, d_k, d_y, d_ymix, d_st, 0, false, s);
    if (r) return r;
    const size_t dsm = (size_t)it->cols->max_nz * NP * sizeof(double);
    const int stage = dsm <= DECIDE_STAGE_MAX;
    hipLaunchKernelGGL(k_decide, dim3(it->cols->ncols), dim3(DECIDE_NT), stage ? dsm : 0, s, it->plan->natom,
                       it->cols->d_off, it->part, *prm, 0, nullptr, nullptr, nullptr, d_st, nullptr, nullptr, 0, stage);
    VK_CHECK(hipGetLastError());
    return 0;
}

// one full loop iteration (attempt + decision + bookkeeping) on stream s
static int loop_body(vk_integrator *it, const vk_atm *atm, const vk_geom *g, const vk_params *prm,
                     const double *d_k, double *d_y, double *d_ymix, double *d_y_prev, double *d_atom_ini,
                     double *d_atom_loss, double *d_atom_loss_prev, vk_colstate *d_st, bool timed, hipStream_t s) {
    const vk_plan *p = it->plan;
    const vk_columns *cl = it->cols;
    int r = 0;
    if (it->photo) {
        // compute_tau / flux / J at the top of the loop (op.py:818-830); the
        // photolysis rows of k are the ones the attempt then reads
        r = vk_launch_photo(it->photo, p, cl, d_y, d_ymix, g->dz, const_cast<double *>(d_k), d_st, 0,
                            it->photo_ini_frq, it->photo_final_frq, prm->yconv_min, s);
        if (r) return r;
    }
    r = ros2_attempt(it, atm, prm, d_k, d_y, d_ymix, d_st, 1, timed, s);
    if (r) return r;
    const size_t dsm = (size_t)cl->max_nz * NP * sizeof(double);
    const int stage = dsm <= DECIDE_STAGE_MAX;
    hipLaunchKernelGGL(k_decide, dim3(cl->ncols), dim3(DECIDE_NT), stage ? dsm : 0, s, p->natom, cl->d_off, it->part,
                       *prm, 1, d_atom_ini, d_atom_loss, d_atom_loss_prev, d_st, it->flags, it->thist, it->tcap, stage);
    if (prm->use_update_mu)
        hipLaunchKernelGGL(k_update_mu, dim3(cl->ncols), dim3(256), 0, s, p->ni, cl->d_off, p->ms, d_ymix, *atm,
                           *g, it->flags, 0, d_st);
    if (esc_on(atm))
        hipLaunchKernelGGL(k_phi_esc, dim3(cl->ncols), dim3(64), 0, s, p->ni, cl->d_off, p->ms, d_y, *atm, it->flags,
                           0, d_st);
    hipLaunchKernelGGL(k_apply, dim3(cl->L), dim3(128), 0, s, p->ni, cl->d_layer_col, it->flags, atm->M, d_ymix, d_y,
                       d_y_prev, d_st, it->yhist, it->history, (size_t)cl->L * p->ni);
    hipLaunchKernelGGL(k_conv, dim3(cl->ncols), dim3(256), 0, s, p->ni, cl->d_off, *prm, *atm, d_y, d_ymix,
                       it->flags, d_st, it->yhist, it->history, (size_t)cl->L * p->ni, it->thist, it->tcap);
    VK_CHECK(hipGetLastError());
    return 0;
}

// everything the captured graph bakes in
struct GraphKey {
    vk_atm atm;
    vk_geom geom;
    vk_params prm;
    const void *ptr[9];
    int factor_variant;
    const void *photo;
    int photo_frq[2];
};

extern "C" int vk_integrate(vk_integrator *it, const vk_atm *atm, const vk_geom *geom, const vk_params *prm,
                            const double *d_k, double *d_y, double *d_ymix, double *d_y_prev, double *d_atom_ini,
                            double *d_atom_loss, double *d_atom_loss_prev, vk_colstate *d_st, int n_attempts,
                            void *stream) {
    if (!it || !atm || !prm || !d_k || !d_y || !d_ymix || !d_y_prev || !d_atom_ini || !d_atom_loss ||
        !d_atom_loss_prev || !d_st || n_attempts < 0)
        return VK_EINVAL;
    if ((prm->use_update_mu || it->photo) && !geom) return VK_EINVAL;
    if (prm->use_conv && it->history > 0 && it->history < prm->conv_step + 1) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    int r = ensure_thist(it, prm->count_max);
    if (!r) r = ensure_bcr(it);
    if (r) return r;
    vk_geom g{};
    if (geom) g = *geom;
    if (!it->use_graph || it->timing) {
        for (int a = 0; a < n_attempts; a++) {
            r = loop_body(it, atm, &g, prm, d_k, d_y, d_ymix, d_y_prev, d_atom_ini, d_atom_loss, d_atom_loss_prev,
                          d_st, it->timing && a == n_attempts - 1, s);
            if (r) return r;
        }
        if (n_attempts > 0 && it->timing) accumulate_factor_time(it);
        return 0;
    }
    // graph path: one captured loop iteration replayed per attempt on the
    // integrator's own stream, ordered after and before the caller's stream
    GraphKey key;
    memset(&key, 0, sizeof(key));
    key.atm = *atm;
    key.geom = g;
    key.prm = *prm;
    const void *ptrs[9] = {d_k, d_y, d_ymix, d_y_prev, d_atom_ini, d_atom_loss, d_atom_loss_prev, d_st, it->thist};
    memcpy(key.ptr, ptrs, sizeof(ptrs));
    key.factor_variant = it->plan->factor_variant;
    key.photo = it->photo;
    key.photo_frq[0] = it->photo_ini_frq;
    key.photo_frq[1] = it->photo_final_frq;
    if (!it->gstream) VK_CHECK(hipStreamCreateWithFlags(&it->gstream, hipStreamNonBlocking));
    if (!it->gexec || it->gkey.size() != sizeof(key) || memcmp(it->gkey.data(), &key, sizeof(key)) != 0) {
        if (it->gexec) { (void)hipGraphExecDestroy(it->gexec); it->gexec = nullptr; }
        hipGraph_t graph = nullptr;
        VK_CHECK(hipStreamBeginCapture(it->gstream, hipStreamCaptureModeThreadLocal));
        r = loop_body(it, atm, &g, prm, d_k, d_y, d_ymix, d_y_prev, d_atom_ini, d_atom_loss, d_atom_loss_prev, d_st,
                      false, it->gstream);
        hipError_t e = hipStreamEndCapture(it->gstream, &graph);
        if (r) { if (graph) (void)hipGraphDestroy(graph); return r; }
        VK_CHECK(e);
        e = hipGraphInstantiate(&it->gexec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        VK_CHECK(e);
        it->gkey.assign((const char *)&key, (const char *)&key + sizeof(key));
    }
    VK_CHECK(hipEventRecord(it->ev0, s));
    VK_CHECK(hipStreamWaitEvent(it->gstream, it->ev0, 0));
    for (int a = 0; a < n_attempts; a++) VK_CHECK(hipGraphLaunch(it->gexec, it->gstream));
    VK_CHECK(hipEventRecord(it->ev1, it->gstream));
    VK_CHECK(hipStreamWaitEvent(s, it->ev1, 0));
    return 0;
}

extern "C" int vk_conv(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const vk_params *prm,
                       const double *d_y, const double *d_ymix, const double *d_yhist, int H, const double *d_thist,
                       int tcap, vk_colstate *d_st, void *stream) {
    if (!plan || !cols || !atm || !prm || !d_y || !d_ymix || !d_st) return VK_EINVAL;
    if (prm->use_conv && (!d_yhist || !d_thist || H < 1 || tcap < 1)) return VK_EINVAL;
    hipLaunchKernelGGL(k_conv, dim3(cols->ncols), dim3(256), 0, (hipStream_t)stream, plan->ni, cols->d_off, *prm,
                       *atm, d_y, d_ymix, nullptr, d_st, d_yhist, prm->use_conv ? H : 0, (size_t)cols->L * plan->ni,
                       d_thist, tcap);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_integrator_history(vk_integrator *it, int write, int count, int q0, double *h_t, double *h_y) {
    if (!it || it->history <= 0 || count < 0 || q0 < 0 || q0 > count || count - q0 > it->history) return VK_EINVAL;
    VK_CHECK(hipDeviceSynchronize());
    if (write) {
        int r = ensure_thist(it, count);
        if (r) return r;
    } else if (count > it->tcap) {
        return VK_EINVAL;
    }
    const hipMemcpyKind dir = write ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    const int nc = it->cols->ncols;
    const size_t Lni = (size_t)it->cols->L * it->plan->ni;
    if (h_t && count > 0) {
        for (int c = 0; c < nc; c++) {
            double *dv = it->thist + (size_t)c * it->tcap, *hv = h_t + (size_t)c * count;
            VK_CHECK(write ? hipMemcpy(dv, hv, sizeof(double) * count, dir) : hipMemcpy(hv, dv, sizeof(double) * count, dir));
        }
    }
    if (h_y) {
        for (int q = q0; q < count; q++) {
            double *dv = it->yhist + (size_t)(q % it->history) * Lni, *hv = h_y + (size_t)(q - q0) * Lni;
            VK_CHECK(write ? hipMemcpy(dv, hv, sizeof(double) * Lni, dir) : hipMemcpy(hv, dv, sizeof(double) * Lni, dir));
        }
    }
    return 0;
}

extern "C" int vk_integrator_set_photo(vk_integrator *it, vk_photo *photo, int ini_frq, int final_frq) {
    if (!it || (photo && (ini_frq < 1 || final_frq < 1))) return VK_EINVAL;
    it->photo = photo;
    it->photo_ini_frq = ini_frq;
    it->photo_final_frq = final_frq;
    return 0;
}

extern "C" int vk_integrator_set_graph(vk_integrator *it, int on) {
    if (!it) return VK_EINVAL;
    it->use_graph = on ? 1 : 0;
    return 0;
}

extern "C" int vk_update_phi_esc(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const double *d_y,
                                 void *stream) {
    if (!plan || !cols || !atm || !d_y || !atm->Dzz || !atm->Hp || !atm->g || !atm->Tco) return VK_EINVAL;
    if (!esc_on(atm)) return 0;
    hipLaunchKernelGGL(k_phi_esc, dim3(cols->ncols), dim3(64), 0, (hipStream_t)stream, plan->ni, cols->d_off, plan->ms,
                       d_y, *atm, nullptr, 1, nullptr);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_update_mu_dz(const vk_plan *plan, const vk_columns *cols, const vk_atm *atm, const vk_geom *geom,
                               const double *d_ymix, const vk_colstate *d_st, int force, void *stream) {
    if (!plan || !cols || !atm || !geom || !d_ymix) return VK_EINVAL;
    if (!force) return VK_EINVAL;
    hipLaunchKernelGGL(k_update_mu, dim3(cols->ncols), dim3(256), 0, (hipStream_t)stream, plan->ni, cols->d_off,
                       plan->ms, d_ymix, *atm, *geom, nullptr, 1, d_st);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_atom_sums(const vk_plan *plan, const vk_columns *cols, const double *d_y, double *d_atom_sum,
                            void *stream) {
    if (!plan || !cols || !d_y || !d_atom_sum) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    double *part = nullptr;
    VK_CHECK(hipMallocAsync((void **)&part, sizeof(double) * (size_t)cols->L * NP, s));
    hipLaunchKernelGGL(k_atom_layer, dim3(cols->L), dim3(64), 0, s, plan->ni, plan->natom, plan->atom_mat, d_y, part);
    hipLaunchKernelGGL(k_atom_col, dim3(cols->ncols), dim3(64), 0, s, plan->natom, cols->d_off, part, d_atom_sum);
    VK_CHECK(hipGetLastError());
    VK_CHECK(hipFreeAsync(part, s));
    return 0;
}
