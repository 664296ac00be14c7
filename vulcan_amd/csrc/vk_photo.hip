// Photochemical radiative transfer on the device: optical depth, two-stream
// diffuse flux and photolysis rates (ODESolver.compute_tau / compute_flux /
// compute_J, op.py:2580-2787), refreshed every update_photo_frq steps of the
// integration loop (op.py:818-830).
//
// Layout per column c with nz layers (layers off[c] .. off[c+1]-1 of the
// stack): the staggered level arrays (tau, direct beam, diffuse up / down
// fluxes; nz + 1 levels) occupy rows off[c] + c .. off[c] + c + nz of
// [L + ncols][nbin]; the layer arrays (single-scattering albedo, actinic
// flux) rows off[c] .. of [L][nbin].  One thread per (column, wavelength
// bin) sweeps the layers: every bin is independent, and consecutive threads
// read consecutive bins of a row (coalesced).
#include <atomic>

#include "vk_internal.h"
#include "vk_device.h"

namespace {

// the tables and state the kernels read (plain data, passed by value)
struct PhotoDev {
    int nbin, nabs, nscat, nbranch, nT, nTJ, din12;
    double dbin1, dbin2, cos_sl, edd, f_diurnal, flux_atol;
    int ncols, L;
    int *abs_sp, *abs_T, *scat_sp, *br_col, *br_T;
    double *bins, *sflux_top, *cross, *cross_T, *cross_scat, *cross_J, *cross_J_T, *hc_bin;
    double *tau, *sflux, *dflux_u, *dflux_d, *w0, *aflux;
    unsigned long long *change;   // per column: max relative change (bits of a positive double)
    int *last, *switched, *due;   // per column: count of the last update, final frequency on, update now
};

}  // namespace

struct vk_photo {
    PhotoDev d;
    std::vector<void *> allocs;
};

namespace {

constexpr int PT = 256;

// per column: is an update due at this attempt (force, or the first attempt
// of a step whose count is a multiple of the frequency in force, op.py:818-827)
__global__ void k_photo_due(int ncols, const vk_colstate *__restrict__ st, int force, int ini_frq, int final_frq,
                            double yconv_min, int *__restrict__ last, int *__restrict__ switched,
                            int *__restrict__ due, unsigned long long *__restrict__ change) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    change[c] = 0ull;
    // a forced update (the set-up call before the loop, vulcan.py:172-174)
    // leaves the loop's own schedule alone: its first step updates again
    if (force) { due[c] = 1; return; }
    const vk_colstate s = st[c];
    if (s.done) { due[c] = 0; return; }
    if (!switched[c] && s.longdy < yconv_min * 10. && s.longdydt < 1.e-6) switched[c] = 1;
    const int frq = switched[c] ? final_frq : ini_frq;
    const int d = (s.count % frq == 0) && last[c] != s.count;
    if (d) last[c] = s.count;
    due[c] = d;
}

// optical depth (op.py:2580-2601) and single-scattering albedo (op.py:2627-2639)
__global__ void __launch_bounds__(PT) k_photo_tau(PhotoDev ph, int ni, const int *__restrict__ off,
                                                  const double *__restrict__ y, const double *__restrict__ ymix,
                                                  const double *__restrict__ dz) {
    const int c = blockIdx.y, b = blockIdx.x * PT + threadIdx.x;
    if (!ph.due[c] || b >= ph.nbin) return;
    const int o = off[c], nz = off[c + 1] - o, nb = ph.nbin;
    const size_t top = (size_t)(o + c + nz) * nb + b;
    ph.tau[top] = 0.0;
    double above = 0.0;
    for (int j = nz - 1; j >= 0; j--) {
        const int l = o + j;
        const double *yl = y + (size_t)l * ni, *ml = ymix + (size_t)l * ni;
        const double dzl = dz[l];
        double t = 0.0, tabs = 0.0, tscat = 0.0;
        for (int a = 0; a < ph.nabs; a++) {
            const int sp = ph.abs_sp[a], tt = ph.abs_T[a];
            const double cr = ph.cross[(size_t)a * nb + b];
            const double crt = tt >= 0 ? ph.cross_T[((size_t)tt * ph.L + l) * nb + b] : cr;
            t += yl[sp] * dzl * crt;
            tabs += ml[sp] * cr;
        }
        for (int q = 0; q < ph.nscat; q++) {
            const int sp = ph.scat_sp[q];
            const double cr = ph.cross_scat[(size_t)q * nb + b];
            t += yl[sp] * dzl * cr;
            tscat += ml[sp] * cr;
        }
        above = t + above;
        ph.tau[(size_t)(o + c + j) * nb + b] = above;
        double w = tscat / (tabs + tscat);
        if (w != w) w = 0.0;                       // nan_to_num: 0/0
        w = fmin(w, 1. - 1.E-8);
        ph.w0[(size_t)l * nb + b] = w;
    }
}

// two-stream coefficients of layer l (op.py:2648-2677, isotropic scattering)
struct TwoStream {
    double chi, xi, phi, i_u, i_d;
};

__device__ inline TwoStream two_stream(const PhotoDev &ph, double mu, double w, double dtau, double dir_lo,
                                       double dir_hi) {
    const double edd = ph.edd;
    const double s1 = sqrt(1. - w);
    const double tran = exp(-1. / edd * s1 * dtau);
    const double zp = 0.5 * (1. + s1), zm = 0.5 * (1. - s1);
    const double ll = -1. * w / (1. / (mu * mu) - 1. / (edd * edd) * (1. - w));
    const double gp = 0.5 * (ll * (1. / edd + 1. / mu)), gm = 0.5 * (ll * (1. / edd - 1. / mu));
    TwoStream r;
    r.chi = zm * zm * (tran * tran) - zp * zp;
    r.xi = zp * zm * (1. - tran * tran);
    r.phi = (zm * zm - zp * zp) * tran;
    r.i_u = r.phi * gp * dir_lo - (r.xi * gm + r.chi * gp) * dir_hi;
    r.i_d = r.phi * gm * dir_hi - (r.chi * gm + r.xi * gp) * dir_lo;
    return r;
}

// direct beam, diffuse fluxes by a downward then an upward sweep, actinic
// flux and its relative change (op.py:2603-2738)
__global__ void __launch_bounds__(PT) k_photo_flux(PhotoDev ph, const int *__restrict__ off) {
    const int c = blockIdx.y;
    if (!ph.due[c]) return;                          // uniform over the workgroup
    // threads past the last bin take part in the reduction below only
    const int b = min(blockIdx.x * PT + threadIdx.x, ph.nbin - 1);
    const bool mine = blockIdx.x * PT + threadIdx.x < ph.nbin;
    double mx = 0.0;
    if (mine) {
        const int o = off[c], nz = off[c + 1] - o, nb = ph.nbin;
        const double mu = -1. * ph.cos_sl, top = ph.sflux_top[b];
        const size_t base = (size_t)(o + c) * nb + b;   // level 0 of the column
        double *tau = ph.tau + base, *sf = ph.sflux + base, *du = ph.dflux_u + base, *dd = ph.dflux_d + base;
        for (int j = 0; j <= nz; j++) sf[(size_t)j * nb] = top * exp(-1. * tau[(size_t)j * nb] / ph.cos_sl);
        // downward: dflux_d[j] from dflux_d[j+1] and the previous update's dflux_u[j]
        for (int j = nz - 1; j >= 0; j--) {
            const size_t lj = (size_t)j * nb;
            const TwoStream t = two_stream(ph, mu, ph.w0[(size_t)(o + j) * nb + b], tau[lj] - tau[lj + nb],
                                           sf[lj] * ph.cos_sl, sf[lj + nb] * ph.cos_sl);
            dd[lj] = 1. / t.chi * (t.phi * dd[lj + nb] - t.xi * du[lj] + t.i_d / mu);
        }
        // upward
        for (int j = 1; j <= nz; j++) {
            const size_t lj = (size_t)(j - 1) * nb;
            const TwoStream t = two_stream(ph, mu, ph.w0[(size_t)(o + j - 1) * nb + b], tau[lj] - tau[lj + nb],
                                           sf[lj] * ph.cos_sl, sf[lj + nb] * ph.cos_sl);
            du[lj + nb] = 1. / t.chi * (t.phi * du[lj] - t.xi * dd[lj + nb] + t.i_u / mu);
        }
        // actinic flux and the largest relative change over the bins above flux_atol
        for (int j = 0; j < nz; j++) {
            const size_t lj = (size_t)j * nb;
            const double ave = 0.5 * (sf[lj] + sf[lj + nb]);
            const double tot = ave + 0.5 * (du[lj] + du[lj + nb] + dd[lj + nb] + dd[lj]) / ph.edd;
            const double af = tot / ph.hc_bin[b];
            double *ap = ph.aflux + (size_t)(o + j) * nb + b;
            const double prev = *ap;
            *ap = af;
            if (af > ph.flux_atol) {
                const double ch = fabs(af - prev) / af;
                if (ch == ch) mx = fmax(mx, ch);
            }
        }
    }
    // positive doubles order as their bit patterns
    __shared__ double red[PT];
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int w = PT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] > 0.0)
        atomicMax(&ph.change[c], (unsigned long long)__double_as_longlong(red[0]));
}

// photolysis rates of every branch (op.py:2751-2787): trapezoidal sums over
// the two uniform bin widths; the branch's rate column of k = J f_diurnal.  One
// workgroup per layer, one wave per branch, the layer's actinic flux in LDS.
__global__ void __launch_bounds__(PT) k_photo_J(PhotoDev ph, const int *__restrict__ layer_col, int nr,
                                                double *__restrict__ k) {
    extern __shared__ double af[];
    const int l = blockIdx.x;
    const int c = layer_col[l];
    if (!ph.due[c]) return;
    const int nb = ph.nbin;
    for (int b = threadIdx.x; b < nb; b += PT) af[b] = ph.aflux[(size_t)l * nb + b];
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int d = ph.din12 >= 0 ? ph.din12 : nb - 1;
    for (int br = w; br < ph.nbranch; br += PT / 64) {
        const int tt = ph.br_T[br];
        const double *cr = tt >= 0 ? ph.cross_J_T + ((size_t)tt * ph.L + l) * nb : ph.cross_J + (size_t)br * nb;
        double s1 = 0.0, s2 = 0.0;
        for (int b = lane; b < d; b += 64) s1 += af[b] * cr[b] * ph.dbin1;
        for (int b = d + lane; b < nb; b += 64) s2 += af[b] * cr[b] * ph.dbin2;
        for (int m = 32; m > 0; m >>= 1) {
            s1 += __shfl_xor(s1, m);
            s2 += __shfl_xor(s2, m);
        }
        if (lane == 0) {
            double J = s1;
            J -= 0.5 * (af[0] * cr[0] + af[d - 1] * cr[d - 1]) * ph.dbin1;
            J += s2;
            J -= 0.5 * (af[d] * cr[d] + af[nb - 1] * cr[nb - 1]) * ph.dbin2;
            k[(size_t)l * nr + ph.br_col[br]] = J * ph.f_diurnal;
        }
    }
}

__global__ void k_photo_change(int ncols, const int *__restrict__ due, const unsigned long long *__restrict__ change,
                               vk_colstate *__restrict__ st) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols || !due[c] || !st) return;
    st[c].aflux_change = __longlong_as_double((long long)change[c]);
}

template <typename T>
int dup(std::vector<void *> &v, T **p, const T *src, size_t n) {
    if (n == 0) n = 1, src = nullptr;
    if (hipMalloc((void **)p, sizeof(T) * n) != hipSuccess) return VK_ENOMEM;
    v.push_back(*p);
    if (src && hipMemcpy(*p, src, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess) return VK_ENOMEM;
    if (!src && hipMemset(*p, 0, sizeof(T) * n) != hipSuccess) return VK_ENOMEM;
    return 0;
}

}  // namespace

extern "C" int vk_photo_create(const vk_plan *plan, const vk_columns *cols, const vk_photo_desc *d, vk_photo **out) {
    if (!plan || !cols || !d || !out || d->nbin < 2 || d->nabs < 0 || d->nscat < 0 || d->nbranch < 0 || d->nT < 0 ||
        d->nTJ < 0)
        return VK_EINVAL;
    if (d->din12 >= d->nbin || d->din12 == 0 || (d->din12 < 0 && d->din12 != -1)) return VK_EINVAL;
    for (int a = 0; a < d->nabs; a++)
        if (d->abs_sp[a] < 0 || d->abs_sp[a] >= plan->ni || d->abs_T[a] >= d->nT) return VK_EINVAL;
    for (int q = 0; q < d->nscat; q++)
        if (d->scat_sp[q] < 0 || d->scat_sp[q] >= plan->ni) return VK_EINVAL;
    for (int r = 0; r < d->nbranch; r++)
        if (d->br_col[r] < 0 || d->br_col[r] >= plan->nr || d->br_T[r] >= d->nTJ) return VK_EINVAL;
    vk_photo *h = new vk_photo();
    PhotoDev *p = &h->d;
    p->nbin = d->nbin; p->nabs = d->nabs; p->nscat = d->nscat; p->nbranch = d->nbranch;
    p->nT = d->nT; p->nTJ = d->nTJ; p->din12 = d->din12;
    p->dbin1 = d->dbin1; p->dbin2 = d->dbin2; p->cos_sl = d->cos_sl; p->edd = d->edd;
    p->f_diurnal = d->f_diurnal; p->flux_atol = d->flux_atol;
    p->ncols = cols->ncols; p->L = cols->L;
    const size_t nb = d->nbin, L = cols->L, lev = (size_t)(cols->L + cols->ncols);
    std::vector<double> hcb(nb);
    for (size_t b = 0; b < nb; b++) hcb[b] = d->hc / d->bins[b];   // hc/bins (op.py:2735)
    auto &v = h->allocs;
    int r = 0;
    r = r ? r : dup(v, &p->abs_sp, d->abs_sp, d->nabs);
    r = r ? r : dup(v, &p->abs_T, d->abs_T, d->nabs);
    r = r ? r : dup(v, &p->scat_sp, d->scat_sp, d->nscat);
    r = r ? r : dup(v, &p->br_col, d->br_col, d->nbranch);
    r = r ? r : dup(v, &p->br_T, d->br_T, d->nbranch);
    r = r ? r : dup(v, &p->bins, d->bins, nb);
    r = r ? r : dup(v, &p->hc_bin, (const double *)hcb.data(), nb);
    r = r ? r : dup(v, &p->sflux_top, d->sflux_top, nb);
    r = r ? r : dup(v, &p->cross, d->cross, (size_t)d->nabs * nb);
    r = r ? r : dup(v, &p->cross_T, d->cross_T, (size_t)d->nT * L * nb);
    r = r ? r : dup(v, &p->cross_scat, d->cross_scat, (size_t)d->nscat * nb);
    r = r ? r : dup(v, &p->cross_J, d->cross_J, (size_t)d->nbranch * nb);
    r = r ? r : dup(v, &p->cross_J_T, d->cross_J_T, (size_t)d->nTJ * L * nb);
    r = r ? r : dup<double>(v, &p->tau, nullptr, lev * nb);
    r = r ? r : dup<double>(v, &p->sflux, nullptr, lev * nb);
    r = r ? r : dup<double>(v, &p->dflux_u, nullptr, lev * nb);
    r = r ? r : dup<double>(v, &p->dflux_d, nullptr, lev * nb);
    r = r ? r : dup<double>(v, &p->w0, nullptr, L * nb);
    r = r ? r : dup<double>(v, &p->aflux, nullptr, L * nb);
    r = r ? r : dup<unsigned long long>(v, &p->change, nullptr, cols->ncols);
    r = r ? r : dup<int>(v, &p->due, nullptr, cols->ncols);
    r = r ? r : dup<int>(v, &p->switched, nullptr, cols->ncols);
    if (!r) {
        std::vector<int> last(cols->ncols, -1);
        r = dup(v, &p->last, (const int *)last.data(), cols->ncols);
    }
    if (r) { vk_photo_destroy(h); return r; }
    *out = h;
    return 0;
}

extern "C" int vk_photo_destroy(vk_photo *p) {
    if (!p) return 0;
    for (void *q : p->allocs) (void)hipFree(q);
    delete p;
    return 0;
}

// tau, fluxes and J for the columns whose update is due (every column with force)
int vk_launch_photo(vk_photo *ph, const vk_plan *plan, const vk_columns *cols, const double *y, const double *ymix,
                    const double *dz, double *k, vk_colstate *st, int force, int ini_frq, int final_frq,
                    double yconv_min, hipStream_t s) {
    const PhotoDev &d = ph->d;
    if (d.ncols != cols->ncols || d.L != cols->L) return VK_EINVAL;
    const int nc = cols->ncols;
    hipLaunchKernelGGL(k_photo_due, dim3((nc + 127) / 128), dim3(128), 0, s, nc, st, force, ini_frq, final_frq,
                       yconv_min, d.last, d.switched, d.due, d.change);
    const dim3 grid((d.nbin + PT - 1) / PT, nc);
    hipLaunchKernelGGL(k_photo_tau, grid, dim3(PT), 0, s, d, plan->ni, cols->d_off, y, ymix, dz);
    hipLaunchKernelGGL(k_photo_flux, grid, dim3(PT), 0, s, d, cols->d_off);
    const size_t smem = sizeof(double) * (size_t)d.nbin;
    hipLaunchKernelGGL(k_photo_J, dim3(cols->L), dim3(PT), smem, s, d, cols->d_layer_col, plan->nr, k);
    hipLaunchKernelGGL(k_photo_change, dim3((nc + 127) / 128), dim3(128), 0, s, nc, d.due, d.change, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_photo_update(vk_photo *p, const vk_plan *plan, const vk_columns *cols, const double *d_y,
                               const double *d_ymix, const double *d_dz, double *d_k, vk_colstate *d_st, int force,
                               int ini_frq, int final_frq, double yconv_min, void *stream) {
    if (!p || !plan || !cols || !d_y || !d_ymix || !d_dz || !d_k) return VK_EINVAL;
    if (!force && (!d_st || ini_frq < 1 || final_frq < 1)) return VK_EINVAL;
    if ((size_t)p->d.nbin * sizeof(double) > 64 * 1024) return VK_ETOOBIG;
    return vk_launch_photo(p, plan, cols, d_y, d_ymix, d_dz, d_k, d_st, force, ini_frq, final_frq, yconv_min,
                           (hipStream_t)stream);
}

extern "C" int vk_photo_read(const vk_photo *p, int what, double *h_out) {
    if (!p || !h_out) return VK_EINVAL;
    const PhotoDev &d = p->d;
    const size_t nb = d.nbin, lev = (size_t)(d.L + d.ncols) * nb, lay = (size_t)d.L * nb;
    const double *src = nullptr;
    size_t n = 0;
    switch (what) {
    case VK_PHOTO_TAU: src = d.tau; n = lev; break;
    case VK_PHOTO_SFLUX: src = d.sflux; n = lev; break;
    case VK_PHOTO_DFLUX_U: src = d.dflux_u; n = lev; break;
    case VK_PHOTO_DFLUX_D: src = d.dflux_d; n = lev; break;
    case VK_PHOTO_AFLUX: src = d.aflux; n = lay; break;
    default: return VK_EINVAL;
    }
    VK_CHECK(hipDeviceSynchronize());
    VK_CHECK(hipMemcpy(h_out, src, sizeof(double) * n, hipMemcpyDeviceToHost));
    return 0;
}
