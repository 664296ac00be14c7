// Block-Thomas factorisation and solves for the Ros2 left-hand side
//
//     lhs = blocktridiag(diag(Lsub_j), D_j, diag(Usup_j))
//
// The off-diagonal blocks are diagonal (diffusion couples a species only to
// itself in the neighbouring layers, op.py:2012-2040), so the Schur update
//     S_j = D_j - diag(Lsub_j) S_{j-1}^-1 diag(Usup_{j-1})
// is elementwise and the dense work is the inversion of each S_j.  That
// inversion is an in-register Gauss-Jordan elimination with partial
// pivoting: a 16 x 16 grid of threads holds the ni x ni block (RPT x CPT
// elements each), one barrier per pivot step, the pivot column travels
// through LDS and the pivot row through intra-wave shuffles (the 16 threads
// sharing a column group are 16 consecutive lanes of one wave).  Register
// indices are static: the pivot column block is a template parameter and
// the (wave-uniform) pivot row is selected by a scalar switch.
//
// The chain over layers is split in two (twisted factorisation): one
// workgroup eliminates upwards from layer 0, another downwards from layer
// nz-1, and the meeting layer nz/2 is inverted last.  Same arithmetic as a
// single chain, half the dependence length; the solves are split the same
// way.
//
// The inverses go to Dinv, stored transposed (Dinv[l][m][i] = S_l^-1[i][m])
// so that the solve sweeps read them coalesced; D is left intact for the
// residuals of iterative refinement.
//
// Sweeps with explicit block inverses, like block elimination without
// pivoting across layers, are less accurate than LAPACK's band LU on
// ill-conditioned blocks (early steps from a far-from-equilibrium state);
// one refinement step with a residual computed in double-double arithmetic
// (k_residual) brings the solution to full FP64 accuracy.
#include "vk_internal.h"

namespace {

constexpr int NT = 256;

template <int N>
__device__ __forceinline__ double pick(const double (&v)[N], int idx) {
    double r = v[0];
#pragma unroll
    for (int q = 1; q < N; q++)
        if (q == idx) r = v[q];
    return r;
}

// Shared state of one Gauss-Jordan inversion.
template <int RPT, int CPT>
struct GJShared {
    double colbuf[2][16 * RPT];
    int pivrow[2];
    int perm[16 * CPT];
    int iperm[16 * RPT];
    int info;
};

// Copies row ra_p of a[][] (ra_p wave-uniform) into out[] with static
// register indices: a scalar switch instead of a per-lane select chain.
template <int RPT, int CPT>
__device__ __forceinline__ void row_of(const double (&a)[RPT][CPT], int ra_p, double (&out)[CPT]) {
    switch (ra_p) {
#define VK_ROW_CASE(R)                                                        \
    case R:                                                                   \
        if (R < RPT) {                                                        \
            _Pragma("unroll") for (int cb = 0; cb < CPT; cb++) out[cb] = a[R < RPT ? R : 0][cb]; \
        }                                                                     \
        break;
        VK_ROW_CASE(0) VK_ROW_CASE(1) VK_ROW_CASE(2) VK_ROW_CASE(3) VK_ROW_CASE(4)
        VK_ROW_CASE(5) VK_ROW_CASE(6) VK_ROW_CASE(7) VK_ROW_CASE(8) VK_ROW_CASE(9)
#undef VK_ROW_CASE
    default:
        break;
    }
}

// One Gauss-Jordan pivot step for pivot column k = 16*KB + kg (KB static).
template <int RPT, int CPT, int KB>
__device__ __forceinline__ void gj_step(double (&a)[RPT][CPT], int ni, int kg, unsigned &used,
                                        GJShared<RPT, CPT> &sh, int tag) {
    const int tid = threadIdx.x;
    const int tr = tid & 15, tc = tid >> 4;
    const int gbase = tid & 63 & ~15;
    const int k = 16 * KB + kg;
    const int b = k & 1;
    if (tc == kg) {
        double best = -1.0;
        int bi = 0x7fffffff;
#pragma unroll
        for (int ra = 0; ra < RPT; ra++) {
            const int i = tr + 16 * ra;
            const double v = a[ra][KB];
            if (i < ni && !((used >> ra) & 1u)) {
                const double av = fabs(v);
                if (bi == 0x7fffffff || av > best) { best = av; bi = i; }
            }
            if (i < ni) sh.colbuf[b][i] = v;
        }
#pragma unroll
        for (int o = 8; o >= 1; o >>= 1) {
            const double ob = __shfl_xor(best, o, 16);
            const int oi = __shfl_xor(bi, o, 16);
            if (oi != 0x7fffffff && (bi == 0x7fffffff || ob > best || (ob == best && oi < bi))) {
                best = ob;
                bi = oi;
            }
        }
        if (tr == 0) {
            sh.pivrow[b] = bi;
            if (!(best > 0.0) && sh.info == 0) sh.info = tag;
        }
    }
    __syncthreads();
    const int p = __builtin_amdgcn_readfirstlane(sh.pivrow[b]);
    const int ra_p = p >> 4, tr_p = p & 15;
    double mine[CPT], rs[CPT];
    row_of<RPT, CPT>(a, ra_p, mine);
#pragma unroll
    for (int cb = 0; cb < CPT; cb++) rs[cb] = __shfl(mine[cb], gbase + tr_p, 64);
    const double inv = 1.0 / sh.colbuf[b][p];
#pragma unroll
    for (int cb = 0; cb < CPT; cb++) rs[cb] *= inv;
    const bool own_k = (tc == kg);
#pragma unroll
    for (int ra = 0; ra < RPT; ra++) {
        const int i = tr + 16 * ra;
        const double f = sh.colbuf[b][i];
        if (ra == ra_p && tr == tr_p) {
#pragma unroll
            for (int cb = 0; cb < CPT; cb++) a[ra][cb] = rs[cb];
            if (own_k) a[ra][KB] = inv;
        } else {
#pragma unroll
            for (int cb = 0; cb < CPT; cb++) a[ra][cb] = fma(-f, rs[cb], a[ra][cb]);
            if (own_k) a[ra][KB] = -f * inv;
        }
    }
    if (tr == tr_p) used |= 1u << ra_p;
    if (tid == 0) sh.perm[k] = p;
}

template <int RPT, int CPT, int KB>
struct GJBlocks {
    static __device__ __forceinline__ void run(double (&a)[RPT][CPT], int ni, unsigned &used,
                                               GJShared<RPT, CPT> &sh, int tag) {
        if (KB >= CPT || 16 * KB >= ni) return;
        const int kend = min(16, ni - 16 * KB);
        for (int kg = 0; kg < kend; kg++) gj_step<RPT, CPT, (KB < CPT ? KB : 0)>(a, ni, kg, used, sh, tag);
        GJBlocks<RPT, CPT, KB + 1>::run(a, ni, used, sh, tag);
    }
};
template <int RPT, int CPT>
struct GJBlocks<RPT, CPT, 10> {
    static __device__ __forceinline__ void run(double (&)[RPT][CPT], int, unsigned &, GJShared<RPT, CPT> &, int) {}
};

// Inverts the ni x ni matrix held in a[][] (row tr + 16 ra, column tc + 16 cb
// per thread) and writes the inverse transposed to X (X[m][i] = inv[i][m]).
// A zero pivot column records `tag` in sh.info (sticky).
template <int RPT, int CPT>
__device__ __forceinline__ void gj_invert(double (&a)[RPT][CPT], int ni, double *__restrict__ X,
                                          GJShared<RPT, CPT> &sh, int tag) {
    const int tid = threadIdx.x;
    const int tr = tid & 15, tc = tid >> 4;
    unsigned used = 0;
    __syncthreads();
    GJBlocks<RPT, CPT, 0>::run(a, ni, used, sh, tag);
    __syncthreads();
    for (int t = tid; t < ni; t += NT) sh.iperm[sh.perm[t]] = t;
    __syncthreads();
    // inv[q_r][p_c] = a[r][c]  ->  X[p_c][q_r]
#pragma unroll
    for (int ra = 0; ra < RPT; ra++) {
        const int i = tr + 16 * ra;
        if (i >= ni) continue;
        const int qi = sh.iperm[i];
#pragma unroll
        for (int cb = 0; cb < CPT; cb++) {
            const int cc = tc + 16 * cb;
            if (cc < ni) X[(size_t)sh.perm[cc] * ni + qi] = a[ra][cb];
        }
    }
    __syncthreads();
}

// a = D_l - diag(cl) Xn^T diag(cn) [- diag(cl2) Xn2^T diag(cn2)] (X stored transposed)
template <int RPT, int CPT>
__device__ __forceinline__ void load_schur(double (&a)[RPT][CPT], int ni, const double *__restrict__ Dl,
                                           const double *cl, const double *Xn, const double *cn,
                                           const double *cl2, const double *Xn2, const double *cn2) {
    const int tr = threadIdx.x & 15, tc = threadIdx.x >> 4;
#pragma unroll
    for (int ra = 0; ra < RPT; ra++) {
        const int i = tr + 16 * ra;
        const double li = (Xn && i < ni) ? cl[i] : 0.0;
        const double li2 = (Xn2 && i < ni) ? cl2[i] : 0.0;
#pragma unroll
        for (int cb = 0; cb < CPT; cb++) {
            const int cc = tc + 16 * cb;
            double v = 0.0;
            if (i < ni && cc < ni) {
                v = Dl[(size_t)i * ni + cc];
                if (Xn) v = v - li * (Xn[(size_t)cc * ni + i] * cn[cc]);
                if (Xn2) v = v - li2 * (Xn2[(size_t)cc * ni + i] * cn2[cc]);
            }
            a[ra][cb] = v;
        }
    }
}

// Twisted block-Thomas: the lower chain (side 0) runs j = 0 .. m-1 with
//   S_j = D_j - diag(Lsub_j) S_{j-1}^-1 diag(Usup_{j-1}),
// the upper chain (side 1) runs j = nz-1 .. m+1 with
//   T_j = D_j - diag(Usup_j) T_{j+1}^-1 diag(Lsub_{j+1}),
// on two workgroups; k_factor_mid then inverts the meeting block
//   M_m = D_m - diag(Lsub_m) S_{m-1}^-1 diag(Usup_{m-1})
//             - diag(Usup_m) T_{m+1}^-1 diag(Lsub_{m+1}),  m = nz/2.
// Same work as one chain, half the dependence length.
template <int RPT, int CPT>
__global__ void __launch_bounds__(NT) k_factor(int ni, const int *__restrict__ off, const double *__restrict__ D,
                                                const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                double *__restrict__ Dinv, int *__restrict__ info,
                                                const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    __shared__ GJShared<RPT, CPT> sh;
    const int tid = threadIdx.x;
    for (int q = tid; q < 2 * 16 * RPT; q += NT) (&sh.colbuf[0][0])[q] = 0.0;
    if (tid == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    double a[RPT][CPT];
    if (side == 0) {
        for (int j = 0; j < m; ++j) {
            const size_t l = (size_t)l0 + j;
            load_schur<RPT, CPT>(a, ni, D + l * nn, Lsub + l * ni, j > 0 ? Dinv + (l - 1) * nn : nullptr,
                                 j > 0 ? Usup + (l - 1) * ni : nullptr, nullptr, nullptr, nullptr);
            gj_invert<RPT, CPT>(a, ni, Dinv + l * nn, sh, j + 1);
        }
    } else {
        for (int j = nz - 1; j > m; --j) {
            const size_t l = (size_t)l0 + j;
            load_schur<RPT, CPT>(a, ni, D + l * nn, Usup + l * ni, j < nz - 1 ? Dinv + (l + 1) * nn : nullptr,
                                 j < nz - 1 ? Lsub + (l + 1) * ni : nullptr, nullptr, nullptr, nullptr);
            gj_invert<RPT, CPT>(a, ni, Dinv + l * nn, sh, j + 1);
        }
    }
    if (tid == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int RPT, int CPT>
__global__ void __launch_bounds__(NT) k_factor_mid(int ni, const int *__restrict__ off,
                                                    const double *__restrict__ D, const double *__restrict__ Lsub,
                                                    const double *__restrict__ Usup, double *__restrict__ Dinv,
                                                    int *__restrict__ info, const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x;
    if (st && st[c].done) return;
    __shared__ GJShared<RPT, CPT> sh;
    const int tid = threadIdx.x;
    for (int q = tid; q < 2 * 16 * RPT; q += NT) (&sh.colbuf[0][0])[q] = 0.0;
    if (tid == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    const size_t l = (size_t)l0 + m;
    double a[RPT][CPT];
    load_schur<RPT, CPT>(a, ni, D + l * nn, Lsub + l * ni, Dinv + (l - 1) * nn, Usup + (l - 1) * ni,
                         Usup + l * ni, m < nz - 1 ? Dinv + (l + 1) * nn : nullptr,
                         m < nz - 1 ? Lsub + (l + 1) * ni : nullptr);
    gj_invert<RPT, CPT>(a, ni, Dinv + l * nn, sh, m + 1);
    if (tid == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

__global__ void k_zero_int(int n, int *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// out_sh[i] = sum_m A[m][i] v[m] (A stored transposed), all threads; two
// threads per row for ni <= 128, four independent accumulators per thread.
// Ends with a barrier: out_sh is ready for every thread.
__device__ __forceinline__ void matvec_T(int ni, const double *__restrict__ A, const double *v, double *out_sh,
                                         double *part_sh) {
    const int tid = threadIdx.x;
    if (ni <= 128) {
        const int i = tid & 127, h = tid >> 7;
        const int half = (ni + 1) >> 1;
        const int m0 = h * half, m1 = min(ni, m0 + half);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (i < ni) {
            int m = m0;
            for (; m + 3 < m1; m += 4) {
                a0 = fma(A[(size_t)m * ni + i], v[m], a0);
                a1 = fma(A[(size_t)(m + 1) * ni + i], v[m + 1], a1);
                a2 = fma(A[(size_t)(m + 2) * ni + i], v[m + 2], a2);
                a3 = fma(A[(size_t)(m + 3) * ni + i], v[m + 3], a3);
            }
            for (; m < m1; m++) a0 = fma(A[(size_t)m * ni + i], v[m], a0);
        }
        const double acc = (a0 + a1) + (a2 + a3);
        if (h == 1 && i < ni) part_sh[i] = acc;
        __syncthreads();
        if (h == 0 && i < ni) out_sh[i] = acc + part_sh[i];
        __syncthreads();
    } else {
        for (int i = tid; i < ni; i += NT) {
            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
            int m = 0;
            for (; m + 3 < ni; m += 4) {
                a0 = fma(A[(size_t)m * ni + i], v[m], a0);
                a1 = fma(A[(size_t)(m + 1) * ni + i], v[m + 1], a1);
                a2 = fma(A[(size_t)(m + 2) * ni + i], v[m + 2], a2);
                a3 = fma(A[(size_t)(m + 3) * ni + i], v[m + 3], a3);
            }
            for (; m < ni; m++) a0 = fma(A[(size_t)m * ni + i], v[m], a0);
            out_sh[i] = (a0 + a1) + (a2 + a3);
        }
        __syncthreads();
    }
}

// forward halves of the twisted solve: z_j (j < m) and w_j (j > m) into tmp
__global__ void __launch_bounds__(NT) k_solve_fwd(int ni, const int *__restrict__ off,
                                                   const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                   const double *__restrict__ Usup, const double *in,
                                                   double *tmp, const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    __shared__ double v[VK_NI_MAX];
    __shared__ double z[VK_NI_MAX];
    __shared__ double part[VK_NI_MAX];
    const int j0 = side == 0 ? 0 : nz - 1, j1 = side == 0 ? m : m, dj = side == 0 ? 1 : -1;
    for (int j = j0; j != j1; j += dj) {
        const size_t l = (size_t)l0 + j;
        const bool first = (j == j0);
        const double *cpl = side == 0 ? Lsub + l * ni : Usup + l * ni;
        for (int i = tid; i < ni; i += NT) {
            double w = in[l * ni + i];
            if (!first) w = w - cpl[i] * z[i];
            v[i] = w;
        }
        __syncthreads();
        matvec_T(ni, Dinv + l * nn, v, z, part);
        for (int i = tid; i < ni; i += NT) tmp[l * ni + i] = z[i];
    }
    // park b_m in the unused slot m so the backward kernel never reads `in`
    // (which may alias `out`)
    if (side == 0) {
        const size_t lm = (size_t)l0 + m;
        for (int i = tid; i < ni; i += NT) tmp[lm * ni + i] = in[lm * ni + i];
    }
}

// middle solution and back substitution outwards: x into out
__global__ void __launch_bounds__(NT) k_solve_bwd(int ni, const int *__restrict__ off,
                                                   const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                   const double *__restrict__ Usup, const double *tmp, double *out,
                                                   const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    __shared__ double v[VK_NI_MAX];
    __shared__ double x[VK_NI_MAX];
    __shared__ double y[VK_NI_MAX];
    __shared__ double part[VK_NI_MAX];
    const size_t lm = (size_t)l0 + m;
    // x_m = M^-1 (b_m - Lsub_m z_{m-1} - Usup_m w_{m+1}); both sides compute it
    for (int i = tid; i < ni; i += NT) {
        double w = tmp[lm * ni + i] - Lsub[lm * ni + i] * tmp[(lm - 1) * ni + i];
        if (m < nz - 1) w = w - Usup[lm * ni + i] * tmp[(lm + 1) * ni + i];
        v[i] = w;
    }
    __syncthreads();
    matvec_T(ni, Dinv + lm * nn, v, x, part);
    if (side == 0)
        for (int i = tid; i < ni; i += NT) out[lm * ni + i] = x[i];
    const int j0 = side == 0 ? m - 1 : m + 1, j1 = side == 0 ? -1 : nz, dj = side == 0 ? -1 : 1;
    for (int j = j0; j != j1; j += dj) {
        const size_t l = (size_t)l0 + j;
        const double *cpl = side == 0 ? Usup + l * ni : Lsub + l * ni;
        for (int i = tid; i < ni; i += NT) v[i] = cpl[i] * x[i];
        __syncthreads();
        matvec_T(ni, Dinv + l * nn, v, y, part);
        for (int i = tid; i < ni; i += NT) {
            const double xi = tmp[l * ni + i] - y[i];
            out[l * ni + i] = xi;
            x[i] = xi;
        }
        __syncthreads();
    }
}

// r = b - A x with A = blocktridiag(diag(Lsub), D, diag(Usup)), accumulated
// in double-double (exact products via fma, error-free sums), one
// workgroup per layer, one thread per row
__device__ inline void dd_sub_prod(double &hi, double &lo, double a, double b) {
    const double p = a * b;
    const double pe = fma(a, b, -p);          // a*b = p + pe exactly
    const double s = hi - p;
    const double bb = s - hi;
    const double se = (hi - (s - bb)) + (-p - bb);   // hi - p = s + se exactly
    hi = s;
    lo = lo + (se - pe);
}

__global__ void __launch_bounds__(NT) k_residual(int ni, const int *__restrict__ off,
                                                  const int *__restrict__ layer_col, const double *__restrict__ D,
                                                  const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r, const vk_colstate *__restrict__ st) {
    const int l = blockIdx.x;
    const int c = layer_col[l];
    if (st && st[c].done) return;
    const int j = l - off[c], nz = off[c + 1] - off[c];
    const double *Dl = D + (size_t)l * ni * ni;
    const double *xl = x + (size_t)l * ni;
    for (int i = threadIdx.x; i < ni; i += NT) {
        double hi = b[(size_t)l * ni + i], lo = 0.0;
        const double *row = Dl + (size_t)i * ni;
        for (int m = 0; m < ni; m++) dd_sub_prod(hi, lo, row[m], xl[m]);
        if (j > 0) dd_sub_prod(hi, lo, Lsub[(size_t)l * ni + i], x[(size_t)(l - 1) * ni + i]);
        if (j < nz - 1) dd_sub_prod(hi, lo, Usup[(size_t)l * ni + i], x[(size_t)(l + 1) * ni + i]);
        r[(size_t)l * ni + i] = hi + lo;
    }
}

__global__ void k_axpy(int n, int ni, const int *__restrict__ layer_col, const double *__restrict__ dx,
                       double *__restrict__ x, const vk_colstate *__restrict__ st) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    if (st && st[layer_col[idx / ni]].done) return;
    x[idx] = x[idx] + dx[idx];
}

template <int N>
int launch_factor_n(int ni, int ncols, const int *off, const double *D, const double *Lsub, const double *Usup,
                    double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    hipLaunchKernelGGL((k_factor<N, N>), dim3(2 * ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
    hipLaunchKernelGGL((k_factor_mid<N, N>), dim3(ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
    return 0;
}

}  // namespace

int vk_launch_factor(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                     const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    const int ni = plan->ni;
    const int n = (ni + 15) / 16;
    const int nc = cols->ncols;
    if (info) hipLaunchKernelGGL(k_zero_int, dim3((nc + 255) / 256), dim3(256), 0, s, nc, info);
    switch (n) {
    case 1: launch_factor_n<1>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 2: launch_factor_n<2>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 3: launch_factor_n<3>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 4: launch_factor_n<4>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 5: launch_factor_n<5>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 6: launch_factor_n<6>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 7: launch_factor_n<7>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 8: launch_factor_n<8>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 9: launch_factor_n<9>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 10: launch_factor_n<10>(ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    default: return VK_ETOOBIG;
    }
    VK_CHECK(hipGetLastError());
    return 0;
}

// out = lhs^-1 in (tmp: L*ni scratch, may not alias in/out; in may alias out)
int vk_launch_solve(const vk_plan *plan, const vk_columns *cols, const double *Dinv, const double *Lsub,
                    const double *Usup, const double *in, double *out, double *tmp, const vk_colstate *st,
                    hipStream_t s) {
    hipLaunchKernelGGL(k_solve_fwd, dim3(2 * cols->ncols), dim3(NT), 0, s, plan->ni, cols->d_off, Dinv, Lsub, Usup, in,
                       tmp, st);
    hipLaunchKernelGGL(k_solve_bwd, dim3(2 * cols->ncols), dim3(NT), 0, s, plan->ni, cols->d_off, Dinv, Lsub, Usup,
                       tmp, out, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

int vk_launch_residual(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                       const double *Usup, const double *x, const double *b, double *r, const vk_colstate *st,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_residual, dim3(cols->L), dim3(NT), 0, s, plan->ni, cols->d_off, cols->d_layer_col, D, Lsub,
                       Usup, x, b, r, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

// x = A^-1 b with `refine` steps of iterative refinement (work, tmp: L*ni scratch each)
int vk_launch_solve_refined(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Dinv,
                            const double *Lsub, const double *Usup, const double *b, double *x, double *work,
                            double *tmp, int refine, const vk_colstate *st, hipStream_t s) {
    int r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, b, x, tmp, st, s);
    if (r) return r;
    const int n = cols->L * plan->ni;
    for (int q = 0; q < refine; q++) {
        r = vk_launch_residual(plan, cols, D, Lsub, Usup, x, b, work, st, s);
        if (r) return r;
        r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, work, work, tmp, st, s);
        if (r) return r;
        hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, n, plan->ni, cols->d_layer_col, work, x,
                           st);
        VK_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" int vk_block_factor(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                               const double *d_Lsub, const double *d_Usup, double *d_Dinv, int *d_info,
                               void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_Dinv || d_D == d_Dinv) return VK_EINVAL;
    return vk_launch_factor(plan, cols, d_D, d_Lsub, d_Usup, d_Dinv, d_info, nullptr, (hipStream_t)stream);
}

extern "C" int vk_block_solve(const vk_plan *plan, const vk_columns *cols, const double *d_Dinv,
                              const double *d_Lsub, const double *d_Usup, double *d_rhs, void *stream) {
    if (!plan || !cols || !d_Dinv || !d_Lsub || !d_Usup || !d_rhs) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    double *tmp = nullptr;
    VK_CHECK(hipMallocAsync((void **)&tmp, sizeof(double) * (size_t)cols->L * plan->ni, s));
    int r = vk_launch_solve(plan, cols, d_Dinv, d_Lsub, d_Usup, d_rhs, d_rhs, tmp, nullptr, s);
    VK_CHECK(hipFreeAsync(tmp, s));
    return r;
}

extern "C" int vk_block_residual(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                                 const double *d_Lsub, const double *d_Usup, const double *d_x, const double *d_b,
                                 double *d_r, void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_x || !d_b || !d_r) return VK_EINVAL;
    return vk_launch_residual(plan, cols, d_D, d_Lsub, d_Usup, d_x, d_b, d_r, nullptr, (hipStream_t)stream);
}
