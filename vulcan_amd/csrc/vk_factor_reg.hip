// Block-Thomas factorisation, register Gauss-Jordan variant (factor variant 0).
//
// The inversion of each Schur complement is an in-register Gauss-Jordan
// elimination with partial pivoting: a 16 x 16 grid of threads holds the
// ni x ni block (RPT x CPT elements each), one barrier per pivot step, the
// pivot column travels through LDS and the pivot row through intra-wave
// shuffles (the 16 threads sharing a column group are 16 consecutive lanes
// of one wave).  Register indices are static: the pivot column block is a
// template parameter and the (wave-uniform) pivot row is selected by a
// scalar switch.
#include "vk_blk.h"

namespace {

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
    load_schur<RPT, CPT>(a, ni, D + l * nn, Lsub + l * ni, m > 0 ? Dinv + (l - 1) * nn : nullptr,
                         m > 0 ? Usup + (l - 1) * ni : nullptr, Usup + l * ni,
                         m < nz - 1 ? Dinv + (l + 1) * nn : nullptr, m < nz - 1 ? Lsub + (l + 1) * ni : nullptr);
    gj_invert<RPT, CPT>(a, ni, Dinv + l * nn, sh, m + 1);
    if (tid == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int N>
void launch_reg(int ni, int ncols, const int *off, const double *D, const double *Lsub, const double *Usup,
                double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    hipLaunchKernelGGL((k_factor<N, N>), dim3(2 * ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
    hipLaunchKernelGGL((k_factor_mid<N, N>), dim3(ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
}

}  // namespace

int vk_launch_factor_register(int nb, int ni, int ncols, const int *off, const double *D, const double *Lsub,
                              const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    switch (nb) {
    case 1: launch_reg<1>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 2: launch_reg<2>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 3: launch_reg<3>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 4: launch_reg<4>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 5: launch_reg<5>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 6: launch_reg<6>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 7: launch_reg<7>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 8: launch_reg<8>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 9: launch_reg<9>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 10: launch_reg<10>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    default: return VK_ETOOBIG;
    }
    return 0;
}
