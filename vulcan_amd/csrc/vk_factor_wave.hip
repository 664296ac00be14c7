// Block-Thomas factorisation, two-wave Gauss-Jordan variant (factor
// variant 3, ni <= 112).
#include "vk_blk.h"

namespace {

// ---------------------------------------------------------------------------
// Two waves per block inversion (factor variant 3, ni <= 112)
//
// 128 lanes form an 8 x 16 grid, lane l = 8*lc + lr, and lane (lr, lc)
// holds rows lr + 8a (a < 2 NB) and columns lc + 16b (b < NB) in registers:
// unblocked Gauss-Jordan with partial pivoting.  Per pivot step the eight
// lanes owning column k (all in one wave) find the pivot (DPP max within the
// lane octet) and publish the old column through LDS; one barrier (two
// waves); the pivot row is exchanged through LDS inside each wave (lane
// (lr, lc) only needs row p from lane (p%8, lc), its own wave); every lane
// then applies 2 NB^2 FMAs.  Small workgroups, so several chains share a CU
// and keep the FP64 pipes of all its SIMDs busy in batched runs.
// ---------------------------------------------------------------------------
constexpr int NTW = 128;

template <int NB>
struct WaveShared {
    static constexpr int PR = 2 * NB, PC = NB;
    double colbuf[2][8][PR];       // old pivot column by step parity, [lr][a]
    unsigned long long key[2];     // pivot key by step parity
    double rowbuf[16][PC];         // pivot row, [lc][b]
    int ipos[16 * NB];
    int perm[16 * NB];
    int info;
};

// Max over the 8 lanes of each aligned lane octet (all lanes active).
__device__ __forceinline__ unsigned long long oct_max_u64(unsigned long long v) {
    unsigned long long u;
    u = dpp_u64<0xB1>(v);    // quad_perm [1,0,3,2]
    v = u > v ? u : v;
    u = dpp_u64<0x4E>(v);    // quad_perm [2,3,0,1]
    v = u > v ? u : v;
    u = dpp_u64<0x141>(v);   // row_half_mirror
    return u > v ? u : v;
}

template <int NB>
__device__ __forceinline__ void wave_invert(double (&a)[2 * NB][NB], int ni, double *__restrict__ X,
                                            WaveShared<NB> &sh, int tag) {
    constexpr int PR = 2 * NB, PC = NB;
    const int l = threadIdx.x, lr = l & 7, lc = l >> 3;
    unsigned used = 0;   // bit a: row lr + 8a already pivoted
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PC; b++) {
#pragma unroll 1
        for (int kg = 0; kg < 16; kg++) {
            const int k = 16 * b + kg;
            if (k >= ni) break;                       // uniform
            const int par = k & 1;
            const bool colgrp = (lc == kg);
            unsigned long long key = 0;
            if (colgrp) {
#pragma unroll
                for (int ra = 0; ra < PR; ra++) {
                    const int r = lr + 8 * ra;
                    if (r < ni && !((used >> ra) & 1u)) {
                        const unsigned long long kk =
                            ((unsigned long long)__double_as_longlong(fabs(a[ra][b])) & ~0xffull) | (unsigned)(255 - r);
                        key = kk > key ? kk : key;
                    }
                    sh.colbuf[par][lr][ra] = a[ra][b];
                }
            }
            key = oct_max_u64(key);
            if (colgrp && lr == 0) sh.key[par] = key;
            __syncthreads();
            key = sh.key[par];
            const int p = key ? 255 - (int)(key & 0xffull) : 0;
            if ((key >> 8) == 0ull && l == 0 && sh.info == 0) sh.info = tag;
            const int pr = p & 7, pa = p >> 3;
            const double inv = 1.0 / sh.colbuf[par][pr][pa];
            // the pivot row, from the lane (pr, lc) of this wave
            if (lr == pr) {
                switch (pa) {
#define VK_WROW(R)                                                                                           \
    case R:                                                                                                  \
        if (R < PR) {                                                                                        \
            _Pragma("unroll") for (int cb = 0; cb < PC; cb++) sh.rowbuf[lc][cb] = a[R < PR ? R : 0][cb];     \
        }                                                                                                    \
        break;
                    VK_WROW(0) VK_WROW(1) VK_WROW(2) VK_WROW(3) VK_WROW(4) VK_WROW(5) VK_WROW(6)
                    VK_WROW(7) VK_WROW(8) VK_WROW(9) VK_WROW(10) VK_WROW(11) VK_WROW(12) VK_WROW(13)
#undef VK_WROW
                default:
                    break;
                }
            }
            __builtin_amdgcn_wave_barrier();
            double rs[PC];
#pragma unroll
            for (int cb = 0; cb < PC; cb++) rs[cb] = sh.rowbuf[lc][cb] * inv;
#pragma unroll
            for (int ra = 0; ra < PR; ra++) {
                const double f = sh.colbuf[par][lr][ra];
#pragma unroll
                for (int cb = 0; cb < PC; cb++) a[ra][cb] = fma(-f, rs[cb], a[ra][cb]);
                if (colgrp) a[ra][b] = -f * inv;      // pivot column
            }
            if (lr == pr) {                           // pivot row
                switch (pa) {
#define VK_SROW(R)                                                                                           \
    case R:                                                                                                  \
        if (R < PR) {                                                                                        \
            _Pragma("unroll") for (int cb = 0; cb < PC; cb++) a[R < PR ? R : 0][cb] = rs[cb];                \
            if (colgrp) a[R < PR ? R : 0][b] = inv;                                                          \
            used |= 1u << R;                                                                                 \
        }                                                                                                    \
        break;
                    VK_SROW(0) VK_SROW(1) VK_SROW(2) VK_SROW(3) VK_SROW(4) VK_SROW(5) VK_SROW(6)
                    VK_SROW(7) VK_SROW(8) VK_SROW(9) VK_SROW(10) VK_SROW(11) VK_SROW(12) VK_SROW(13)
#undef VK_SROW
                default:
                    break;
                }
            }
            if (l == 0) {
                sh.perm[k] = p;
                sh.ipos[p] = k;
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    // inv[ipos[r]][perm[c]] = A[r][c]  ->  X[perm[c]][ipos[r]]
#pragma unroll
    for (int ra = 0; ra < PR; ra++) {
        const int r = lr + 8 * ra;
        if (r >= ni) continue;
        const int q = sh.ipos[r];
#pragma unroll
        for (int cb = 0; cb < PC; cb++) {
            const int c = lc + 16 * cb;
            if (c < ni) X[(size_t)sh.perm[c] * ni + q] = a[ra][cb];
        }
    }
}

// a = D_l - diag(cl) Xn^T diag(cn) [- diag(cl2) Xn2^T diag(cn2)] in the two-wave layout
template <int NB>
__device__ __forceinline__ void wave_load_schur(double (&a)[2 * NB][NB], int ni, const double *__restrict__ Dl,
                                                const double *cl, const double *Xn, const double *cn,
                                                const double *cl2, const double *Xn2, const double *cn2) {
    constexpr int PR = 2 * NB, PC = NB;
    const int l = threadIdx.x, lr = l & 7, lc = l >> 3;
#pragma unroll
    for (int ra = 0; ra < PR; ra++) {
        const int i = lr + 8 * ra;
        const bool rowok = i < ni;
        const double li = (Xn && rowok) ? cl[i] : 0.0;
        const double li2 = (Xn2 && rowok) ? cl2[i] : 0.0;
#pragma unroll
        for (int cb = 0; cb < PC; cb++) {
            const int cc = lc + 16 * cb;
            double v = 0.0;
            if (rowok && cc < ni) {
                v = Dl[(size_t)i * ni + cc];
                if (Xn) v = v - li * (Xn[(size_t)cc * ni + i] * cn[cc]);
                if (Xn2) v = v - li2 * (Xn2[(size_t)cc * ni + i] * cn2[cc]);
            }
            a[ra][cb] = v;
        }
    }
}

template <int NB>
__global__ void __launch_bounds__(NTW) k_factor_wave(int ni, const int *__restrict__ off, const double *__restrict__ D,
                                                     const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                     double *__restrict__ Dinv, int *__restrict__ info,
                                                     const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    __shared__ WaveShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    double a[2 * NB][NB];
    if (side == 0) {
        for (int j = 0; j < m; ++j) {
            const size_t l = (size_t)l0 + j;
            wave_load_schur<NB>(a, ni, D + l * nn, Lsub + l * ni, j > 0 ? Dinv + (l - 1) * nn : nullptr,
                                j > 0 ? Usup + (l - 1) * ni : nullptr, nullptr, nullptr, nullptr);
            wave_invert<NB>(a, ni, Dinv + l * nn, sh, j + 1);
            __threadfence();   // the next layer reads this inverse back
        }
    } else {
        for (int j = nz - 1; j > m; --j) {
            const size_t l = (size_t)l0 + j;
            wave_load_schur<NB>(a, ni, D + l * nn, Usup + l * ni, j < nz - 1 ? Dinv + (l + 1) * nn : nullptr,
                                j < nz - 1 ? Lsub + (l + 1) * ni : nullptr, nullptr, nullptr, nullptr);
            wave_invert<NB>(a, ni, Dinv + l * nn, sh, j + 1);
            __threadfence();
        }
    }
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int NB>
__global__ void __launch_bounds__(NTW) k_factor_wave_mid(int ni, const int *__restrict__ off,
                                                         const double *__restrict__ D, const double *__restrict__ Lsub,
                                                         const double *__restrict__ Usup, double *__restrict__ Dinv,
                                                         int *__restrict__ info, const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x;
    if (st && st[c].done) return;
    __shared__ WaveShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    const size_t l = (size_t)l0 + m;
    double a[2 * NB][NB];
    wave_load_schur<NB>(a, ni, D + l * nn, Lsub + l * ni, m > 0 ? Dinv + (l - 1) * nn : nullptr,
                        m > 0 ? Usup + (l - 1) * ni : nullptr, Usup + l * ni,
                        m < nz - 1 ? Dinv + (l + 1) * nn : nullptr, m < nz - 1 ? Lsub + (l + 1) * ni : nullptr);
    wave_invert<NB>(a, ni, Dinv + l * nn, sh, m + 1);
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int N>
void launch_wave(int ni, int ncols, const int *off, const double *D, const double *Lsub, const double *Usup,
                 double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    hipLaunchKernelGGL((k_factor_wave<N>), dim3(2 * ncols), dim3(NTW), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
    hipLaunchKernelGGL((k_factor_wave_mid<N>), dim3(ncols), dim3(NTW), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
}

}  // namespace

int vk_launch_factor_wave(int nb, int ni, int ncols, const int *off, const double *D, const double *Lsub,
                          const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    switch (nb) {
    case 1: launch_wave<1>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 2: launch_wave<2>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 3: launch_wave<3>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 4: launch_wave<4>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 5: launch_wave<5>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 6: launch_wave<6>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 7: launch_wave<7>(ni, ncols, off, D, Lsub, Usup, Dinv, info, st, s); break;
    default: return VK_EINVAL;
    }
    return 0;
}
