// Block-Thomas factorisation, blocked Gauss-Jordan variants (factor
// variants 1 and 2: one or two resident workgroups per CU).  The inversion
// itself is blk_invert (panels of 16 pivots on one wave, rank-16 Schur
// updates on FP64 MFMA).  Included by vk_factor_blk1.hip and
// vk_factor_blk2.hip, one translation unit per occupancy.
#pragma once
#include "vk_blk.h"

namespace {

// 8-wave workgroups (blocks above 112 rows) run one per CU
// (order: the chain each workgroup runs, 2c + side; null = blockIdx.x)
template <int NB, int MINW>
__global__ void __launch_bounds__(blk_nt(NB), blk_nt(NB) > NT ? 1 : MINW) k_factor_blk(int ni, const int *__restrict__ off,
                                                    const int *__restrict__ order, const double *__restrict__ D,
                                                    const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                    double *__restrict__ Dinv, int *__restrict__ info,
                                                    const vk_colstate *__restrict__ st) {
    const int chain = order ? order[blockIdx.x] : (int)blockIdx.x;
    const int c = chain >> 1, side = chain & 1;
    if (st && st[c].done) return;
    __shared__ BlkShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    f64x4 acc[BlkGrid<NB>::TBR][BlkGrid<NB>::TBC];
    // one loop for both chains (a single copy of the inversion code, which
    // is large: the kernel's instruction footprint matters with two
    // workgroups per CU): side 0 runs j = 0 .. m-1 coupled through Lsub_j /
    // Usup_{j-1}, side 1 runs j = nz-1 .. m+1 through Usup_j / Lsub_{j+1}
    const int dir = side == 0 ? 1 : -1, j0 = side == 0 ? 0 : nz - 1, cnt = side == 0 ? m : nz - 1 - m;
    const double *cown = side == 0 ? Lsub : Usup, *cprev = side == 0 ? Usup : Lsub;
    for (int t = 0; t < cnt; ++t) {
        const int j = j0 + dir * t;
        const size_t l = (size_t)l0 + j, lp = (size_t)((long)l - dir);
        blk_load_schur<NB>(acc, ni, D + l * nn, cown + l * ni, t > 0 ? Dinv + lp * nn : nullptr,
                           t > 0 ? cprev + lp * ni : nullptr, nullptr, nullptr, nullptr);
        blk_invert<NB>(acc, ni, Dinv + l * nn, sh, j + 1);
    }
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int NB, int MINW>
__global__ void __launch_bounds__(blk_nt(NB), blk_nt(NB) > NT ? 1 : MINW) k_factor_blk_mid(int ni, const int *__restrict__ off,
                                                        const double *__restrict__ D, const double *__restrict__ Lsub,
                                                        const double *__restrict__ Usup, double *__restrict__ Dinv,
                                                        int *__restrict__ info, const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x;
    if (st && st[c].done) return;
    __shared__ BlkShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    const size_t l = (size_t)l0 + m;
    f64x4 acc[BlkGrid<NB>::TBR][BlkGrid<NB>::TBC];
    blk_load_schur<NB>(acc, ni, D + l * nn, Lsub + l * ni, m > 0 ? Dinv + (l - 1) * nn : nullptr,
                       m > 0 ? Usup + (l - 1) * ni : nullptr, Usup + l * ni,
                       m < nz - 1 ? Dinv + (l + 1) * nn : nullptr, m < nz - 1 ? Lsub + (l + 1) * ni : nullptr);
    blk_invert<NB>(acc, ni, Dinv + l * nn, sh, m + 1);
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int N, int MINW>
void launch_blk(int ni, int ncols, const int *off, const int *order, const double *D, const double *Lsub,
                const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    hipLaunchKernelGGL((k_factor_blk<N, MINW>), dim3(2 * ncols), dim3(blk_nt(N)), 0, s, ni, off, order, D, Lsub, Usup,
                       Dinv, info, st);
    hipLaunchKernelGGL((k_factor_blk_mid<N, MINW>), dim3(ncols), dim3(blk_nt(N)), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                       st);
}

template <int MINW>
int launch_blocked(int nb, int ni, int ncols, const int *off, const int *order, const double *D, const double *Lsub,
                   const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    switch (nb) {
    case 1: launch_blk<1, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 2: launch_blk<2, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 3: launch_blk<3, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 4: launch_blk<4, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 5: launch_blk<5, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 6: launch_blk<6, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 7: launch_blk<7, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 8: launch_blk<8, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 9: launch_blk<9, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    case 10: launch_blk<10, MINW>(ni, ncols, off, order, D, Lsub, Usup, Dinv, info, st, s); break;
    default: return VK_ETOOBIG;
    }
    return 0;
}

}  // namespace
