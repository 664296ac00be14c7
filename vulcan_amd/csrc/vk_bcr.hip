// Block cyclic reduction factorisation and solves (factor variant 4).
#include <algorithm>
#include <atomic>

#include "vk_blk.h"

namespace {

// ---------------------------------------------------------------------------
// Block cyclic reduction (factor variant 4)
//
// Every level eliminates every other active layer of each column at once
// (the even positions of an odd count, the odd ones of an even count, so
// that floor(n/2) layers survive): their blocks are inverted in parallel
// (blk_invert), and each surviving layer j with eliminated neighbours l
// (left) and r (right) takes
//     P_j = B_j A_l^-1,   Q_j = C_j A_r^-1,
//     A_j -= P_j C_l + Q_j B_r,   B_j = -P_j B_l,   C_j = -Q_j C_r
// (B: coupling to the previous active layer, C: to the next; diagonal at
// level 0, dense above).  The chain over layers becomes floor(log2 nz)
// levels of independent work and one inversion on top: the Schur-complement updates above level 0 are dense
// n x n products on FP64 MFMA, level 0 is elementwise.  P and Q are kept
// for the right-hand side reduction of the solves.
//
// Storage: A row-major (the survivors' Schur complements, formed from D by
// the level-0 update; level 0 inverts from D itself); P, Q, B, C
// transposed (Pt[b][a] = P[a][b]), so that the solve sweeps, one thread per
// row a, read them coalesced; inverses transposed as in the chain variants.
// ---------------------------------------------------------------------------

// invert A[layer] for the listed layers (the eliminated ones of a level, or
// the top layer of each column) into Dinv[layer] (transposed storage)
template <int NB>
__global__ void __launch_bounds__(blk_nt(NB), blk_nt(NB) > NT ? 1 : 2) k_bcr_invert(int ni, int n, const int4 *__restrict__ el,
                                                       const int2 *__restrict__ top, const int *__restrict__ off,
                                                       const double *__restrict__ A, double *__restrict__ Dinv,
                                                       int *__restrict__ info, const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    const int layer = el ? el[q].x : top[q].x, c = el ? el[q].w : top[q].y;
    if (st && st[c].done) return;
    __shared__ BlkShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const size_t nn = (size_t)ni * ni;
    f64x4 acc[BlkGrid<NB>::TBR][BlkGrid<NB>::TBC];
    blk_load_schur<NB>(acc, ni, A + layer * nn, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    blk_invert<NB>(acc, ni, Dinv + layer * nn, sh, layer - off[c] + 1);
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

// level 0: couplings are the diagonals Lsub (B) and Usup (C); elementwise.
// Survivors: 32 x 32 tiles (a: block row, b: block column), the grid's y
// dimension over a block's tiles; P, Q, B, C (transposed storage, index
// tt = b*ni + a) are written in the loop order with a fastest, the Schur
// update of A (row-major) goes through LDS and is written with b fastest:
// A = D - update (A holds nothing before level 0: the level's eliminated
// blocks are inverted from D directly).
// Workgroups n .. n+ne-1 form the back-substitution operators of the
// eliminated layers instead: XB = A_k^-1 diag(Lsub_k), XC = A_k^-1 diag(Usup_k)
// (over tt, the block's range split over the grid's y dimension).
__global__ void k_bcr_update0(int ni, int n, const int4 *__restrict__ sv, const int *__restrict__ sflag,
                              int slot0, const double *__restrict__ Lsub, const double *__restrict__ Usup,
                              const double *__restrict__ Dinv, const double *__restrict__ D, double *__restrict__ A,
                              double *__restrict__ Bt, double *__restrict__ Ct, double *__restrict__ Pt,
                              double *__restrict__ Qt, const int4 *__restrict__ el, int ne,
                              double *__restrict__ XBt, double *__restrict__ XCt, const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n + ne) return;
    if (q >= n) {
        const int4 e = el[q - n];
        if (st && st[e.w].done) return;
        const size_t nn = (size_t)ni * ni, k = e.x;
        const double *X = Dinv + k * nn;     // X[b][a] = inv[a][b]
        const int span = (ni * ni + gridDim.y - 1) / gridDim.y, t0 = blockIdx.y * span;
        const int t1 = min(ni * ni, t0 + span);
        for (int tt = t0 + threadIdx.x; tt < t1; tt += blockDim.x) {
            const int b = tt / ni;
            XBt[k * nn + tt] = X[tt] * Lsub[k * ni + b];
            XCt[k * nn + tt] = X[tt] * Usup[k * ni + b];
        }
        return;
    }
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const int fl = sflag[q];
    const size_t nn = (size_t)ni * ni, j = e.x;
    const size_t slot = (size_t)(slot0 + q) * nn;
    const double *XL = e.y >= 0 ? Dinv + (size_t)e.y * nn : nullptr;   // X[m][i] = inv[i][m]
    const double *XR = e.z >= 0 ? Dinv + (size_t)e.z * nn : nullptr;
    __shared__ double up[32][33];
    const int tpr = (ni + 31) / 32;
    for (int tile = blockIdx.y; tile < tpr * tpr; tile += gridDim.y) {
        const int a0 = 32 * (tile % tpr), b0 = 32 * (tile / tpr);
        for (int u = threadIdx.x; u < 32 * 32; u += blockDim.x) {
            const int al = u & 31, bl = u >> 5, a = a0 + al, b = b0 + bl;
            double upd = 0.0;
            if (a < ni && b < ni) {
                const size_t tt = (size_t)b * ni + a;
                if (XL) {
                    const double p = Lsub[j * ni + a] * XL[tt];
                    Pt[slot + tt] = p;
                    upd = p * Usup[(size_t)e.y * ni + b];
                    Bt[j * nn + tt] = (fl & 1) ? -(p * Lsub[(size_t)e.y * ni + b]) : 0.0;
                } else {
                    Bt[j * nn + tt] = 0.0;
                }
                if (XR) {
                    const double qv = Usup[j * ni + a] * XR[tt];
                    Qt[slot + tt] = qv;
                    upd = upd + qv * Lsub[(size_t)e.z * ni + b];
                    Ct[j * nn + tt] = (fl & 2) ? -(qv * Usup[(size_t)e.z * ni + b]) : 0.0;
                } else {
                    Ct[j * nn + tt] = 0.0;
                }
            }
            up[al][bl] = upd;
        }
        __syncthreads();
        for (int u = threadIdx.x; u < 32 * 32; u += blockDim.x) {
            const int bl = u & 31, al = u >> 5, a = a0 + al, b = b0 + bl;
            if (a < ni && b < ni) {
                const size_t o = j * nn + (size_t)a * ni + b;
                A[o] = D[o] - up[al][bl];
            }
        }
        __syncthreads();
    }
}

// Levels >= 1: dense couplings, products on FP64 MFMA, one workgroup per
// (layer, 32-row slab, 32 output columns): a slab's output rows depend
// only on the same rows of the left operands, so the slabs are
// independent, and the column groups split a slab's MFMA chain and right
// operands over several workgroups.  Two launches
// after the level's inversions, each one product deep:
//
// k_bcr_pq, survivors j (x < n):       P_j = B_j A_l^-1,  Q_j = C_j A_r^-1
//           eliminated k (x >= n):     the back-substitution operators
//                                      XB_k = A_k^-1 B_k,  XC_k = A_k^-1 C_k
//           (an eliminated layer's couplings are final: only survivors'
//           couplings change at this level)
// k_bcr_update, survivors j, slice 0:    A_j -= P_j C_l + Q_j B_r
//                            slice 1:    B_j = -P_j B_l,  C_j = -Q_j C_r
//           (slice 0 reads the eliminated neighbours' couplings, slice 1
//           writes the survivor's own: the two slices are independent)
// Grid z: column group (k_bcr_pq), slice * BCR_NCS<NB> + column group
// (k_bcr_update).
template <int NB>
constexpr int BCR_NBH = NB >= 2 ? 2 : 1;                        // column blocks per workgroup
template <int NB>
constexpr int BCR_NCS = (NB + BCR_NBH<NB> - 1) / BCR_NBH<NB>;   // workgroups across a block's columns

template <int NB, int KC>
__global__ void __launch_bounds__(NT) k_bcr_pq(int ni, int n, const int4 *__restrict__ sv, int slot0,
                                               const double *__restrict__ Dinv, const double *__restrict__ Bt,
                                               const double *__restrict__ Ct, double *__restrict__ Pt,
                                               double *__restrict__ Qt, const int4 *__restrict__ el, int ne,
                                               double *__restrict__ XBt, double *__restrict__ XCt,
                                               const vk_colstate *__restrict__ st) {
    constexpr int NBH = BCR_NBH<NB>;
    const int q = blockIdx.x, r0 = 32 * blockIdx.y, c0 = 16 * NBH * (int)blockIdx.z;
    if (q >= n + ne || r0 >= ni || c0 >= ni) return;
    const int4 e = q >= n ? el[q - n] : sv[q];
    if (st && st[e.w].done) return;
    __shared__ SlabShared<NBH, KC> g;
    const size_t nn = (size_t)ni * ni;
    f64x4 acc[(NBH + 1) / 2], acc2[(NBH + 1) / 2];
    slab_zero<NBH>(acc);
    slab_zero<NBH>(acc2);
    if (q >= n) {
        const size_t k = e.x;
        // A_k^-1 (a, c) = X[c][a]: the transposed-storage inverse is the At operand
        slab_gemm2<NBH, false, KC>(acc, acc2, ni, r0, Dinv + k * nn, Bt + k * nn, Dinv + k * nn, Ct + k * nn, g, c0);
        slab_store<NBH>(acc, ni, r0, XBt + k * nn, true, 1.0, nullptr, c0);
        slab_store<NBH>(acc2, ni, r0, XCt + k * nn, true, 1.0, nullptr, c0);
        return;
    }
    const size_t j = e.x, slot = (size_t)(slot0 + q) * nn;
    const bool hl = e.y >= 0, hr = e.z >= 0;
    // B_j(a, c) = Bt_j[c][a], inv(c, b) = X[b][c]
    if (hl) slab_gemm2<NBH, false, KC>(acc, acc2, ni, r0, Bt + j * nn, Dinv + (size_t)e.y * nn,
                                      hr ? Ct + j * nn : nullptr, hr ? Dinv + (size_t)e.z * nn : nullptr, g, c0);
    else if (hr) slab_gemm2<NBH, false, KC>(acc2, acc, ni, r0, Ct + j * nn, Dinv + (size_t)e.z * nn, nullptr, nullptr,
                                           g, c0);
    if (hl) slab_store<NBH>(acc, ni, r0, Pt + slot, true, 1.0, nullptr, c0);
    if (hr) slab_store<NBH>(acc2, ni, r0, Qt + slot, true, 1.0, nullptr, c0);
}

template <int NB, int KC>
__global__ void __launch_bounds__(NT) k_bcr_update(int ni, int n, const int4 *__restrict__ sv,
                                                   const int *__restrict__ sflag, int slot0, double *__restrict__ A,
                                                   double *__restrict__ Bt, double *__restrict__ Ct,
                                                   const double *__restrict__ Pt, const double *__restrict__ Qt,
                                                   const vk_colstate *__restrict__ st) {
    constexpr int NBH = BCR_NBH<NB>;
    const int q = blockIdx.x, r0 = 32 * blockIdx.y, cg = (int)blockIdx.z % BCR_NCS<NB>;
    const int slice = (int)blockIdx.z / BCR_NCS<NB>, c0 = 16 * NBH * cg;
    if (q >= n || r0 >= ni || c0 >= ni) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    __shared__ SlabShared<NBH, KC> g;
    const size_t nn = (size_t)ni * ni, j = e.x;
    const double *Ptj = Pt + (size_t)(slot0 + q) * nn, *Qtj = Qt + (size_t)(slot0 + q) * nn;
    f64x4 acc[(NBH + 1) / 2], acc2[(NBH + 1) / 2];
    slab_zero<NBH>(acc);
    slab_zero<NBH>(acc2);
    if (slice == 0) {
        // one accumulator for both products
        const bool hl = e.y >= 0, hr = e.z >= 0;
        if (hl) slab_gemm2<NBH, true, KC>(acc, acc2, ni, r0, Ptj, Ct + (size_t)e.y * nn, hr ? Qtj : nullptr,
                                         hr ? Bt + (size_t)e.z * nn : nullptr, g, c0);
        else if (hr) slab_gemm2<NBH, true, KC>(acc, acc2, ni, r0, Qtj, Bt + (size_t)e.z * nn, nullptr, nullptr, g, c0);
        slab_store<NBH>(acc, ni, r0, A + j * nn, false, -1.0, A + j * nn, c0);
        return;
    }
    // the new couplings reach past the eliminated neighbours (zero where there is no layer beyond)
    const int fl = sflag[q];
    const bool nl = fl & 1, nr = fl & 2;
    if (nl) slab_gemm2<NBH, false, KC>(acc, acc2, ni, r0, Ptj, Bt + (size_t)e.y * nn, nr ? Qtj : nullptr,
                                      nr ? Ct + (size_t)e.z * nn : nullptr, g, c0);
    else if (nr) slab_gemm2<NBH, false, KC>(acc2, acc, ni, r0, Qtj, Ct + (size_t)e.z * nn, nullptr, nullptr, g, c0);
    slab_store<NBH>(acc, ni, r0, Bt + j * nn, true, -1.0, nullptr, c0);
    slab_store<NBH>(acc2, ni, r0, Ct + j * nn, true, -1.0, nullptr, c0);
}

__global__ void k_copy_d(long n, const double *__restrict__ a, double *__restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Matrix-vector products of the solves: out[a] = sum_b M[a][b] v[b] with M
// stored transposed (Mt[b][a]), 256 threads: row a = tid % 128 (+128 ...),
// half h = tid / 128 of the b range, four independent accumulators per
// thread so that the loads of a row stay in flight; the two halves are
// combined by the caller through LDS.
constexpr int BCR_ST = 256;

__device__ __forceinline__ double tmv_part(const double *__restrict__ Mt, const double *v, int ni, int a, int h) {
    const int hb = (ni + 1) >> 1, b0 = h * hb, b1 = min(ni, b0 + hb);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = b0;
    for (; b + 3 < b1; b += 4) {
        s0 = fma(Mt[(size_t)b * ni + a], v[b], s0);
        s1 = fma(Mt[(size_t)(b + 1) * ni + a], v[b + 1], s1);
        s2 = fma(Mt[(size_t)(b + 2) * ni + a], v[b + 2], s2);
        s3 = fma(Mt[(size_t)(b + 3) * ni + a], v[b + 3], s3);
    }
    for (; b < b1; b++) s0 = fma(Mt[(size_t)b * ni + a], v[b], s0);
    return (s0 + s1) + (s2 + s3);
}

// solves: forward reduction r_j -= P_j r_l + Q_j r_r at level s
__global__ void __launch_bounds__(BCR_ST) k_bcr_fwd(int ni, int n, const int4 *__restrict__ sv, int slot0,
                                                    const double *__restrict__ P, const double *__restrict__ Q,
                                                    double *__restrict__ r, const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const size_t nn = (size_t)ni * ni;
    const double *Pj = P + (size_t)(slot0 + q) * nn, *Qj = Q + (size_t)(slot0 + q) * nn;
    __shared__ double rl[VK_NI_MAX], rr[VK_NI_MAX], pp[2][VK_NI_MAX];
    for (int i = threadIdx.x; i < ni; i += BCR_ST) {
        rl[i] = e.y >= 0 ? r[(size_t)e.y * ni + i] : 0.0;
        rr[i] = e.z >= 0 ? r[(size_t)e.z * ni + i] : 0.0;
    }
    __syncthreads();
    const int h = threadIdx.x >> 7;
    for (int a = threadIdx.x & 127; a < ni; a += 128) {
        const double sp = e.y >= 0 ? tmv_part(Pj, rl, ni, a, h) : 0.0;
        const double sq = e.z >= 0 ? tmv_part(Qj, rr, ni, a, h) : 0.0;
        pp[h][a] = sp + sq;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < ni; a += BCR_ST) r[(size_t)e.x * ni + a] = r[(size_t)e.x * ni + a] - (pp[0][a] + pp[1][a]);
}

// x_k = A_k^-1 r_k - XB_k x_l - XC_k x_r for the eliminated layers of a level
// (XB = A_k^-1 B_k, XC = A_k^-1 C_k, formed at factorisation); top: x = A^-1 r
__global__ void __launch_bounds__(BCR_ST) k_bcr_bwd(int ni, int n, const int4 *__restrict__ el,
                                                    const int2 *__restrict__ top, const double *__restrict__ Dinv,
                                                    const double *__restrict__ XB, const double *__restrict__ XC,
                                                    const double *__restrict__ r, double *__restrict__ x,
                                                    const vk_colstate *__restrict__ st, double *acc) {
    const int q = blockIdx.x;
    if (q >= n) return;
    int k, left = -1, right = -1, c;
    if (el) { const int4 e = el[q]; k = e.x; left = e.y; right = e.z; c = e.w; }
    else { k = top[q].x; c = top[q].y; }
    if (st && st[c].done) return;
    const size_t nn = (size_t)ni * ni;
    __shared__ double rk[VK_NI_MAX], xl[VK_NI_MAX], xr[VK_NI_MAX], pp[2][VK_NI_MAX];
    for (int i = threadIdx.x; i < ni; i += BCR_ST) {
        rk[i] = r[(size_t)k * ni + i];
        xl[i] = left >= 0 ? x[(size_t)left * ni + i] : 0.0;
        xr[i] = right >= 0 ? x[(size_t)right * ni + i] : 0.0;
    }
    __syncthreads();
    const int h = threadIdx.x >> 7;
    for (int a = threadIdx.x & 127; a < ni; a += 128) {
        double v = tmv_part(Dinv + (size_t)k * nn, rk, ni, a, h);
        if (left >= 0) v = v - tmv_part(XB + (size_t)k * nn, xl, ni, a, h);
        if (right >= 0) v = v - tmv_part(XC + (size_t)k * nn, xr, ni, a, h);
        pp[h][a] = v;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < ni; a += BCR_ST) {
        const double t = pp[0][a] + pp[1][a];
        x[(size_t)k * ni + a] = t;
        if (acc) acc[(size_t)k * ni + a] = acc[(size_t)k * ni + a] + t;
    }
}

// Solve sweeps for ni <= 128: each block's products are split over row
// groups of 32 (one workgroup each, grid y), so a block's operators are
// read by several CUs at once; 256 threads, thread t owns row a0 + t % 32
// and part h = t / 32 of the b range (chunks of ceil(ni/8) <= 16).  Every
// operator entry a thread needs is loaded into registers before the first
// product, and the eight partial sums of a row are combined through LDS in
// part order.
constexpr int MV_NT = 256, MV_RG = 32, MV_SP = 8, MV_CH = 16, MV_NIMAX = 128;

__device__ __forceinline__ void mv_load(double (&m)[MV_CH], const double *__restrict__ Mt, int ni, int a, int h,
                                        int ch) {
#pragma unroll
    for (int i = 0; i < MV_CH; i++) {
        const int b = h * ch + i;
        m[i] = (i < ch && b < ni && a < ni) ? Mt[(size_t)b * ni + a] : 0.0;
    }
}

__device__ __forceinline__ double mv_dot(const double (&m)[MV_CH], const double *v, int ni, int h, int ch) {
    double s0 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < MV_CH; i += 2) {
        const int b = h * ch + i;
        if (i < ch && b < ni) s0 = fma(m[i], v[b], s0);
        if (i + 1 < ch && b + 1 < ni) s1 = fma(m[i + 1], v[b + 1], s1);
    }
    return s0 + s1;
}

// rin: the right-hand sides read (the solve's input at level 0, else r
// itself); at level 0 (sflag non-null) the sweep also places the
// eliminated layers' right-hand sides into r, each layer once: by its
// right survivor, or by its left one where it ends the column (no layer
// beyond it: sflag bit 1 clear).
__global__ void __launch_bounds__(MV_NT) k_bcr_fwd_s(int ni, int n, const int4 *__restrict__ sv, int slot0,
                                                     const double *__restrict__ P, const double *__restrict__ Q,
                                                     const double *rin, double *r, const int *__restrict__ sflag,
                                                     const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x, a0 = MV_RG * blockIdx.y;
    if (q >= n || a0 >= ni) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const size_t nn = (size_t)ni * ni;
    const int lr = threadIdx.x % MV_RG, a = a0 + lr, h = threadIdx.x / MV_RG, ch = (ni + MV_SP - 1) / MV_SP;
    double mp[MV_CH], mq[MV_CH];
    mv_load(mp, P + (size_t)(slot0 + q) * nn, e.y >= 0 ? ni : 0, a, h, ch);
    mv_load(mq, Q + (size_t)(slot0 + q) * nn, e.z >= 0 ? ni : 0, a, h, ch);
    __shared__ double rl[MV_NIMAX], rr[MV_NIMAX], pp[MV_SP][MV_RG];
    for (int i = threadIdx.x; i < ni; i += MV_NT) {
        rl[i] = e.y >= 0 ? rin[(size_t)e.y * ni + i] : 0.0;
        rr[i] = e.z >= 0 ? rin[(size_t)e.z * ni + i] : 0.0;
    }
    __syncthreads();
    pp[h][lr] = mv_dot(mp, rl, ni, h, ch) + mv_dot(mq, rr, ni, h, ch);
    __syncthreads();
    if (threadIdx.x < MV_RG && a < ni) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < MV_SP; k++) t += pp[k][threadIdx.x];
        r[(size_t)e.x * ni + a] = rin[(size_t)e.x * ni + a] - t;
        if (sflag) {
            if (e.y >= 0) r[(size_t)e.y * ni + a] = rl[a];
            if (e.z >= 0 && !(sflag[q] & 2)) r[(size_t)e.z * ni + a] = rr[a];
        }
    }
}

__global__ void __launch_bounds__(MV_NT) k_bcr_bwd_s(int ni, int n, const int4 *__restrict__ el,
                                                     const int2 *__restrict__ top, const double *__restrict__ Dinv,
                                                     const double *__restrict__ XB, const double *__restrict__ XC,
                                                     const double *__restrict__ r, double *__restrict__ x,
                                                     const vk_colstate *__restrict__ st, double *acc) {
    const int q = blockIdx.x, a0 = MV_RG * blockIdx.y;
    if (q >= n || a0 >= ni) return;
    int k, left = -1, right = -1, c;
    if (el) { const int4 e = el[q]; k = e.x; left = e.y; right = e.z; c = e.w; }
    else { k = top[q].x; c = top[q].y; }
    if (st && st[c].done) return;
    const size_t nn = (size_t)ni * ni;
    const int lr = threadIdx.x % MV_RG, a = a0 + lr, h = threadIdx.x / MV_RG, ch = (ni + MV_SP - 1) / MV_SP;
    // x_k = A_k^-1 r_k - XB x_l - XC x_r: one pass over three operators
    double mx[MV_CH], mb[MV_CH], mc[MV_CH];
    mv_load(mx, Dinv + (size_t)k * nn, ni, a, h, ch);   // X[m][a] = inv[a][m]: the transposed inverse
    mv_load(mb, XB + (size_t)k * nn, left >= 0 ? ni : 0, a, h, ch);
    mv_load(mc, XC + (size_t)k * nn, right >= 0 ? ni : 0, a, h, ch);
    __shared__ double rk[MV_NIMAX], xl[MV_NIMAX], xr[MV_NIMAX], pp[MV_SP][MV_RG];
    for (int i = threadIdx.x; i < ni; i += MV_NT) {
        rk[i] = r[(size_t)k * ni + i];
        xl[i] = left >= 0 ? x[(size_t)left * ni + i] : 0.0;
        xr[i] = right >= 0 ? x[(size_t)right * ni + i] : 0.0;
    }
    __syncthreads();
    pp[h][lr] = (mv_dot(mx, rk, ni, h, ch) - mv_dot(mb, xl, ni, h, ch)) - mv_dot(mc, xr, ni, h, ch);
    __syncthreads();
    if (threadIdx.x < MV_RG && a < ni) {
        double t = 0.0;
#pragma unroll
        for (int p2 = 0; p2 < MV_SP; p2++) t += pp[p2][threadIdx.x];
        x[(size_t)k * ni + a] = t;
        if (acc) acc[(size_t)k * ni + a] = acc[(size_t)k * ni + a] + t;
    }
}

// The last level of one column when it holds two layers, the top survivor
// s and the eliminated layer e to its right: r_s -= Q_s r_e, x_s = A_s^-1
// r_s, x_e = A_e^-1 r_e - XB_e x_s, in one workgroup of MV_NT-thread row
// groups (ni <= 96) instead of three launches.  Each product is formed as
// in k_bcr_fwd_s / k_bcr_bwd_s (same partition of the columns, same order
// of the sums; their product with an absent neighbour is exactly +0.0 and
// stands here as that constant), so the results are theirs.
constexpr int TAIL_G = 3, TAIL_NIMAX = TAIL_G * MV_RG;
__global__ void __launch_bounds__(MV_NT *TAIL_G) k_bcr_tail_s(int ni, const int4 *__restrict__ sv, int slot0,
                                                              const int4 *__restrict__ el, const double *__restrict__ Q,
                                                              const double *__restrict__ Dinv,
                                                              const double *__restrict__ XB, double *r, double *x,
                                                              const vk_colstate *__restrict__ st, double *acc) {
    const int4 es = sv[0], ee = el[0];
    if (st && st[es.w].done) return;
    const size_t nn = (size_t)ni * ni;
    const int g = threadIdx.x / MV_NT, tt = threadIdx.x % MV_NT;
    const int lr = tt % MV_RG, a = MV_RG * g + lr, h = tt / MV_RG, ch = (ni + MV_SP - 1) / MV_SP;
    const int ks = es.x, ke = ee.x;
    // every operator up front: Q_s, A_s^-1, A_e^-1, XB_e
    double mq[MV_CH], mxs[MV_CH], mxe[MV_CH], mb[MV_CH];
    mv_load(mq, Q + (size_t)slot0 * nn, ni, a, h, ch);
    mv_load(mxs, Dinv + (size_t)ks * nn, ni, a, h, ch);
    mv_load(mxe, Dinv + (size_t)ke * nn, ni, a, h, ch);
    mv_load(mb, XB + (size_t)ke * nn, ni, a, h, ch);
    __shared__ double re[TAIL_NIMAX], rs[TAIL_NIMAX], xs[TAIL_NIMAX], pp[TAIL_G][MV_SP][MV_RG];
    for (int i = threadIdx.x; i < ni; i += blockDim.x) re[i] = r[(size_t)ke * ni + i];
    __syncthreads();
    auto total = [&]() {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < MV_SP; k++) t += pp[g][k][lr];
        return t;
    };
    // forward: r_s -= P_s r_l + Q_s r_r, no left neighbour
    pp[g][h][lr] = 0.0 + mv_dot(mq, re, ni, h, ch);
    __syncthreads();
    if (h == 0 && a < ni) {
        const double v = r[(size_t)ks * ni + a] - total();
        r[(size_t)ks * ni + a] = v;
        rs[a] = v;
    }
    __syncthreads();
    // top: x_s = A_s^-1 r_s, no neighbours
    pp[g][h][lr] = (mv_dot(mxs, rs, ni, h, ch) - 0.0) - 0.0;
    __syncthreads();
    if (h == 0 && a < ni) {
        const double t = total();
        x[(size_t)ks * ni + a] = t;
        if (acc) acc[(size_t)ks * ni + a] = acc[(size_t)ks * ni + a] + t;
        xs[a] = t;
    }
    __syncthreads();
    // back-substitution of e: x_e = A_e^-1 r_e - XB_e x_s, no right neighbour
    pp[g][h][lr] = (mv_dot(mxe, re, ni, h, ch) - mv_dot(mb, xs, ni, h, ch)) - 0.0;
    __syncthreads();
    if (h == 0 && a < ni) {
        const double t = total();
        x[(size_t)ke * ni + a] = t;
        if (acc) acc[(size_t)ke * ni + a] = acc[(size_t)ke * ni + a] + t;
    }
}

// the two product launches of a level >= 1
template <int NB, int KC>
void launch_bcr_products(int ni, int ns, int ne, int ny, const int4 *sv, const int *sflag, int slot0,
                         const int4 *el, const double *Dinv, const vk_bcr_ws *ws, const vk_colstate *st,
                         hipStream_t s) {
    hipLaunchKernelGGL((k_bcr_pq<NB, KC>), dim3(ns + ne, ny, BCR_NCS<NB>), dim3(NT), 0, s, ni, ns, sv, slot0, Dinv,
                       ws->B, ws->C, ws->P, ws->Q, el, ne, ws->XB, ws->XC, st);
    hipLaunchKernelGGL((k_bcr_update<NB, KC>), dim3(ns, ny, 2 * BCR_NCS<NB>), dim3(NT), 0, s, ni, ns, sv, sflag, slot0,
                       ws->A, ws->B, ws->C, ws->P, ws->Q, st);
}

// compute units of the current device, queried once per device; 0 if the
// query fails
int device_cus() {
    static std::atomic<int> cus[64];   // zero-initialised (static storage)
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int n = cus[dev].load(std::memory_order_relaxed);
    if (n <= 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) return 0;
        cus[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

template <int NB>
int launch_bcr_factor_n(int ni, const vk_columns *cols, const double *D, const double *Lsub, const double *Usup,
                        double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws, hipStream_t s) {
    const vk_bcr_sched &b = cols->bcr;
    // deep chunks where every workgroup of a launch is resident at once,
    // at most two per CU (with three, measured slower than shallow chunks)
    constexpr long per_cu = sizeof(SlabShared<BCR_NBH<NB>, SlabKC<NB>::deep>) <= 80 * 1024 ? 2 : 1;
    const long cap = per_cu * device_cus();
    const int ny = (ni + 31) / 32;
    const int tpr = (ni + 31) / 32;
    for (int lev = 0; lev < b.nlev; lev++) {
        const int e0 = b.elim_off[lev], ne = b.elim_off[lev + 1] - e0;
        const int s0 = b.surv_off[lev], ns = b.surv_off[lev + 1] - s0;
        // level 0 inverts its eliminated blocks from D (the level-0 update writes A)
        if (ne > 0)
            hipLaunchKernelGGL((k_bcr_invert<NB>), dim3(ne), dim3(blk_nt(NB)), 0, s, ni, ne, b.d_elim + e0,
                               (const int2 *)nullptr, cols->d_off, lev == 0 ? D : ws->A, Dinv, info, st);
        if (ns > 0) {
            if (lev == 0)
                hipLaunchKernelGGL(k_bcr_update0, dim3(ns + ne, std::min(tpr * tpr, 16)), dim3(NT), 0, s, ni, ns,
                                   b.d_surv + s0, b.d_sflag + s0, s0, Lsub, Usup, Dinv, D, ws->A, ws->B, ws->C,
                                   ws->P, ws->Q, b.d_elim + e0, ne, ws->XB, ws->XC, st);
            else if ((long)(ns + ne) * ny * BCR_NCS<NB> <= cap && 2L * ns * ny * BCR_NCS<NB> <= cap)
                launch_bcr_products<NB, SlabKC<NB>::deep>(ni, ns, ne, ny, b.d_surv + s0, b.d_sflag + s0, s0,
                                                          b.d_elim + e0, Dinv, ws, st, s);
            else
                launch_bcr_products<NB, SlabKC<NB>::shallow>(ni, ns, ne, ny, b.d_surv + s0, b.d_sflag + s0, s0,
                                                             b.d_elim + e0, Dinv, ws, st, s);
        }
    }
    hipLaunchKernelGGL((k_bcr_invert<NB>), dim3(b.ntop), dim3(blk_nt(NB)), 0, s, ni, b.ntop, (const int4 *)nullptr, b.d_top,
                       cols->d_off, ws->A, Dinv, info, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

int vk_launch_bcr_factor(int nb, int ni, const vk_columns *cols, const double *D, const double *Lsub,
                         const double *Usup, double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws,
                         hipStream_t s) {
    switch (nb) {
    case 1: return launch_bcr_factor_n<1>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 2: return launch_bcr_factor_n<2>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 3: return launch_bcr_factor_n<3>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 4: return launch_bcr_factor_n<4>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 5: return launch_bcr_factor_n<5>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 6: return launch_bcr_factor_n<6>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 7: return launch_bcr_factor_n<7>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 8: return launch_bcr_factor_n<8>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 9: return launch_bcr_factor_n<9>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case 10: return launch_bcr_factor_n<10>(ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    default: return VK_ETOOBIG;
    }
}

int vk_launch_bcr_solve(int ni, const vk_columns *cols, const double *Dinv, const double *Lsub, const double *Usup,
                        const double *in, double *out, double *tmp, const vk_colstate *st, const vk_bcr_ws *ws,
                        double *acc, hipStream_t s) {
    const vk_bcr_sched &b = cols->bcr;
    const long n = (long)cols->L * ni;
    // ni <= 128: the row-group sweeps (several workgroups per block); their
    // level 0 reads `in` and fills tmp itself, the generic sweeps start
    // from a copy
    const bool small = ni <= MV_NIMAX;
    auto bwd = small ? k_bcr_bwd_s : k_bcr_bwd;
    const int bt = small ? MV_NT : BCR_ST, gy = small ? (ni + MV_RG - 1) / MV_RG : 1;
    if (!small) hipLaunchKernelGGL(k_copy_d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, in, tmp);
    // the last level and the top in one launch (one column, ni <= 96)
    const bool tail = small && b.tail && ni <= TAIL_NIMAX;
    for (int lev = 0; lev < b.nlev; lev++) {
        const int s0 = b.surv_off[lev], ns = b.surv_off[lev + 1] - s0;
        if (ns <= 0) continue;
        if (tail && lev == b.nlev - 1) {
            const int e0 = b.elim_off[lev];
            hipLaunchKernelGGL(k_bcr_tail_s, dim3(1), dim3(MV_NT * ((ni + MV_RG - 1) / MV_RG)), 0, s, ni,
                               b.d_surv + s0, s0, b.d_elim + e0, ws->Q, Dinv, ws->XB, tmp, out, st, acc);
            break;
        }
        if (small)
            hipLaunchKernelGGL(k_bcr_fwd_s, dim3(ns, gy), dim3(bt), 0, s, ni, ns, b.d_surv + s0, s0, ws->P, ws->Q,
                               lev == 0 ? in : tmp, tmp, lev == 0 ? b.d_sflag + s0 : (const int *)nullptr, st);
        else
            hipLaunchKernelGGL(k_bcr_fwd, dim3(ns, gy), dim3(bt), 0, s, ni, ns, b.d_surv + s0, s0, ws->P, ws->Q, tmp,
                               st);
    }
    if (!tail)
        hipLaunchKernelGGL(bwd, dim3(b.ntop, gy), dim3(bt), 0, s, ni, b.ntop, (const int4 *)nullptr, b.d_top, Dinv,
                           ws->XB, ws->XC, tmp, out, st, acc);
    for (int lev = b.nlev - 1 - (tail ? 1 : 0); lev >= 0; lev--) {
        const int e0 = b.elim_off[lev], ne = b.elim_off[lev + 1] - e0;
        if (ne > 0)
            hipLaunchKernelGGL(bwd, dim3(ne, gy), dim3(bt), 0, s, ni, ne, b.d_elim + e0, (const int2 *)nullptr, Dinv,
                               ws->XB, ws->XC, tmp, out, st, acc);
    }
    VK_CHECK(hipGetLastError());
    return 0;
}
