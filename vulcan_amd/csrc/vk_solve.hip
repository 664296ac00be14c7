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

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion on FP64 MFMA (factor variant 1)
//
// The pivot steps are grouped into panels of 16 columns.  One wave runs the
// 16 pivot steps of a panel on the panel columns alone (rows spread over the
// lanes, pivot search by a wave-wide max over (|value|, row) keys, pivot row
// broadcast with readlane: no workgroup barrier inside a panel).  The other
// columns then receive the whole panel at once, as the rank-16 Schur update
//     A[i][j] = (i in P ? 0 : A[i][j]) + sum_s Pan[i][s] A_old[p_s][j]
// (P: the panel's pivot rows p_s, Pan: the factored panel columns), which is
// a 16-deep GEMM and runs on v_mfma_f64_16x16x4f64.  Three barriers per
// panel instead of one per pivot step.  The matrix lives in the MFMA
// accumulator layout: 16 x 16 tiles, wave (wr, wc) of the 2 x 2 wave grid
// owns row blocks wr, wr+2, ... and column blocks wc, wc+2, ...; lane l,
// register r of a tile holds element (4*r + l/16, l%16) (operands: A(i, k)
// in lane i + 16 k, B(k, j) in lane j + 16 k; checked by
// vk_selftest_mfma_f64).
// Pivoting is implicit (rows are not swapped): ipos[r] is the pivot step at
// which row r was chosen, perm[k] the pivot row of step k, and the inverse is
// inv[ipos[r]][perm[c]] = A[r][c] at the end.
// ---------------------------------------------------------------------------
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <int NB>
struct BlkShared {
    static constexpr int N16 = 16 * NB;
    double pan[2][16][N16 + 1];   // panel columns, [k][row], double-buffered by panel parity
    double rb[16][N16];           // old pivot rows of the panel, [s][col]
    double prow[16];              // current pivot row of the panel factorisation
    int ipos[N16];
    int perm[N16];
    int info;
};

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v & 0xffffffffull), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, 0xf, 0xf, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo;
}

__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v, int lane) {
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v & 0xffffffffull), lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// Maximum over the 64 lanes of a wave (all lanes active), returned
// wave-uniform: DPP within rows of 16 (quad swaps, half-row and row
// mirrors), then the four row maxima through readlane.
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    unsigned long long u;
    u = dpp_u64<0xB1>(v);    // quad_perm [1,0,3,2]
    v = u > v ? u : v;
    u = dpp_u64<0x4E>(v);    // quad_perm [2,3,0,1]
    v = u > v ? u : v;
    u = dpp_u64<0x141>(v);   // row_half_mirror
    v = u > v ? u : v;
    u = dpp_u64<0x140>(v);   // row_mirror
    v = u > v ? u : v;
    unsigned long long m = readlane_u64(v, 0);
    u = readlane_u64(v, 16);
    m = u > m ? u : m;
    u = readlane_u64(v, 32);
    m = u > m ? u : m;
    u = readlane_u64(v, 48);
    return u > m ? u : m;
}

// Pivot steps p0 .. p0+kend-1 on the panel pan[k][row] (one wave).
template <int NB>
__device__ __forceinline__ void blk_panel(BlkShared<NB> &sh, double (*pan)[16 * NB + 1], int ni, int p0, int kend,
                                          int tag) {
    constexpr int N16 = 16 * NB;
    constexpr int RPL = (N16 + 63) / 64;
    const int l = threadIdx.x & 63;
    double v[RPL][16];
    bool av[RPL];
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const int r = l + 64 * q;
#pragma unroll
        for (int k = 0; k < 16; k++) v[q][k] = (r < N16) ? pan[k][r] : 0.0;
        av[q] = (r < ni) && sh.ipos[r] < 0;
    }
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
        if (kk >= kend) continue;   // uniform; keeps the trip count static
        unsigned long long best = 0;
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            const int r = l + 64 * q;
            if (av[q]) {
                const unsigned long long key =
                    ((unsigned long long)__double_as_longlong(fabs(v[q][kk])) & ~0xffull) | (unsigned)(255 - r);
                best = key > best ? key : best;
            }
        }
        best = wave_max_u64(best);
        const unsigned blo = (unsigned)(best & 0xffffffffull);
        const unsigned bhi = (unsigned)(best >> 32);
        const int p = (blo | bhi) ? 255 - (int)(blo & 0xffu) : 0;
        if (((blo >> 8) | bhi) == 0u && l == 0 && sh.info == 0) sh.info = tag;   // zero (or no) pivot
        const int pl = p & 63;
        // the owner lane hands the pivot row to the wave through LDS (in order within the wave)
#pragma unroll
        for (int q = 0; q < RPL; q++)
            if (l + 64 * q == p) {
#pragma unroll
                for (int k = 0; k < 16; k++) sh.prow[k] = v[q][k];
            }
        __builtin_amdgcn_wave_barrier();
        double rs[16];
#pragma unroll
        for (int k = 0; k < 16; k++) rs[k] = sh.prow[k];
        const double inv = 1.0 / rs[kk];
#pragma unroll
        for (int k = 0; k < 16; k++) rs[k] = rs[k] * inv;
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            const double f = v[q][kk];
#pragma unroll
            for (int k = 0; k < 16; k++) v[q][k] = fma(-f, rs[k], v[q][k]);
            v[q][kk] = -f * inv;
        }
        // the pivot row itself: the scaled row, inv in the pivot column
#pragma unroll
        for (int q = 0; q < RPL; q++)
            if (l + 64 * q == p) {
#pragma unroll
                for (int k = 0; k < 16; k++) v[q][k] = rs[k];
                v[q][kk] = inv;
                av[q] = false;
            }
        __builtin_amdgcn_wave_barrier();
        if (l == pl) sh.ipos[p] = p0 + kk;
        if (l == 0) sh.perm[p0 + kk] = p;
    }
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const int r = l + 64 * q;
        if (r < N16) {
#pragma unroll
            for (int k = 0; k < 16; k++) pan[k][r] = v[q][k];
        }
    }
    for (int s = kend; s < 16; s++)
        for (int c = l; c < N16; c += 64) sh.rb[s][c] = 0.0;
}

// Inverts the matrix held in acc (MFMA tile layout) and writes the inverse
// transposed to X (X[m][i] = inv[i][m]).  A zero pivot records `tag` in
// sh.info (sticky).
template <int NB>
__device__ __forceinline__ void blk_invert(f64x4 (&acc)[(NB + 1) / 2][(NB + 1) / 2], int ni, double *__restrict__ X,
                                           BlkShared<NB> &sh, int tag) {
    constexpr int N16 = 16 * NB, TB = (NB + 1) / 2;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int wr = w >> 1, wc = w & 1;
    const int lg = l >> 4, lc = l & 15;
    for (int t = tid; t < N16; t += NT) sh.ipos[t] = -1;
    __syncthreads();
    const int npan = (ni + 15) >> 4;
    for (int pb = 0; pb < npan; pb++) {
        double(*pan)[N16 + 1] = sh.pan[pb & 1];
        const int p0 = 16 * pb, kend = min(16, ni - p0);
        // the owners of column block pb publish the panel
#pragma unroll
        for (int tj = 0; tj < TB; tj++) {
            if (wc + 2 * tj != pb) continue;   // static register index, uniform test
#pragma unroll
            for (int ti = 0; ti < TB; ti++) {
                const int bi = wr + 2 * ti;
                if (bi < NB) {
#pragma unroll
                    for (int r = 0; r < 4; r++) pan[lc][16 * bi + 4 * r + lg] = acc[ti][tj][r];
                }
            }
        }
        __syncthreads();
        if (w == 0) blk_panel<NB>(sh, pan, ni, p0, kend, tag);
        __syncthreads();
        // panel tiles take the factored panel; the other tiles hand their
        // pivot-row entries to rb and clear them
#pragma unroll
        for (int ti = 0; ti < TB; ti++) {
            const int bi = wr + 2 * ti;
            if (bi >= NB) continue;
            int sr[4];
#pragma unroll
            for (int r = 0; r < 4; r++) sr[r] = sh.ipos[16 * bi + 4 * r + lg] - p0;
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                if (bj >= NB) continue;
                if (bj == pb) {
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[ti][tj][r] = pan[lc][16 * bi + 4 * r + lg];
                } else {
#pragma unroll
                    for (int r = 0; r < 4; r++)
                        if (sr[r] >= 0 && sr[r] < 16) {
                            sh.rb[sr[r]][16 * bj + lc] = acc[ti][tj][r];
                            acc[ti][tj][r] = 0.0;
                        }
                }
            }
        }
        __syncthreads();
        // rank-16 update on MFMA
#pragma unroll
        for (int ti = 0; ti < TB; ti++) {
            const int bi = wr + 2 * ti;
            if (bi >= NB) continue;
            double a[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) a[s4] = pan[4 * s4 + (l >> 4)][16 * bi + lc];
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                if (bj >= NB || bj == pb) continue;
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4], sh.rb[4 * s4 + (l >> 4)][16 * bj + lc],
                                                                        acc[ti][tj], 0, 0, 0);
            }
        }
    }
    // inv[ipos[r]][perm[c]] = A[r][c]  ->  X[perm[c]][ipos[r]]
#pragma unroll
    for (int ti = 0; ti < TB; ti++) {
        const int bi = wr + 2 * ti;
        if (bi >= NB) continue;
        int qi[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 16 * bi + 4 * r + lg;
            qi[r] = i < ni ? sh.ipos[i] : -1;
        }
#pragma unroll
        for (int tj = 0; tj < TB; tj++) {
            const int bj = wc + 2 * tj;
            const int cc = 16 * bj + lc;
            if (bj >= NB || cc >= ni) continue;
            const size_t rowoff = (size_t)sh.perm[cc] * ni;
#pragma unroll
            for (int r = 0; r < 4; r++)
                if (qi[r] >= 0) X[rowoff + qi[r]] = acc[ti][tj][r];
        }
    }
    __syncthreads();
}

// acc = D_l - diag(cl) Xn^T diag(cn) [- diag(cl2) Xn2^T diag(cn2)] in the tile layout
template <int NB>
__device__ __forceinline__ void blk_load_schur(f64x4 (&acc)[(NB + 1) / 2][(NB + 1) / 2], int ni,
                                               const double *__restrict__ Dl, const double *cl, const double *Xn,
                                               const double *cn, const double *cl2, const double *Xn2,
                                               const double *cn2) {
    constexpr int TB = (NB + 1) / 2;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int wr = w >> 1, wc = w & 1;
    const int lg = l >> 4, lc = l & 15;
#pragma unroll
    for (int ti = 0; ti < TB; ti++) {
        const int bi = wr + 2 * ti;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 16 * bi + 4 * r + lg;
            const bool rowok = bi < NB && i < ni;
            const double li = (Xn && rowok) ? cl[i] : 0.0;
            const double li2 = (Xn2 && rowok) ? cl2[i] : 0.0;
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                const int cc = 16 * bj + lc;
                double v = 0.0;
                if (rowok && bj < NB && cc < ni) {
                    v = Dl[(size_t)i * ni + cc];
                    if (Xn) v = v - li * (Xn[(size_t)cc * ni + i] * cn[cc]);
                    if (Xn2) v = v - li2 * (Xn2[(size_t)cc * ni + i] * cn2[cc]);
                }
                acc[ti][tj][r] = v;
            }
        }
    }
}

template <int NB, int MINW>
__global__ void __launch_bounds__(NT, MINW) k_factor_blk(int ni, const int *__restrict__ off, const double *__restrict__ D,
                                                    const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                    double *__restrict__ Dinv, int *__restrict__ info,
                                                    const vk_colstate *__restrict__ st) {
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    __shared__ BlkShared<NB> sh;
    if (threadIdx.x == 0) sh.info = 0;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    f64x4 acc[(NB + 1) / 2][(NB + 1) / 2];
    if (side == 0) {
        for (int j = 0; j < m; ++j) {
            const size_t l = (size_t)l0 + j;
            blk_load_schur<NB>(acc, ni, D + l * nn, Lsub + l * ni, j > 0 ? Dinv + (l - 1) * nn : nullptr,
                               j > 0 ? Usup + (l - 1) * ni : nullptr, nullptr, nullptr, nullptr);
            blk_invert<NB>(acc, ni, Dinv + l * nn, sh, j + 1);
        }
    } else {
        for (int j = nz - 1; j > m; --j) {
            const size_t l = (size_t)l0 + j;
            blk_load_schur<NB>(acc, ni, D + l * nn, Usup + l * ni, j < nz - 1 ? Dinv + (l + 1) * nn : nullptr,
                               j < nz - 1 ? Lsub + (l + 1) * ni : nullptr, nullptr, nullptr, nullptr);
            blk_invert<NB>(acc, ni, Dinv + l * nn, sh, j + 1);
        }
    }
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

template <int NB, int MINW>
__global__ void __launch_bounds__(NT, MINW) k_factor_blk_mid(int ni, const int *__restrict__ off,
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
    f64x4 acc[(NB + 1) / 2][(NB + 1) / 2];
    blk_load_schur<NB>(acc, ni, D + l * nn, Lsub + l * ni, m > 0 ? Dinv + (l - 1) * nn : nullptr,
                       m > 0 ? Usup + (l - 1) * ni : nullptr, Usup + l * ni,
                       m < nz - 1 ? Dinv + (l + 1) * nn : nullptr, m < nz - 1 ? Lsub + (l + 1) * ni : nullptr);
    blk_invert<NB>(acc, ni, Dinv + l * nn, sh, m + 1);
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

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

// ---------------------------------------------------------------------------
// Block cyclic reduction (factor variant 4)
//
// Every level eliminates the odd-positioned active layers of each column at
// once: their blocks are inverted in parallel (blk_invert), and each
// surviving layer j with eliminated neighbours l (left) and r (right) takes
//     P_j = B_j A_l^-1,   Q_j = C_j A_r^-1,
//     A_j -= P_j C_l + Q_j B_r,   B_j = -P_j B_l,   C_j = -Q_j C_r
// (B: coupling to the previous active layer, C: to the next; diagonal at
// level 0, dense above).  The chain over layers becomes log2(nz) levels of
// independent work: the Schur-complement updates above level 0 are dense
// n x n products on FP64 MFMA, level 0 is elementwise.  P and Q are kept
// for the right-hand side reduction of the solves.
// ---------------------------------------------------------------------------

// acc += op(A) op(B) over k < ni, A(a, c) = Ab[a*sar + c*sac], B(c, b) = Bb[c*sbr + b*sbc],
// operands staged through LDS in slices of 16 (tile layout of blk_invert).
template <int NB>
struct GemmShared {
    double sa[16][16 * NB + 1];
    double sb[16][16 * NB];
};

template <int NB>
__device__ __forceinline__ void blk_gemm_acc(f64x4 (&acc)[(NB + 1) / 2][(NB + 1) / 2], int ni,
                                             const double *__restrict__ Ab, long sar, long sac,
                                             const double *__restrict__ Bb, long sbr, long sbc, GemmShared<NB> &g) {
    constexpr int N16 = 16 * NB, TB = (NB + 1) / 2;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int wr = w >> 1, wc = w & 1, lc = l & 15;
    for (int k0 = 0; k0 < ni; k0 += 16) {
        __syncthreads();
        for (int e = tid; e < 16 * N16; e += NT) {
            const int kk = e / N16, rr = e % N16, c = k0 + kk;
            g.sa[kk][rr] = (rr < ni && c < ni) ? Ab[(long)rr * sar + (long)c * sac] : 0.0;
            g.sb[kk][rr] = (rr < ni && c < ni) ? Bb[(long)c * sbr + (long)rr * sbc] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int ti = 0; ti < TB; ti++) {
            const int bi = wr + 2 * ti;
            if (bi >= NB) continue;
            double a[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) a[s4] = g.sa[4 * s4 + (l >> 4)][16 * bi + lc];
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                if (bj >= NB) continue;
#pragma unroll
                for (int s4 = 0; s4 < 4; s4++)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s4], g.sb[4 * s4 + (l >> 4)][16 * bj + lc],
                                                                        acc[ti][tj], 0, 0, 0);
            }
        }
    }
    __syncthreads();
}

template <int NB>
__device__ __forceinline__ void blk_zero(f64x4 (&acc)[(NB + 1) / 2][(NB + 1) / 2]) {
#pragma unroll
    for (int ti = 0; ti < (NB + 1) / 2; ti++)
#pragma unroll
        for (int tj = 0; tj < (NB + 1) / 2; tj++) acc[ti][tj] = f64x4{0.0, 0.0, 0.0, 0.0};
}

// out[a][b] = sgn * acc (+ add[a][b] if add), row-major ni x ni
template <int NB>
__device__ __forceinline__ void blk_store(const f64x4 (&acc)[(NB + 1) / 2][(NB + 1) / 2], int ni,
                                          double *__restrict__ out, double sgn, const double *add) {
    constexpr int TB = (NB + 1) / 2;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int wr = w >> 1, wc = w & 1, lg = l >> 4, lc = l & 15;
#pragma unroll
    for (int ti = 0; ti < TB; ti++) {
        const int bi = wr + 2 * ti;
        if (bi >= NB) continue;
#pragma unroll
        for (int tj = 0; tj < TB; tj++) {
            const int bj = wc + 2 * tj, cc = 16 * bj + lc;
            if (bj >= NB || cc >= ni) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int i = 16 * bi + 4 * r + lg;
                if (i < ni) {
                    const size_t e = (size_t)i * ni + cc;
                    out[e] = add ? add[e] + sgn * acc[ti][tj][r] : sgn * acc[ti][tj][r];
                }
            }
        }
    }
}

// invert A[layer] for the listed layers (the eliminated ones of a level, or
// the top layer of each column) into Dinv[layer] (transposed storage)
template <int NB>
__global__ void __launch_bounds__(NT, 2) k_bcr_invert(int ni, int n, const int4 *__restrict__ el,
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
    f64x4 acc[(NB + 1) / 2][(NB + 1) / 2];
    blk_load_schur<NB>(acc, ni, A + layer * nn, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    blk_invert<NB>(acc, ni, Dinv + layer * nn, sh, layer - off[c] + 1);
    if (threadIdx.x == 0 && info && sh.info) atomicCAS(&info[c], 0, sh.info);
}

// level 0: couplings are the diagonals Lsub (B) and Usup (C); elementwise
__global__ void k_bcr_update0(int ni, int n, const int4 *__restrict__ sv, const int *__restrict__ sflag,
                              int slot0, const double *__restrict__ Lsub, const double *__restrict__ Usup,
                              const double *__restrict__ Dinv, double *__restrict__ A, double *__restrict__ B,
                              double *__restrict__ Cc, double *__restrict__ P, double *__restrict__ Q,
                              const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const int fl = sflag[q];
    const size_t nn = (size_t)ni * ni, j = e.x;
    const size_t slot = (size_t)(slot0 + q) * nn;
    const double *XL = e.y >= 0 ? Dinv + (size_t)e.y * nn : nullptr;   // X[m][i] = inv[i][m]
    const double *XR = e.z >= 0 ? Dinv + (size_t)e.z * nn : nullptr;
    for (int t = threadIdx.x; t < ni * ni; t += blockDim.x) {
        const int a = t / ni, b = t % ni;
        double upd = 0.0;
        if (XL) {
            const double p = Lsub[j * ni + a] * XL[(size_t)b * ni + a];
            P[slot + t] = p;
            upd = p * Usup[(size_t)e.y * ni + b];
            B[j * nn + t] = (fl & 1) ? -(p * Lsub[(size_t)e.y * ni + b]) : 0.0;
        } else {
            B[j * nn + t] = 0.0;
        }
        if (XR) {
            const double qv = Usup[j * ni + a] * XR[(size_t)b * ni + a];
            Q[slot + t] = qv;
            upd = upd + qv * Lsub[(size_t)e.z * ni + b];
            Cc[j * nn + t] = (fl & 2) ? -(qv * Usup[(size_t)e.z * ni + b]) : 0.0;
        } else {
            Cc[j * nn + t] = 0.0;
        }
        A[j * nn + t] = A[j * nn + t] - upd;
    }
}

// levels >= 1: dense couplings, products on MFMA; one workgroup per surviving layer
template <int NB>
__global__ void __launch_bounds__(NT, 2) k_bcr_update(int ni, int n, const int4 *__restrict__ sv,
                                                       const int *__restrict__ sflag, int slot0,
                                                       const double *__restrict__ Dinv, double *__restrict__ A,
                                                       double *__restrict__ B, double *__restrict__ Cc,
                                                       double *__restrict__ P, double *__restrict__ Q,
                                                       const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const int fl = sflag[q];
    const size_t nn = (size_t)ni * ni, j = e.x;
    double *Pj = P + (size_t)(slot0 + q) * nn, *Qj = Q + (size_t)(slot0 + q) * nn;
    __shared__ GemmShared<NB> g;
    f64x4 acc[(NB + 1) / 2][(NB + 1) / 2];
    // P = B_j X_l, Q = C_j X_r (inverse stored transposed: inv(c, b) = X[b][c])
    if (e.y >= 0) {
        blk_zero<NB>(acc);
        blk_gemm_acc<NB>(acc, ni, B + j * nn, ni, 1, Dinv + (size_t)e.y * nn, 1, ni, g);
        blk_store<NB>(acc, ni, Pj, 1.0, nullptr);
    }
    if (e.z >= 0) {
        blk_zero<NB>(acc);
        blk_gemm_acc<NB>(acc, ni, Cc + j * nn, ni, 1, Dinv + (size_t)e.z * nn, 1, ni, g);
        blk_store<NB>(acc, ni, Qj, 1.0, nullptr);
    }
    __threadfence_block();
    __syncthreads();
    // A_j -= P C_l + Q B_r
    blk_zero<NB>(acc);
    if (e.y >= 0) blk_gemm_acc<NB>(acc, ni, Pj, ni, 1, Cc + (size_t)e.y * nn, ni, 1, g);
    if (e.z >= 0) blk_gemm_acc<NB>(acc, ni, Qj, ni, 1, B + (size_t)e.z * nn, ni, 1, g);
    blk_store<NB>(acc, ni, A + j * nn, -1.0, A + j * nn);
    // new couplings: B_j = -P B_l (to l's left neighbour), C_j = -Q C_r
    blk_zero<NB>(acc);
    if (fl & 1) blk_gemm_acc<NB>(acc, ni, Pj, ni, 1, B + (size_t)e.y * nn, ni, 1, g);
    blk_store<NB>(acc, ni, B + j * nn, -1.0, nullptr);
    blk_zero<NB>(acc);
    if (fl & 2) blk_gemm_acc<NB>(acc, ni, Qj, ni, 1, Cc + (size_t)e.z * nn, ni, 1, g);
    blk_store<NB>(acc, ni, Cc + j * nn, -1.0, nullptr);
}

__global__ void k_copy_d(long n, const double *__restrict__ a, double *__restrict__ b) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// solves: forward reduction r_j -= P_j r_l + Q_j r_r at level s
__global__ void k_bcr_fwd(int ni, int n, const int4 *__restrict__ sv, int slot0, const double *__restrict__ P,
                          const double *__restrict__ Q, double *__restrict__ r, const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    const int4 e = sv[q];
    if (st && st[e.w].done) return;
    const size_t nn = (size_t)ni * ni;
    const double *Pj = P + (size_t)(slot0 + q) * nn, *Qj = Q + (size_t)(slot0 + q) * nn;
    __shared__ double rl[VK_NI_MAX], rr[VK_NI_MAX];
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        rl[i] = e.y >= 0 ? r[(size_t)e.y * ni + i] : 0.0;
        rr[i] = e.z >= 0 ? r[(size_t)e.z * ni + i] : 0.0;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < ni; a += blockDim.x) {
        double v = r[(size_t)e.x * ni + a];
        if (e.y >= 0) {
            double s = 0.0;
            for (int b = 0; b < ni; b++) s = fma(Pj[(size_t)a * ni + b], rl[b], s);
            v = v - s;
        }
        if (e.z >= 0) {
            double s = 0.0;
            for (int b = 0; b < ni; b++) s = fma(Qj[(size_t)a * ni + b], rr[b], s);
            v = v - s;
        }
        r[(size_t)e.x * ni + a] = v;
    }
}

// x_k = A_k^-1 (r_k - B_k x_l - C_k x_r) for the eliminated layers of level s
// (level 0: diagonal couplings Lsub, Usup); top: x = A^-1 r
__global__ void k_bcr_bwd(int ni, int n, const int4 *__restrict__ el, const int2 *__restrict__ top, int level,
                          const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                          const double *__restrict__ Usup, const double *__restrict__ B, const double *__restrict__ Cc,
                          const double *__restrict__ r, double *__restrict__ x, const vk_colstate *__restrict__ st) {
    const int q = blockIdx.x;
    if (q >= n) return;
    int k, left = -1, right = -1, c;
    if (el) { const int4 e = el[q]; k = e.x; left = e.y; right = e.z; c = e.w; }
    else { k = top[q].x; c = top[q].y; }
    if (st && st[c].done) return;
    const size_t nn = (size_t)ni * ni;
    __shared__ double w[VK_NI_MAX], xl[VK_NI_MAX], xr[VK_NI_MAX];
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        xl[i] = left >= 0 ? x[(size_t)left * ni + i] : 0.0;
        xr[i] = right >= 0 ? x[(size_t)right * ni + i] : 0.0;
    }
    __syncthreads();
    for (int a = threadIdx.x; a < ni; a += blockDim.x) {
        double v = r[(size_t)k * ni + a];
        if (level == 0) {
            if (left >= 0) v = v - Lsub[(size_t)k * ni + a] * xl[a];
            if (right >= 0) v = v - Usup[(size_t)k * ni + a] * xr[a];
        } else {
            if (left >= 0) {
                double s = 0.0;
                for (int b = 0; b < ni; b++) s = fma(B[(size_t)k * nn + (size_t)a * ni + b], xl[b], s);
                v = v - s;
            }
            if (right >= 0) {
                double s = 0.0;
                for (int b = 0; b < ni; b++) s = fma(Cc[(size_t)k * nn + (size_t)a * ni + b], xr[b], s);
                v = v - s;
            }
        }
        w[a] = v;
    }
    __syncthreads();
    const double *X = Dinv + (size_t)k * nn;   // X[m][a] = inv[a][m]
    for (int a = threadIdx.x; a < ni; a += blockDim.x) {
        double s = 0.0;
        for (int m = 0; m < ni; m++) s = fma(X[(size_t)m * ni + a], w[m], s);
        x[(size_t)k * ni + a] = s;
    }
}

// D = A B for A[i][k] = i + 0.25 k + 1, B[k][j] = 1 - 0.5 k + 0.125 j (16 x 4 x 16)
// through one v_mfma_f64_16x16x4f64 with the operand layout the blocked
// inversion assumes; out[i*16+j] = D[i][j].
__global__ void k_selftest_mfma_f64(double *out) {
    const int l = threadIdx.x & 63;
    const double a = (double)(l & 15) + 0.25 * (double)(l >> 4) + 1.0;
    const double b = 1.0 - 0.5 * (double)(l >> 4) + 0.125 * (double)(l & 15);
    f64x4 acc = {0.0, 0.0, 0.0, 0.0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; r++) out[(4 * r + (l >> 4)) * 16 + (l & 15)] = acc[r];
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

__global__ void k_zero_int(int n, int *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Twisted block solves, 1024-thread workgroups.  Thread (i, h) = (tid % RP,
// tid / RP) owns species row i and the h-th of SP column chunks of each
// inverse (RP = 128 or 256 rows, SP = 8 or 4 parts), so one layer's
// matrix-vector product is SP-way split and the
// Dinv chunks of the next three layers are in flight while the current layer
// is processed (the chain over layers is latency bound, not bandwidth bound).
// Both sweeps stay within the column: layer indices run over [0, nz).
// ---------------------------------------------------------------------------
constexpr int NTS = 1024;

// rows per part (a power of two >= 16 NB) and column chunk per thread
template <int NB>
struct SolveShape {
    static constexpr int RP = 16 * NB <= 128 ? 128 : 256;
    static constexpr int SP = NTS / RP;
    static constexpr int CH = (16 * NB + SP - 1) / SP;
};

template <int CH>
__device__ __forceinline__ void load_chunk(double (&r)[CH], const double *__restrict__ Xt, int ni, int i, int m0) {
#pragma unroll
    for (int q = 0; q < CH; q++) {
        const int mm = m0 + q;
        r[q] = (i < ni && mm < ni) ? Xt[(size_t)mm * ni + i] : 0.0;
    }
}

// layers in flight per thread: fewer for larger blocks (128 registers per thread at 1024 threads)
template <int NB>
struct SolveDepth {
    static constexpr int value = NB <= 5 ? 3 : (NB <= 7 ? 2 : 1);
};

template <int DEP, int CH>
__device__ __forceinline__ void rotate(double (&b)[DEP][CH]) {
#pragma unroll
    for (int d = 0; d + 1 < DEP; d++) {
#pragma unroll
        for (int q = 0; q < CH; q++) b[d][q] = b[d + 1][q];
    }
}

// y_i = sum_m Xt[m][i] v[m] from the registers of the eight row parts; the
// result is valid in threads h == 0 (i < ni).  Two barriers.
template <int SP, int RP, int CH>
__device__ __forceinline__ double solve_matvec(const double (&cur)[CH], const double *v, double (*part)[RP], int i,
                                               int h, int m0) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < CH; q++) acc = fma(cur[q], v[m0 + q], acc);
    part[h][i] = acc;
    __syncthreads();
    double y = part[0][i];
#pragma unroll
    for (int hh = 1; hh < SP; hh++) y += part[hh][i];
    return y;
}

template <int NB>
__global__ void __launch_bounds__(NTS) k_solve_fwd2(int ni, const int *__restrict__ off,
                                                     const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                     const double *__restrict__ Usup, const double *in, double *tmp,
                                                     const vk_colstate *__restrict__ st) {
    constexpr int RP = SolveShape<NB>::RP, SP = SolveShape<NB>::SP, CH = SolveShape<NB>::CH;
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x, i = tid % RP, h = tid / RP, m0 = h * CH;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    __shared__ double v[SP * CH];
    __shared__ double part[SP][RP];
    for (int q = tid; q < SP * CH; q += NTS) v[q] = 0.0;
    const int j0 = side == 0 ? 0 : nz - 1, dj = side == 0 ? 1 : -1;
    const int nsteps = side == 0 ? m : nz - 1 - m;
    const double *cpl = side == 0 ? Lsub : Usup;
    constexpr int DEP = SolveDepth<NB>::value;
    double buf[DEP][CH];
#pragma unroll
    for (int d = 0; d < DEP; d++)
        if (d < nsteps) load_chunk<CH>(buf[d], Dinv + (size_t)(l0 + j0 + d * dj) * nn, ni, i, m0);
    double zi = 0.0;
    __syncthreads();
    for (int s = 0; s < nsteps; s++) {
        const size_t l = (size_t)(l0 + j0 + dj * s);
        if (h == 0 && i < ni) {
            double w = in[l * ni + i];
            if (s > 0) w = w - cpl[l * ni + i] * zi;
            v[i] = w;
        }
        __syncthreads();
        const double z = solve_matvec<SP, RP, CH>(buf[0], v, part, i, h, m0);
        if (h == 0 && i < ni) {
            tmp[l * ni + i] = z;
            zi = z;
        }
        rotate<DEP, CH>(buf);
        if (s + DEP < nsteps) load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)(l0 + j0 + dj * (s + DEP)) * nn, ni, i, m0);
        __syncthreads();
    }
    // park b_m in the unused slot m so the backward kernel never reads `in`
    // (which may alias `out`)
    if (side == 0 && h == 0 && i < ni) {
        const size_t lm = (size_t)l0 + m;
        tmp[lm * ni + i] = in[lm * ni + i];
    }
}

template <int NB>
__global__ void __launch_bounds__(NTS) k_solve_bwd2(int ni, const int *__restrict__ off,
                                                     const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                     const double *__restrict__ Usup, const double *tmp, double *out,
                                                     const vk_colstate *__restrict__ st) {
    constexpr int RP = SolveShape<NB>::RP, SP = SolveShape<NB>::SP, CH = SolveShape<NB>::CH;
    const int c = blockIdx.x >> 1, side = blockIdx.x & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x, i = tid % RP, h = tid / RP, m0 = h * CH;
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    __shared__ double v[SP * CH];
    __shared__ double part[SP][RP];
    for (int q = tid; q < SP * CH; q += NTS) v[q] = 0.0;
    const size_t lm = (size_t)l0 + m;
    const int dj = side == 0 ? -1 : 1;
    const int nsteps = 1 + (side == 0 ? m : nz - 1 - m);     // the middle layer, then the chain outwards
    const double *cpl = side == 0 ? Usup : Lsub;
    constexpr int DEP = SolveDepth<NB>::value;
    double buf[DEP][CH];
#pragma unroll
    for (int d = 0; d < DEP; d++)
        if (d < nsteps) load_chunk<CH>(buf[d], Dinv + (size_t)((long)lm + (long)d * dj) * nn, ni, i, m0);
    // x_m = M^-1 (b_m - Lsub_m z_{m-1} - Usup_m w_{m+1}); both sides compute it
    if (h == 0 && i < ni) {
        double w = tmp[lm * ni + i];
        if (m > 0) w = w - Lsub[lm * ni + i] * tmp[(lm - 1) * ni + i];
        if (m < nz - 1) w = w - Usup[lm * ni + i] * tmp[(lm + 1) * ni + i];
        v[i] = w;
    }
    __syncthreads();
    double xi = solve_matvec<SP, RP, CH>(buf[0], v, part, i, h, m0);
    if (side == 0 && h == 0 && i < ni) out[lm * ni + i] = xi;
    rotate<DEP, CH>(buf);
    if (DEP < nsteps) load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)((long)lm + (long)DEP * dj) * nn, ni, i, m0);
    __syncthreads();
    for (int s = 1; s < nsteps; s++) {
        const size_t l = lm + (long)dj * s;
        if (h == 0 && i < ni) v[i] = cpl[l * ni + i] * xi;
        __syncthreads();
        const double y = solve_matvec<SP, RP, CH>(buf[0], v, part, i, h, m0);
        if (h == 0 && i < ni) {
            xi = tmp[l * ni + i] - y;
            out[l * ni + i] = xi;
        }
        rotate<DEP, CH>(buf);
        if (s + DEP < nsteps)
            load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)((long)lm + (long)dj * (s + DEP)) * nn, ni, i, m0);
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

// r = b - A x, 32 rows of one layer per workgroup: the rows of D are staged
// in LDS with coalesced loads, 8 threads share a row (each a double-double
// partial over an eighth of the columns), and the partials, b and the two
// coupling terms are combined in double-double.
constexpr int RES_ROWS = 32, RES_PARTS = NT / RES_ROWS;

__device__ inline void dd_add(double &hi, double &lo, double a_hi, double a_lo) {
    const double s = hi + a_hi;
    const double bb = s - hi;
    const double e = (hi - (s - bb)) + (a_hi - bb);     // hi + a_hi = s + e exactly
    hi = s;
    lo = lo + (e + a_lo);
}

__global__ void __launch_bounds__(NT) k_residual(int ni, const int *__restrict__ off,
                                                  const int *__restrict__ layer_col, const double *__restrict__ D,
                                                  const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r, const vk_colstate *__restrict__ st) {
    extern __shared__ double rsm[];
    double *Ds = rsm;                        // RES_ROWS x ni
    double *xs = rsm + RES_ROWS * ni;        // ni
    double *ph = xs + ni;                    // RES_ROWS x RES_PARTS partial sums (hi)
    double *pl = ph + NT;                    // (lo)
    const int l = blockIdx.x;
    const int c = layer_col[l];
    if (st && st[c].done) return;
    const int j = l - off[c], nz = off[c + 1] - off[c];
    const int i0 = blockIdx.y * RES_ROWS;
    const int nrows = min(RES_ROWS, ni - i0);
    const double *Dl = D + ((size_t)l * ni + i0) * ni;
    const int tid = threadIdx.x;
    for (int e = tid; e < nrows * ni; e += NT) Ds[e] = Dl[e];
    for (int m = tid; m < ni; m += NT) xs[m] = x[(size_t)l * ni + m];
    __syncthreads();
    const int rr = tid / RES_PARTS, pp = tid % RES_PARTS;
    const int chunk = (ni + RES_PARTS - 1) / RES_PARTS;
    double hi = 0.0, lo = 0.0;
    if (rr < nrows) {
        const int m1 = min(ni, (pp + 1) * chunk);
        for (int m = pp * chunk; m < m1; m++) dd_sub_prod(hi, lo, Ds[rr * ni + m], xs[m]);
    }
    ph[tid] = hi;
    pl[tid] = lo;
    __syncthreads();
    if (pp == 0 && rr < nrows) {
        const int i = i0 + rr;
        double h = b[(size_t)l * ni + i], o = 0.0;
        for (int q = 0; q < RES_PARTS; q++) dd_add(h, o, ph[tid + q], pl[tid + q]);
        if (j > 0) dd_sub_prod(h, o, Lsub[(size_t)l * ni + i], x[(size_t)(l - 1) * ni + i]);
        if (j < nz - 1) dd_sub_prod(h, o, Usup[(size_t)l * ni + i], x[(size_t)(l + 1) * ni + i]);
        r[(size_t)l * ni + i] = h + o;
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
void launch_factor_n(int variant, int ni, int ncols, const int *off, const double *D, const double *Lsub,
                     const double *Usup, double *Dinv, int *info, const vk_colstate *st, hipStream_t s) {
    if (variant == VK_FACTOR_BLOCKED) {
        hipLaunchKernelGGL((k_factor_blk<N, 1>), dim3(2 * ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                           st);
        hipLaunchKernelGGL((k_factor_blk_mid<N, 1>), dim3(ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                           st);
    } else if (variant == VK_FACTOR_WAVE && N <= 7) {
        hipLaunchKernelGGL((k_factor_wave<(N <= 7 ? N : 7)>), dim3(2 * ncols), dim3(NTW), 0, s, ni, off, D, Lsub,
                           Usup, Dinv, info, st);
        hipLaunchKernelGGL((k_factor_wave_mid<(N <= 7 ? N : 7)>), dim3(ncols), dim3(NTW), 0, s, ni, off, D, Lsub,
                           Usup, Dinv, info, st);
    } else if (variant == VK_FACTOR_BLOCKED_OCC2) {
        hipLaunchKernelGGL((k_factor_blk<N, 2>), dim3(2 * ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                           st);
        hipLaunchKernelGGL((k_factor_blk_mid<N, 2>), dim3(ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                           st);
    } else {
        hipLaunchKernelGGL((k_factor<N, N>), dim3(2 * ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info, st);
        hipLaunchKernelGGL((k_factor_mid<N, N>), dim3(ncols), dim3(NT), 0, s, ni, off, D, Lsub, Usup, Dinv, info,
                           st);
    }
}

}  // namespace

namespace {
template <int NB>
int launch_bcr_factor_n(int ni, const vk_columns *cols, const double *D, const double *Lsub, const double *Usup,
                        double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws, hipStream_t s) {
    const vk_bcr_sched &b = cols->bcr;
    const long nd = (long)cols->L * ni * ni;
    hipLaunchKernelGGL(k_copy_d, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, nd, D, ws->A);
    for (int lev = 0; lev < b.nlev; lev++) {
        const int e0 = b.elim_off[lev], ne = b.elim_off[lev + 1] - e0;
        const int s0 = b.surv_off[lev], ns = b.surv_off[lev + 1] - s0;
        if (ne > 0)
            hipLaunchKernelGGL((k_bcr_invert<NB>), dim3(ne), dim3(NT), 0, s, ni, ne, b.d_elim + e0,
                               (const int2 *)nullptr, cols->d_off, ws->A, Dinv, info, st);
        if (ns > 0) {
            if (lev == 0)
                hipLaunchKernelGGL(k_bcr_update0, dim3(ns), dim3(NT), 0, s, ni, ns, b.d_surv + s0, b.d_sflag + s0, s0,
                                   Lsub, Usup, Dinv, ws->A, ws->B, ws->C, ws->P, ws->Q, st);
            else
                hipLaunchKernelGGL((k_bcr_update<NB>), dim3(ns), dim3(NT), 0, s, ni, ns, b.d_surv + s0,
                                   b.d_sflag + s0, s0, Dinv, ws->A, ws->B, ws->C, ws->P, ws->Q, st);
        }
    }
    hipLaunchKernelGGL((k_bcr_invert<NB>), dim3(b.ntop), dim3(NT), 0, s, ni, b.ntop, (const int4 *)nullptr, b.d_top,
                       cols->d_off, ws->A, Dinv, info, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

int launch_bcr_solve(int ni, const vk_columns *cols, const double *Dinv, const double *Lsub, const double *Usup,
                     const double *in, double *out, double *tmp, const vk_colstate *st, const vk_bcr_ws *ws,
                     hipStream_t s) {
    const vk_bcr_sched &b = cols->bcr;
    const long n = (long)cols->L * ni;
    hipLaunchKernelGGL(k_copy_d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, in, tmp);
    const int bt = ni <= 64 ? 64 : (ni <= 128 ? 128 : 192);
    for (int lev = 0; lev < b.nlev; lev++) {
        const int s0 = b.surv_off[lev], ns = b.surv_off[lev + 1] - s0;
        if (ns > 0)
            hipLaunchKernelGGL(k_bcr_fwd, dim3(ns), dim3(bt), 0, s, ni, ns, b.d_surv + s0, s0, ws->P, ws->Q, tmp, st);
    }
    hipLaunchKernelGGL(k_bcr_bwd, dim3(b.ntop), dim3(bt), 0, s, ni, b.ntop, (const int4 *)nullptr, b.d_top, 0, Dinv,
                       Lsub, Usup, ws->B, ws->C, tmp, out, st);
    for (int lev = b.nlev - 1; lev >= 0; lev--) {
        const int e0 = b.elim_off[lev], ne = b.elim_off[lev + 1] - e0;
        if (ne > 0)
            hipLaunchKernelGGL(k_bcr_bwd, dim3(ne), dim3(bt), 0, s, ni, ne, b.d_elim + e0, (const int2 *)nullptr, lev,
                               Dinv, Lsub, Usup, ws->B, ws->C, tmp, out, st);
    }
    VK_CHECK(hipGetLastError());
    return 0;
}
}  // namespace

int vk_launch_factor(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                     const double *Usup, double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws,
                     hipStream_t s) {
    const int ni = plan->ni;
    const int n = (ni + 15) / 16;
    const int nc = cols->ncols;
    const int v = plan->factor_variant;
    if (info) hipLaunchKernelGGL(k_zero_int, dim3((nc + 255) / 256), dim3(256), 0, s, nc, info);
    if (v == VK_FACTOR_BCR) {
        if (!ws || !ws->A) return VK_EINVAL;
        switch (n) {
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
    switch (n) {
    case 1: launch_factor_n<1>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 2: launch_factor_n<2>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 3: launch_factor_n<3>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 4: launch_factor_n<4>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 5: launch_factor_n<5>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 6: launch_factor_n<6>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 7: launch_factor_n<7>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 8: launch_factor_n<8>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 9: launch_factor_n<9>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    case 10: launch_factor_n<10>(v, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s); break;
    default: return VK_ETOOBIG;
    }
    VK_CHECK(hipGetLastError());
    return 0;
}

// out = lhs^-1 in (tmp: L*ni scratch, may not alias in/out; in may alias out)
namespace {
template <int NB>
void launch_solve_n(int ni, int ncols, const int *off, const double *Dinv, const double *Lsub, const double *Usup,
                    const double *in, double *out, double *tmp, const vk_colstate *st, hipStream_t s) {
    hipLaunchKernelGGL((k_solve_fwd2<NB>), dim3(2 * ncols), dim3(NTS), 0, s, ni, off, Dinv, Lsub, Usup, in, tmp, st);
    hipLaunchKernelGGL((k_solve_bwd2<NB>), dim3(2 * ncols), dim3(NTS), 0, s, ni, off, Dinv, Lsub, Usup, tmp, out, st);
}
}  // namespace

int vk_launch_solve(const vk_plan *plan, const vk_columns *cols, const double *Dinv, const double *Lsub,
                    const double *Usup, const double *in, double *out, double *tmp, const vk_colstate *st,
                    const vk_bcr_ws *ws, hipStream_t s) {
    const int ni = plan->ni, nc = cols->ncols;
    const int *off = cols->d_off;
    if (plan->factor_variant == VK_FACTOR_BCR) {
        if (!ws || !ws->A) return VK_EINVAL;
        return launch_bcr_solve(ni, cols, Dinv, Lsub, Usup, in, out, tmp, st, ws, s);
    }
    switch ((ni + 15) / 16) {
    case 1: launch_solve_n<1>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 2: launch_solve_n<2>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 3: launch_solve_n<3>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 4: launch_solve_n<4>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 5: launch_solve_n<5>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 6: launch_solve_n<6>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 7: launch_solve_n<7>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 8: launch_solve_n<8>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 9: launch_solve_n<9>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    case 10: launch_solve_n<10>(ni, nc, off, Dinv, Lsub, Usup, in, out, tmp, st, s); break;
    default: return VK_ETOOBIG;
    }
    VK_CHECK(hipGetLastError());
    return 0;
}

int vk_launch_residual(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                       const double *Usup, const double *x, const double *b, double *r, const vk_colstate *st,
                       hipStream_t s) {
    const int ni = plan->ni;
    const size_t smem = sizeof(double) * ((size_t)RES_ROWS * ni + ni + 2 * NT);
    hipLaunchKernelGGL(k_residual, dim3(cols->L, (ni + RES_ROWS - 1) / RES_ROWS), dim3(NT), smem, s, ni, cols->d_off,
                       cols->d_layer_col, D, Lsub, Usup, x, b, r, st);
    VK_CHECK(hipGetLastError());
    return 0;
}

// x = A^-1 b with `refine` steps of iterative refinement (work, tmp: L*ni scratch each)
int vk_launch_solve_refined(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Dinv,
                            const double *Lsub, const double *Usup, const double *b, double *x, double *work,
                            double *tmp, int refine, const vk_colstate *st, const vk_bcr_ws *ws, hipStream_t s) {
    int r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, b, x, tmp, st, ws, s);
    if (r) return r;
    const int n = cols->L * plan->ni;
    for (int q = 0; q < refine; q++) {
        r = vk_launch_residual(plan, cols, D, Lsub, Usup, x, b, work, st, s);
        if (r) return r;
        r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, work, work, tmp, st, ws, s);
        if (r) return r;
        hipLaunchKernelGGL(k_axpy, dim3((n + 255) / 256), dim3(256), 0, s, n, plan->ni, cols->d_layer_col, work, x,
                           st);
        VK_CHECK(hipGetLastError());
    }
    return 0;
}

// stream-ordered workspace of the cyclic reduction (freed with the object)
struct BcrTmp {
    vk_bcr_ws ws;
    hipStream_t s;
    explicit BcrTmp(hipStream_t st) : s(st) {}
    int alloc(const vk_plan *plan, const vk_columns *cols) {
        const size_t nn = (size_t)plan->ni * plan->ni, L = (size_t)cols->L, ns = (size_t)cols->bcr.nsurv + 1;
        if (hipMallocAsync((void **)&ws.A, sizeof(double) * L * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.B, sizeof(double) * L * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.C, sizeof(double) * L * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.P, sizeof(double) * ns * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.Q, sizeof(double) * ns * nn, s) != hipSuccess)
            return VK_ENOMEM;
        return 0;
    }
    ~BcrTmp() {
        for (double *p : {ws.A, ws.B, ws.C, ws.P, ws.Q})
            if (p) (void)hipFreeAsync(p, s);
    }
};

extern "C" int vk_block_solve_system(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                                     const double *d_Lsub, const double *d_Usup, const double *d_b, double *d_x,
                                     int refine, int *d_info, void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_b || !d_x || refine < 0 || d_b == d_x) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)plan->ni * plan->ni, L = (size_t)cols->L;
    double *Dinv = nullptr, *work = nullptr, *tmp = nullptr;
    BcrTmp bt(s);
    int r = 0;
    if (hipMallocAsync((void **)&Dinv, sizeof(double) * L * nn, s) != hipSuccess ||
        hipMallocAsync((void **)&work, sizeof(double) * L * plan->ni, s) != hipSuccess ||
        hipMallocAsync((void **)&tmp, sizeof(double) * L * plan->ni, s) != hipSuccess)
        r = VK_ENOMEM;
    if (!r && plan->factor_variant == VK_FACTOR_BCR) r = bt.alloc(plan, cols);
    if (!r) r = vk_launch_factor(plan, cols, d_D, d_Lsub, d_Usup, Dinv, d_info, nullptr, &bt.ws, s);
    if (!r)
        r = vk_launch_solve_refined(plan, cols, d_D, Dinv, d_Lsub, d_Usup, d_b, d_x, work, tmp, refine, nullptr,
                                    &bt.ws, s);
    for (double *p : {Dinv, work, tmp})
        if (p) (void)hipFreeAsync(p, s);
    return r;
}

extern "C" int vk_block_factor(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                               const double *d_Lsub, const double *d_Usup, double *d_Dinv, int *d_info,
                               void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_Dinv || d_D == d_Dinv) return VK_EINVAL;
    if (plan->factor_variant == VK_FACTOR_BCR) return VK_EINVAL;
    return vk_launch_factor(plan, cols, d_D, d_Lsub, d_Usup, d_Dinv, d_info, nullptr, nullptr, (hipStream_t)stream);
}

extern "C" int vk_block_solve(const vk_plan *plan, const vk_columns *cols, const double *d_Dinv,
                              const double *d_Lsub, const double *d_Usup, double *d_rhs, void *stream) {
    if (!plan || !cols || !d_Dinv || !d_Lsub || !d_Usup || !d_rhs) return VK_EINVAL;
    if (plan->factor_variant == VK_FACTOR_BCR) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    double *tmp = nullptr;
    VK_CHECK(hipMallocAsync((void **)&tmp, sizeof(double) * (size_t)cols->L * plan->ni, s));
    int r = vk_launch_solve(plan, cols, d_Dinv, d_Lsub, d_Usup, d_rhs, d_rhs, tmp, nullptr, nullptr, s);
    VK_CHECK(hipFreeAsync(tmp, s));
    return r;
}

extern "C" int vk_block_residual(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                                 const double *d_Lsub, const double *d_Usup, const double *d_x, const double *d_b,
                                 double *d_r, void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_x || !d_b || !d_r) return VK_EINVAL;
    return vk_launch_residual(plan, cols, d_D, d_Lsub, d_Usup, d_x, d_b, d_r, nullptr, (hipStream_t)stream);
}

namespace {
// a well-conditioned block system of the columns' shape for timing
__global__ void k_fill_tune(int ni, long n_d, double *__restrict__ D, long n_c, double *__restrict__ Lsub,
                            double *__restrict__ Usup) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n_d) {
        const long e = t % ((long)ni * ni);
        const int i = (int)(e / ni), j = (int)(e % ni);
        D[t] = (i == j) ? 2.0 * ni : 1.0 / (1.0 + i + 2.0 * j);
    }
    if (t < n_c) {
        Lsub[t] = -0.25;
        Usup[t] = -0.5;
    }
}
}  // namespace

static bool variant_ok(const vk_plan *plan, int v) {
    return v >= 0 && v < VK_FACTOR_NVARIANTS && (v != VK_FACTOR_WAVE || plan->ni <= 112);
}

// cyclic reduction pays off while the columns leave most of the device idle
static constexpr int BCR_TUNE_MAX_COLS = 64;

extern "C" int vk_plan_set_factor_variant(vk_plan *plan, int variant) {
    if (!plan || !variant_ok(plan, variant)) return VK_EINVAL;
    plan->factor_variant = variant;
    return 0;
}

extern "C" int vk_plan_factor_variant(const vk_plan *plan) { return plan ? plan->factor_variant : VK_EINVAL; }

extern "C" int vk_tune_factor(vk_plan *plan, const vk_columns *cols, int reps, float *ms, void *stream) {
    if (!plan || !cols || reps < 1 || !ms) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    const size_t nn = (size_t)plan->ni * plan->ni, L = (size_t)cols->L;
    const bool try_bcr = cols->ncols <= BCR_TUNE_MAX_COLS;
    double *D = nullptr, *Dinv = nullptr, *Lsub = nullptr, *Usup = nullptr, *x = nullptr, *tmp = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    BcrTmp bt(s);
    int rc = 0;
    if (hipMallocAsync((void **)&D, sizeof(double) * L * nn, s) != hipSuccess ||
        hipMallocAsync((void **)&Dinv, sizeof(double) * L * nn, s) != hipSuccess ||
        hipMallocAsync((void **)&Lsub, sizeof(double) * L * plan->ni, s) != hipSuccess ||
        hipMallocAsync((void **)&Usup, sizeof(double) * L * plan->ni, s) != hipSuccess ||
        hipMallocAsync((void **)&x, sizeof(double) * L * plan->ni, s) != hipSuccess ||
        hipMallocAsync((void **)&tmp, sizeof(double) * L * plan->ni, s) != hipSuccess ||
        hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        (try_bcr && bt.alloc(plan, cols))) {
        rc = VK_ENOMEM;
    } else {
        const long nd = (long)(L * nn);
        hipLaunchKernelGGL(k_fill_tune, dim3((unsigned)((nd + 255) / 256)), dim3(256), 0, s, plan->ni, nd, D,
                           (long)(L * plan->ni), Lsub, Usup);
        const int keep = plan->factor_variant;
        int bestv = keep;
        float bestms = 0.f;
        bool have = false;
        // one repetition: a factorisation and the two solves of an attempt
        auto rep = [&]() -> int {
            int r = vk_launch_factor(plan, cols, D, Lsub, Usup, Dinv, nullptr, nullptr, &bt.ws, s);
            for (int q = 0; q < 2 && !r; q++) r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, Lsub, x, tmp, nullptr,
                                                                  &bt.ws, s);
            return r;
        };
        for (int v = 0; v < VK_FACTOR_NVARIANTS && rc == 0; v++) {
            ms[v] = -1.0f;
            if (!variant_ok(plan, v) || (v == VK_FACTOR_BCR && !try_bcr)) continue;
            plan->factor_variant = v;
            rc = rep();   // warm-up
            if (rc) break;
            hipError_t e = hipEventRecord(e0, s);
            if (e != hipSuccess) { rc = -(int)e; break; }
            for (int q = 0; q < reps && rc == 0; q++) rc = rep();
            if (rc) break;
            e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e != hipSuccess) { rc = -(int)e; break; }
            ms[v] = t / reps;
            if (!have || ms[v] < bestms) { bestms = ms[v]; bestv = v; have = true; }
        }
        plan->factor_variant = rc ? keep : bestv;
    }
    for (double *p : {D, Dinv, Lsub, Usup, x, tmp})
        if (p) (void)hipFreeAsync(p, s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return rc;
}

namespace {
__global__ void k_selftest_wave_max(const unsigned long long *in, unsigned long long *out) {
    out[threadIdx.x] = wave_max_u64(in[threadIdx.x]);
}
}  // namespace

extern "C" int vk_selftest_wave_max_u64(const unsigned long long *d_in, unsigned long long *d_out, void *stream) {
    if (!d_in || !d_out) return VK_EINVAL;
    hipLaunchKernelGGL(k_selftest_wave_max, dim3(1), dim3(64), 0, (hipStream_t)stream, d_in, d_out);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_selftest_mfma_f64(double *d_out, void *stream) {
    if (!d_out) return VK_EINVAL;
    hipLaunchKernelGGL(k_selftest_mfma_f64, dim3(1), dim3(64), 0, (hipStream_t)stream, d_out);
    VK_CHECK(hipGetLastError());
    return 0;
}
