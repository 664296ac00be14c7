// Device building blocks shared by the block-inversion kernels: the blocked
// Gauss-Jordan inversion on FP64 MFMA tiles (chain factorisation and cyclic
// reduction) and the slab products of the cyclic-reduction updates.
#pragma once
#include "vk_internal.h"

namespace {

constexpr int NT = 256;

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion on FP64 MFMA (factor variant 1)
//
// The pivot steps are grouped into panels of 16 columns.  The 16 pivot
// steps of a panel run on the panel columns alone, spread over all four
// waves (a thread per (column, row group)), with two barriers per step:
// one to combine the per-wave pivot keys (|value| bits with the row in the
// low byte), one to publish the pivot row.  The other columns then receive
// the whole panel at once, as the rank-16 Schur update
//     A[i][j] = (i in P ? 0 : A[i][j]) + sum_s Pan[i][s] A_old[p_s][j]
// (P: the panel's pivot rows p_s, Pan: the factored panel columns), which is
// a 16-deep GEMM and runs on v_mfma_f64_16x16x4f64.  The matrix lives in the MFMA
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
    double prow2[2][16];          // pivot rows of the panel factorisation, by step parity
    double colk[2][N16];          // column kk of the panel before step kk, by step parity
    unsigned long long wkey[2][4];   // per-wave pivot keys, by step parity
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

// Pivot steps p0 .. p0+kend-1 on the panel pan[k][row] with all four waves:
// thread t holds column c = t & 15 of rows g + 16m (g = t >> 4, m < NB).
// Per step: the holders of column kk form their pivot keys (|value| bits
// with the row in the low byte, so ties go to the lowest row) and publish
// the column; one barrier; every thread reduces the four wave keys, the
// holders of the pivot row publish it; a second barrier; every thread
// updates its NB entries.  Same operations, in the same order per entry,
// as the one-wave panel.
template <int NB>
__device__ __forceinline__ void blk_panel4(BlkShared<NB> &sh, double (*pan)[16 * NB + 1], int ni, int p0, int kend,
                                           int tag) {
    constexpr int N16 = 16 * NB;
    const int tid = threadIdx.x, c = tid & 15, g = tid >> 4, w = tid >> 6, l = tid & 63;
    double v[NB];
    bool av[NB];
#pragma unroll
    for (int m = 0; m < NB; m++) {
        const int r = g + 16 * m;
        v[m] = pan[c][r];
        av[m] = (r < ni) && sh.ipos[r] < 0;
    }
    bool zero = false;
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
        if (kk >= kend) continue;   // uniform; keeps the trip count static
        const int par = kk & 1;
        unsigned long long key = 0;
        if (c == kk) {
#pragma unroll
            for (int m = 0; m < NB; m++) {
                const int r = g + 16 * m;
                const unsigned long long k1 =
                    ((unsigned long long)__double_as_longlong(fabs(v[m])) & ~0xffull) | (unsigned)(255 - r);
                if (av[m]) key = k1 > key ? k1 : key;
                sh.colk[par][r] = v[m];
            }
        }
        // the wave's maximum: column kk sits in lanes kk, 16+kk, 32+kk, 48+kk
        unsigned long long wk = readlane_u64(key, kk), u;
        u = readlane_u64(key, 16 + kk);
        wk = u > wk ? u : wk;
        u = readlane_u64(key, 32 + kk);
        wk = u > wk ? u : wk;
        u = readlane_u64(key, 48 + kk);
        wk = u > wk ? u : wk;
        if (l == 0) sh.wkey[par][w] = wk;
        __syncthreads();
        unsigned long long best = sh.wkey[par][0];
#pragma unroll
        for (int q = 1; q < 4; q++) {
            const unsigned long long b2 = sh.wkey[par][q];
            best = b2 > best ? b2 : best;
        }
        const unsigned blo = (unsigned)(best & 0xffffffffull), bhi = (unsigned)(best >> 32);
        const int p = (blo | bhi) ? 255 - (int)(blo & 0xffu) : 0;
        zero = zero || (((blo >> 8) | bhi) == 0u);
        const int pg = p & 15, pm = p >> 4;
        double f[NB];
#pragma unroll
        for (int m = 0; m < NB; m++) f[m] = sh.colk[par][g + 16 * m];
        if (g == pg) {
            double pv = v[0];
#pragma unroll
            for (int m = 1; m < NB; m++) pv = (m == pm) ? v[m] : pv;
            sh.prow2[par][c] = pv;
        }
        __syncthreads();
        const double piv = sh.prow2[par][kk];
        const double inv = 1.0 / piv;
        const double rsc = sh.prow2[par][c] * inv;
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const double nv = (c == kk) ? -f[m] * inv : fma(-f[m], rsc, v[m]);
            const bool isp = (g == pg) && (m == pm);
            v[m] = isp ? ((c == kk) ? inv : rsc) : nv;
            if (isp) av[m] = false;
        }
        if (tid == 0) {
            sh.ipos[p] = p0 + kk;
            sh.perm[p0 + kk] = p;
        }
    }
    if (zero && tid == 0 && sh.info == 0) sh.info = tag;
#pragma unroll
    for (int m = 0; m < NB; m++) pan[c][g + 16 * m] = v[m];
    for (int s2 = kend; s2 < 16; s2++)
        for (int cc = tid; cc < N16; cc += NT) sh.rb[s2][cc] = 0.0;
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
        blk_panel4<NB>(sh, pan, ni, p0, kend, tag);
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

// Products for the cyclic-reduction updates, one 32-row output slab per
// workgroup: acc += op(A)[r0:r0+32, :] op(B) over k < ni with
//     A(a, c) = Ab[a*sar + c*sac],  B(c, b) = Bb[c*sbr + b*sbc].
// Output tiles are 16 x 16 (MFMA C layout: lane l, register r holds row
// 4r + l/16, column l%16 of the tile); wave w owns row block w & 1 of the
// slab and column blocks (w >> 1) + 2t.  K runs in chunks of SLAB_KC staged
// through LDS; the next chunk's global loads are issued into registers
// before the current chunk's MFMAs, and each chunk is read with the
// operand's unit-stride index fastest across threads (coalesced whichever
// way the operand is stored).
constexpr int SLAB_KC = 32;

template <int NB>
struct SlabShared {
    double sa[SLAB_KC][33];
    double sb[SLAB_KC][16 * NB + 1];
};

template <int NB>
struct SlabRegs {
    static constexpr int PA = SLAB_KC * 32 / NT, PB = SLAB_KC * 16 * NB / NT;
    double a[PA], b[PB];
};

// element e of a chunk of KC x n: (k, r), the unit-stride index fastest
__device__ __forceinline__ void chunk_kr(int e, int n, bool rfast, int &k, int &r) {
    const int k1 = e / n, r1 = e - k1 * n;
    const int r2 = e / SLAB_KC, k2 = e - r2 * SLAB_KC;
    k = rfast ? k1 : k2;
    r = rfast ? r1 : r2;
}

template <int NB>
__device__ __forceinline__ void slab_load(SlabRegs<NB> &R, int ni, int r0, int k0, const double *__restrict__ Ab,
                                          long sar, long sac, const double *__restrict__ Bb, long sbr, long sbc) {
    constexpr int N16 = 16 * NB;
    const bool afast = sar == 1, bfast = sbc == 1;
#pragma unroll
    for (int q = 0; q < SlabRegs<NB>::PA; q++) {
        int k, r;
        chunk_kr(threadIdx.x + NT * q, 32, afast, k, r);
        const int a = r0 + r, c = k0 + k;
        R.a[q] = (a < ni && c < ni) ? Ab[(long)a * sar + (long)c * sac] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < SlabRegs<NB>::PB; q++) {
        int k, r;
        chunk_kr(threadIdx.x + NT * q, N16, bfast, k, r);
        const int c = k0 + k;
        R.b[q] = (r < ni && c < ni) ? Bb[(long)c * sbr + (long)r * sbc] : 0.0;
    }
}

template <int NB>
__device__ __forceinline__ void slab_stage(const SlabRegs<NB> &R, SlabShared<NB> &g, long sar, long sbc) {
    const bool afast = sar == 1, bfast = sbc == 1;
#pragma unroll
    for (int q = 0; q < SlabRegs<NB>::PA; q++) {
        int k, r;
        chunk_kr(threadIdx.x + NT * q, 32, afast, k, r);
        g.sa[k][r] = R.a[q];
    }
#pragma unroll
    for (int q = 0; q < SlabRegs<NB>::PB; q++) {
        int k, r;
        chunk_kr(threadIdx.x + NT * q, 16 * NB, bfast, k, r);
        g.sb[k][r] = R.b[q];
    }
}

template <int NB>
__device__ __forceinline__ void slab_gemm(f64x4 (&acc)[(NB + 1) / 2], int ni, int r0, const double *__restrict__ Ab,
                                          long sar, long sac, const double *__restrict__ Bb, long sbr, long sbc,
                                          SlabShared<NB> &g) {
    constexpr int TB = (NB + 1) / 2;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int bi = w & 1, wc = w >> 1, lc = l & 15, lk = l >> 4;
    const int nch = (ni + SLAB_KC - 1) / SLAB_KC;
    SlabRegs<NB> R;
    slab_load<NB>(R, ni, r0, 0, Ab, sar, sac, Bb, sbr, sbc);
    for (int ch = 0; ch < nch; ch++) {
        __syncthreads();   // the previous chunk's reads of g are done
        slab_stage<NB>(R, g, sar, sbc);
        __syncthreads();
        if (ch + 1 < nch) slab_load<NB>(R, ni, r0, (ch + 1) * SLAB_KC, Ab, sar, sac, Bb, sbr, sbc);
#pragma unroll
        for (int k4 = 0; k4 < SLAB_KC / 4; k4++) {
            const double a = g.sa[4 * k4 + lk][16 * bi + lc];
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                if (bj >= NB) continue;
                acc[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, g.sb[4 * k4 + lk][16 * bj + lc], acc[tj], 0, 0, 0);
            }
        }
    }
    __syncthreads();
}

template <int NB>
__device__ __forceinline__ void slab_zero(f64x4 (&acc)[(NB + 1) / 2]) {
#pragma unroll
    for (int tj = 0; tj < (NB + 1) / 2; tj++) acc[tj] = f64x4{0.0, 0.0, 0.0, 0.0};
}

// out(i, c) = sgn * acc (+ add(i, c) if add) for the slab rows i, ni x ni,
// stored row-major (out[i*ni + c]) or transposed (out[c*ni + i])
template <int NB>
__device__ __forceinline__ void slab_store(const f64x4 (&acc)[(NB + 1) / 2], int ni, int r0, double *__restrict__ out,
                                           bool transposed, double sgn, const double *add) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int bi = w & 1, wc = w >> 1, lg = l >> 4, lc = l & 15;
#pragma unroll
    for (int tj = 0; tj < (NB + 1) / 2; tj++) {
        const int bj = wc + 2 * tj, cc = 16 * bj + lc;
        if (bj >= NB || cc >= ni) continue;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = r0 + 16 * bi + 4 * r + lg;
            if (i < ni) {
                const size_t e = transposed ? (size_t)cc * ni + i : (size_t)i * ni + cc;
                out[e] = add ? add[e] + sgn * acc[tj][r] : sgn * acc[tj][r];
            }
        }
    }
}

}  // namespace
