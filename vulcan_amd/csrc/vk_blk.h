// Device building blocks shared by the block-inversion kernels: the blocked
// Gauss-Jordan inversion on FP64 MFMA tiles (chain factorisation and cyclic
// reduction) and the slab products of the cyclic-reduction updates.
#pragma once
#include "vk_internal.h"

namespace {

constexpr int NT = 256;

// Threads of a blocked-inversion workgroup: 4 waves up to 112 rows, 8 waves
// above (the accumulator tiles of 128-160 rows would not fit 4 waves'
// registers), with the waves in a 2 x (waves / 2) grid over the tiles.
constexpr int blk_nt(int nb) { return nb >= 7 ? 512 : 256; }
template <int NB>
struct BlkGrid {
    static constexpr int NTH = blk_nt(NB), NW = NTH / 64, WC = NW / 2;
    static constexpr int TBR = (NB + 1) / 2, TBC = (NB + WC - 1) / WC;   // tiles per wave: rows, columns
};

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion on FP64 MFMA (factor variant 1)
//
// The pivot steps are grouped into panels of 16 columns.  One wave runs the
// 16 pivot steps of a panel on the panel columns alone (rows spread over the
// lanes, pivot search by a wave-wide max over 32-bit (|value|, row) keys,
// pivot row handed over through LDS: no workgroup barrier inside a panel;
// the panel wave rotates with the panel and the workgroup index, so that
// co-resident workgroups do not run their panels on the same SIMD; a version
// spreading the panel over four waves, two barriers per step, measured
// slower).  The other columns then receive the whole panel at once, as the
// rank-16 Schur update
//     A[i][j] = (i in P ? 0 : A[i][j]) + sum_s Pan[i][s] A_old[p_s][j]
// (P: the panel's pivot rows p_s, Pan: the factored panel columns), which is
// a 16-deep GEMM and runs on v_mfma_f64_16x16x4f64.  Three barriers per
// panel.  The matrix lives in the MFMA
// accumulator layout: 16 x 16 tiles, wave (wr, wc) of the 2 x WC wave grid
// (WC = 2 or 4, BlkGrid) owns row blocks wr, wr+2, ... and column blocks
// wc, wc+WC, ...; lane l,
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
    // above 128 rows the panel wave parks its accumulator tiles here during
    // the panel (its registers then hold the panel alone)
    static constexpr bool PARK = NB >= 9;
    double pan[2][16][N16 + 1];   // panel columns, [k][row], double-buffered by panel parity
    double rb[16][N16];           // old pivot rows of the panel, [s][col]
    double dump[blk_nt(NB)];      // per-thread target of discarded writes
    double prow[16];              // current pivot row of the panel factorisation
    int ipos[N16];
    int perm[N16];
    int info;
    double park[PARK ? 4 * ((NB + 1) / 2) * ((NB + 3) / 4) : 1][64];
};

template <int CTRL>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v & 0xffffffffull), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(v >> 32), CTRL, 0xf, 0xf, false);
    return ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo;
}

// Maximum over the 64 lanes (all active), wave-uniform: DPP max within rows
// of 16 (quad swaps, half-row and row mirrors), then row_bcast:15 and
// row_bcast:31 carry the row maxima up to lane 63, read once.
__device__ __forceinline__ unsigned wave_max_u32(unsigned v) {
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0xB1, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x4E, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x141, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x140, 0xf, 0xf, false));
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x142, 0xa, 0xf, false));   // rows 1, 3 += row 0, 2
    v = max(v, (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, 0x143, 0xc, 0xf, false));   // rows 2, 3 += row 1
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// 1 / d: the hardware reciprocal estimate refined by two Newton-Raphson
// steps (e = 1 - d r, r = r + r e), within about an ulp; the IEEE division
// sequence (scale, estimate, refinement, fix-up) is twice as long on the
// pivot step's path.
__device__ __forceinline__ double blk_recip(double d) {
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    return fma(r, e, r);
}

// Pivot steps p0 .. p0+kend-1 on the panel pan[k][row] (one wave).  The
// pivot key is the upper word of |value| with the row in its low byte: the
// largest magnitude to 12 mantissa bits, ties to the lowest row.  The pivot
// bookkeeping (ipos of the chosen rows, perm of the steps) stays in
// registers until the panel ends: no LDS write waits inside a step.
//
// A step subtracts (v[kk] / pivot) x the pivot row from every other row.
// The update is linear in the row, so a chosen pivot row is kept unscaled
// (its own multiplier is 0 and its pivot-column entry 1): later steps update
// it like any row, and the panel's end multiplies it by its 1 / pivot,
// giving the Gauss-Jordan values (the scaled row, 1 / pivot in the pivot
// column) without scaling the row inside the step.
template <int NB>
__device__ __forceinline__ void blk_panel(BlkShared<NB> &sh, double (*pan)[16 * NB + 1], int ni, int p0, int kend,
                                          int tag) {
    constexpr int N16 = 16 * NB;
    constexpr int RPL = (N16 + 63) / 64;
    const int l = threadIdx.x & 63;
    double v[RPL][16];
    // this lane's row q: -1 still available, -2 not a candidate (padding or
    // chosen in an earlier panel), >= 0 the pivot step that chose it
    int code[RPL];
    double pinv[RPL];   // 1 / pivot of a row chosen in this panel
    bool zero = false;
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const int r = l + 64 * q;
#pragma unroll
        for (int k = 0; k < 16; k++) v[q][k] = (r < N16) ? pan[k][r] : 0.0;
        code[q] = (r < ni) && sh.ipos[r] < 0 ? -1 : -2;
        pinv[q] = 1.0;
    }
#pragma unroll
    for (int kk = 0; kk < 16; kk++) {
        if (kk >= kend) continue;   // uniform; keeps the trip count static
        unsigned best = 0;
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            const int r = l + 64 * q;
            const unsigned key =
                ((unsigned)(__double_as_longlong(fabs(v[q][kk])) >> 32) & ~0xffu) | (unsigned)(255 - r);
            best = code[q] == -1 && key > best ? key : best;
        }
        best = wave_max_u32(best);
        const int p = best ? 255 - (int)(best & 0xffu) : 0;
        zero = zero || (best >> 8) == 0u;   // zero (or no) pivot
        const int pl = p & 63, qp = p >> 6;
        // the owner lane hands the pivot row to the wave through LDS (in order within the wave)
#pragma unroll
        for (int q = 0; q < RPL; q++)
            if (q == qp && l == pl) {
#pragma unroll
                for (int k = 0; k < 16; k++) sh.prow[k] = v[q][k];
            }
        // lanes exchange data here: without the barrier the compiler may let the
        // other lanes read before the owner's (exec-masked) writes
        __builtin_amdgcn_wave_barrier();
        double rs[16];
#pragma unroll
        for (int k = 0; k < 16; k++) rs[k] = sh.prow[k];
        const double inv = blk_recip(rs[kk]);
#pragma unroll
        for (int q = 0; q < RPL; q++) {
            const bool own = q == qp && l == pl;
            const double f = own ? 0.0 : v[q][kk] * inv;
#pragma unroll
            for (int k = 0; k < 16; k++) v[q][k] = fma(-f, rs[k], v[q][k]);
            v[q][kk] = own ? 1.0 : -f;
            code[q] = own ? p0 + kk : code[q];
            pinv[q] = own ? inv : pinv[q];
        }
        if (l == 0) sh.perm[p0 + kk] = p;
    }
    if (zero && l == 0 && sh.info == 0) sh.info = tag;
#pragma unroll
    for (int q = 0; q < RPL; q++) {
        const int r = l + 64 * q;
        if (code[q] >= 0) sh.ipos[r] = code[q];
        if (r < N16) {
            // rows chosen in this panel take their 1 / pivot (the others 1.0, exact)
#pragma unroll
            for (int k = 0; k < 16; k++) pan[k][r] = v[q][k] * pinv[q];
        }
    }
    for (int s = kend; s < 16; s++)
        for (int c = l; c < N16; c += 64) sh.rb[s][c] = 0.0;
}


// Inverts the matrix held in acc (MFMA tile layout) and writes the inverse
// transposed to X (X[m][i] = inv[i][m]).  A zero pivot records `tag` in
// sh.info (sticky).
template <int NB>
__device__ __forceinline__ void blk_invert(f64x4 (&acc)[BlkGrid<NB>::TBR][BlkGrid<NB>::TBC], int ni,
                                           double *__restrict__ X, BlkShared<NB> &sh, int tag) {
    using G = BlkGrid<NB>;
    constexpr int N16 = 16 * NB, TBR = G::TBR, TBC = G::TBC, WC = G::WC;
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const int wr = w / WC, wc = w % WC;
    const int lg = l >> 4, lc = l & 15;
    for (int t = tid; t < N16; t += G::NTH) sh.ipos[t] = -1;
    __syncthreads();
    const int npan = (ni + 15) >> 4;
    for (int pb = 0; pb < npan; pb++) {
        double(*pan)[N16 + 1] = sh.pan[pb & 1];
        const int p0 = 16 * pb, kend = min(16, ni - p0);
        // the owners of column block pb publish the panel
#pragma unroll
        for (int tj = 0; tj < TBC; tj++) {
            if (wc + WC * tj != pb) continue;   // static register index, uniform test
#pragma unroll
            for (int ti = 0; ti < TBR; ti++) {
                const int bi = wr + 2 * ti;
                if (bi < NB) {
#pragma unroll
                    for (int r = 0; r < 4; r++) pan[lc][16 * bi + 4 * r + lg] = acc[ti][tj][r];
                }
            }
        }
        __syncthreads();
        if (w == ((pb + (int)blockIdx.x) & (G::NW - 1))) {
            if constexpr (BlkShared<NB>::PARK) {
#pragma unroll
                for (int ti = 0; ti < TBR; ti++)
#pragma unroll
                    for (int tj = 0; tj < TBC; tj++)
#pragma unroll
                        for (int r = 0; r < 4; r++) sh.park[(ti * TBC + tj) * 4 + r][l] = acc[ti][tj][r];
            }
            blk_panel<NB>(sh, pan, ni, p0, kend, tag);
            if constexpr (BlkShared<NB>::PARK) {
#pragma unroll
                for (int ti = 0; ti < TBR; ti++)
#pragma unroll
                    for (int tj = 0; tj < TBC; tj++)
#pragma unroll
                        for (int r = 0; r < 4; r++) acc[ti][tj][r] = sh.park[(ti * TBC + tj) * 4 + r][l];
            }
        }
        __syncthreads();
        // panel tiles take the factored panel; the other tiles hand their
        // pivot-row entries to rb and clear them
#pragma unroll
        for (int ti = 0; ti < TBR; ti++) {
            const int bi = wr + 2 * ti;
            if (bi >= NB) continue;
            int sr[4];
#pragma unroll
            for (int r = 0; r < 4; r++) sr[r] = sh.ipos[16 * bi + 4 * r + lg] - p0;
#pragma unroll
            for (int tj = 0; tj < TBC; tj++) {
                const int bj = wc + WC * tj;
                if (bj >= NB) continue;
                if (bj == pb) {
#pragma unroll
                    for (int r = 0; r < 4; r++) acc[ti][tj][r] = pan[lc][16 * bi + 4 * r + lg];
                } else {
                    // branch-free: entries of other rows go to the thread's own dump slot
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const bool piv = (unsigned)sr[r] < 16u;
                        double *dst = piv ? &sh.rb[sr[r]][16 * bj + lc] : &sh.dump[tid];
                        *dst = acc[ti][tj][r];
                        acc[ti][tj][r] = piv ? 0.0 : acc[ti][tj][r];
                    }
                }
            }
        }
        __syncthreads();
        // rank-16 update on MFMA
#pragma unroll
        for (int ti = 0; ti < TBR; ti++) {
            const int bi = wr + 2 * ti;
            if (bi >= NB) continue;
            double a[4];
#pragma unroll
            for (int s4 = 0; s4 < 4; s4++) a[s4] = pan[4 * s4 + (l >> 4)][16 * bi + lc];
#pragma unroll
            for (int tj = 0; tj < TBC; tj++) {
                const int bj = wc + WC * tj;
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
    for (int ti = 0; ti < TBR; ti++) {
        const int bi = wr + 2 * ti;
        if (bi >= NB) continue;
        int qi[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 16 * bi + 4 * r + lg;
            qi[r] = i < ni ? sh.ipos[i] : -1;
        }
#pragma unroll
        for (int tj = 0; tj < TBC; tj++) {
            const int bj = wc + WC * tj;
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
//
// Branch-free: every thread loads all its entries unconditionally (padding
// rows and columns read element 0 and clear it with a bit mask, giving the
// +0.0 they hold), so a row block's loads are all in flight before the first
// wait.  With a conditional load per entry (or a select on the loaded
// value, which the compiler turns back into a branch) it waited on each
// entry's loads in turn.  Xn and Xn2 are uniform.
__device__ __forceinline__ double blk_masked(double v, unsigned long long m) {
    return __longlong_as_double(__double_as_longlong(v) & (long long)m);
}

template <int NB>
__device__ __forceinline__ void blk_load_schur(f64x4 (&acc)[BlkGrid<NB>::TBR][BlkGrid<NB>::TBC], int ni,
                                               const double *__restrict__ Dl, const double *cl, const double *Xn,
                                               const double *cn, const double *cl2, const double *Xn2,
                                               const double *cn2) {
    using G = BlkGrid<NB>;
    constexpr int TBR = G::TBR, TBC = G::TBC, WC = G::WC;
    int tid = threadIdx.x;
    // opaque to the optimiser: the indices below are formed here, at each
    // call, rather than hoisted out of the caller's loop and kept live in
    // registers through the inversion
    asm volatile("" : "+v"(tid));
    const int w = tid >> 6, l = tid & 63;
    const int wr = w / WC, wc = w % WC;
    const int lg = l >> 4, lc = l & 15;
    // column cc = 16 bj + lc of each column tile (index 0 and mask 0 outside the block)
    int cj[TBC];
    unsigned long long cm[TBC];
#pragma unroll
    for (int tj = 0; tj < TBC; tj++) {
        const int bj = wc + WC * tj, cc = 16 * bj + lc;
        const bool ok = bj < NB && cc < ni;
        cj[tj] = ok ? cc : 0;
        cm[tj] = ok ? ~0ull : 0ull;
    }
    // row i = 16 bi + 4 r + lg of each entry (index 0 and mask 0 outside the block)
    int ir[TBR][4];
    unsigned long long rm[TBR][4];
#pragma unroll
    for (int ti = 0; ti < TBR; ti++) {
        const int bi = wr + 2 * ti;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int i = 16 * bi + 4 * r + lg;
            const bool rok = bi < NB && i < ni;
            ir[ti][r] = rok ? i : 0;
            rm[ti][r] = rok ? ~0ull : 0ull;
        }
    }
    // all of D's loads first, then (uniform) the Schur terms' loads
#pragma unroll
    for (int ti = 0; ti < TBR; ti++)
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int tj = 0; tj < TBC; tj++)
                acc[ti][tj][r] = blk_masked(Dl[(size_t)ir[ti][r] * ni + cj[tj]], rm[ti][r] & cm[tj]);
    // the Schur terms in groups of GR row tiles (all at once up to 96 rows;
    // one tile at a time above, where their loads would not fit the registers)
    constexpr int GR = TBR * TBC <= 9 ? TBR : 1;
    auto sub = [&](const double *c_l, const double *X, const double *c_n) {
        double cnv[TBC];
#pragma unroll
        for (int tj = 0; tj < TBC; tj++) cnv[tj] = c_n[cj[tj]];
#pragma unroll
        for (int t0 = 0; t0 < TBR; t0 += GR) {
            double li[GR][4], x[GR][4][TBC];
#pragma unroll
            for (int g = 0; g < GR; g++)
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    li[g][r] = c_l[ir[t0 + g][r]];
#pragma unroll
                    for (int tj = 0; tj < TBC; tj++) x[g][r][tj] = X[(size_t)cj[tj] * ni + ir[t0 + g][r]];
                }
#pragma unroll
            for (int g = 0; g < GR; g++)
#pragma unroll
                for (int r = 0; r < 4; r++)
#pragma unroll
                    for (int tj = 0; tj < TBC; tj++)
                        acc[t0 + g][tj][r] = blk_masked(acc[t0 + g][tj][r] - li[g][r] * (x[g][r][tj] * cnv[tj]),
                                                        rm[t0 + g][r] & cm[tj]);
        }
    };
    if (Xn) sub(cl, Xn, cn);
    if (Xn2) sub(cl2, Xn2, cn2);
}

// Products for the cyclic-reduction updates, one 32-row output slab (and
// the 16 NB output columns from c0) per workgroup, two products at a time:
//     acc1 += A1[r0:r0+32, :] B1[:, c0:],   acc2 (or acc1 if SAME) += A2[r0:r0+32, :] B2[:, c0:]
// over k < ni, with every operand stored transposed:
//     A(a, c) = At[c*ni + a],  B(c, b) = Bt[b*ni + c].
// A null A2 skips the second product.  Output tiles are 16 x 16 (MFMA C
// layout: lane l, register r holds row 4r + l/16, column l%16 of the tile);
// wave w owns row block w & 1 of the slab and column blocks (w >> 1) + 2t.
// K runs in chunks of KC staged through LDS; the next chunk's global loads
// are issued into registers before the current chunk's MFMAs, and each
// chunk is read with the operand's unit-stride index fastest across threads
// (A: the row a, B: the inner index c).  Each chunk costs about one global
// load latency, so deep chunks suit launches whose workgroups are all
// resident at once, shallow ones keep more workgroups per CU resident.
// Two staged operand pairs take 16 KC (34 + 16 NB) bytes of LDS (NB: the
// slab's column blocks); SlabKC is indexed by the block's column blocks.
template <int NB>
struct SlabKC {
    static constexpr int shallow = NB <= 7 ? 24 : 16;   // <= 56 KiB
    static constexpr int deep = NB <= 6 ? 48 : 32;      // <= 98 KiB (gfx950: 160 KiB per workgroup)
};

template <int NB, int KC>
struct SlabShared {
    double sa[2][KC][33];
    double sb[2][KC][16 * NB + 1];
};

template <int NB, int KC>
struct SlabRegs {
    static_assert(KC % 8 == 0, "chunk depth: a multiple of 8");
    static constexpr int PA = KC * 32 / NT, PB = (KC * 16 * NB + NT - 1) / NT;
    double a[2][PA], b[2][PB];
};

template <int NB, int KC>
__device__ __forceinline__ void slab_load(SlabRegs<NB, KC> &R, int u, int ni, int r0, int k0,
                                          const double *__restrict__ At, const double *__restrict__ Bt, int c0) {
#pragma unroll
    for (int q = 0; q < SlabRegs<NB, KC>::PA; q++) {
        const int e = threadIdx.x + NT * q, k = e >> 5, r = e & 31;   // a = r0 + r fastest
        const int a = r0 + r, c = k0 + k;
        R.a[u][q] = (a < ni && c < ni) ? At[(size_t)c * ni + a] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < SlabRegs<NB, KC>::PB; q++) {
        const int e = threadIdx.x + NT * q, r = e / KC, k = e - r * KC;   // c = k0 + k fastest
        const int c = k0 + k, b = c0 + r;
        R.b[u][q] = (e < KC * 16 * NB && b < ni && c < ni) ? Bt[(size_t)b * ni + c] : 0.0;
    }
}

template <int NB, int KC>
__device__ __forceinline__ void slab_stage(const SlabRegs<NB, KC> &R, int u, SlabShared<NB, KC> &g) {
#pragma unroll
    for (int q = 0; q < SlabRegs<NB, KC>::PA; q++) {
        const int e = threadIdx.x + NT * q;
        g.sa[u][e >> 5][e & 31] = R.a[u][q];
    }
#pragma unroll
    for (int q = 0; q < SlabRegs<NB, KC>::PB; q++) {
        const int e = threadIdx.x + NT * q, r = e / KC;
        if (e < KC * 16 * NB) g.sb[u][e - r * KC][r] = R.b[u][q];
    }
}

template <int NB, bool SAME, int KC>
__device__ __forceinline__ void slab_gemm2(f64x4 (&acc1)[(NB + 1) / 2], f64x4 (&acc2)[(NB + 1) / 2], int ni, int r0,
                                           const double *__restrict__ A1, const double *__restrict__ B1,
                                           const double *__restrict__ A2, const double *__restrict__ B2,
                                           SlabShared<NB, KC> &g, int c0) {
    constexpr int TB = (NB + 1) / 2;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int bi = w & 1, wc = w >> 1, lc = l & 15, lk = l >> 4;
    const int nch = (ni + KC - 1) / KC;
    const bool two = A2 != nullptr;
    SlabRegs<NB, KC> R;
    slab_load<NB, KC>(R, 0, ni, r0, 0, A1, B1, c0);
    if (two) slab_load<NB, KC>(R, 1, ni, r0, 0, A2, B2, c0);
    for (int ch = 0; ch < nch; ch++) {
        __syncthreads();   // the previous chunk's reads of g are done
        slab_stage<NB, KC>(R, 0, g);
        if (two) slab_stage<NB, KC>(R, 1, g);
        __syncthreads();
        if (ch + 1 < nch) {
            slab_load<NB, KC>(R, 0, ni, r0, (ch + 1) * KC, A1, B1, c0);
            if (two) slab_load<NB, KC>(R, 1, ni, r0, (ch + 1) * KC, A2, B2, c0);
        }
#pragma unroll
        for (int k4 = 0; k4 < KC / 4; k4++) {
            const double a1 = g.sa[0][4 * k4 + lk][16 * bi + lc];
#pragma unroll
            for (int tj = 0; tj < TB; tj++) {
                const int bj = wc + 2 * tj;
                if (bj >= NB) continue;
                acc1[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, g.sb[0][4 * k4 + lk][16 * bj + lc], acc1[tj], 0, 0,
                                                                 0);
            }
            if (two) {
                const double a2 = g.sa[1][4 * k4 + lk][16 * bi + lc];
#pragma unroll
                for (int tj = 0; tj < TB; tj++) {
                    const int bj = wc + 2 * tj;
                    if (bj >= NB) continue;
                    if (SAME)
                        acc1[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, g.sb[1][4 * k4 + lk][16 * bj + lc],
                                                                         acc1[tj], 0, 0, 0);
                    else
                        acc2[tj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a2, g.sb[1][4 * k4 + lk][16 * bj + lc],
                                                                         acc2[tj], 0, 0, 0);
                }
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

// out(i, c) = sgn * acc (+ add(i, c) if add) for the slab rows i and the
// columns c0 .. c0 + 16 NB - 1, ni x ni, stored row-major (out[i*ni + c])
// or transposed (out[c*ni + i])
template <int NB>
__device__ __forceinline__ void slab_store(const f64x4 (&acc)[(NB + 1) / 2], int ni, int r0, double *__restrict__ out,
                                           bool transposed, double sgn, const double *add, int c0) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int bi = w & 1, wc = w >> 1, lg = l >> 4, lc = l & 15;
#pragma unroll
    for (int tj = 0; tj < (NB + 1) / 2; tj++) {
        const int bj = wc + 2 * tj, cc = c0 + 16 * bj + lc;
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
