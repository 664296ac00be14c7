// Block-Thomas factorisation and solves for the Ros2 left-hand side
//
//     lhs = blocktridiag(diag(Lsub_j), D_j, diag(Usup_j))
//
// The off-diagonal blocks are diagonal (diffusion couples a species only to
// itself in the neighbouring layers, op.py:2012-2040), so the Schur update
//     S_j = D_j - diag(Lsub_j) S_{j-1}^-1 diag(Usup_{j-1})
// is elementwise and the dense work is the inversion of each S_j (kernels
// in vk_factor_reg.hip, vk_factor_blk.hip, vk_factor_wave.hip; the block
// cyclic reduction alternative in vk_bcr.hip).  This file holds the solves,
// the residual of iterative refinement and the dispatch over variants.
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
#include <atomic>

#include "vk_blk.h"

namespace {

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

__global__ void k_zero_int(int n, int *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// Twisted block solves, 1024-thread workgroups.  Thread (i, h) = (tid % RP,
// tid / RP) owns species row i and the h-th of SP column chunks of each
// inverse (SolveShape below), so one layer's matrix-vector product is
// SP-way split, and the Dinv chunks of the next layers (SolveDepth) are in
// flight while the current layer is processed (the chain over layers is
// latency bound, not bandwidth bound).
// Both sweeps stay within the column: layer indices run over [0, nz).
// ---------------------------------------------------------------------------
constexpr int NTS = 1024;

// rows per part (64, 128, or the padded block size 16 NB above 128
// species), parts (16 or 8; 7 or 6 for NB 9 or 10: threads from SP * RP on
// hold no part) and column chunk per thread (CH <= 27 doubles, so that one
// layer's chunk stays in registers at 1024 threads; CH <= 16 (NB - 1) + 1,
// the fewest species of NB blocks, so a chunk always fits in the block)
template <int NB>
struct SolveShape {
    static constexpr int RP = 16 * NB <= 64 ? 64 : (16 * NB <= 128 ? 128 : 16 * NB);
    static constexpr int SP = NTS / RP;
    static constexpr int CH = (16 * NB + SP - 1) / SP;
    static_assert(SP * RP <= NTS && SP * CH >= 16 * NB && CH <= 16 * (NB - 1) + 1, "solve shape");
};

// Where a thread's chunk of a layer's inverse lies: rows mb .. mb+CH-1 of
// the transposed inverse (column ib).  A part's chunk starts at m0 = h CH;
// the last part's is moved back to end at row ni-1 (mb = ni - CH) and its
// first d = m0 - mb entries, which belong to the part before, are zeroed;
// rows i >= ni read column ni-1 (their sums are never used) and threads
// without a part read part 0's chunk.  Every load stays inside the block.
struct ChunkMap {
    int voff;   // byte offset of (mb, ib) in the block
    int d;      // leading entries to zero
    int mb;     // first row of the chunk
};

template <int SP, int CH>
__device__ __forceinline__ ChunkMap chunk_map(int ni, int i, int h) {
    const int m0 = (h < SP ? h : 0) * CH;
    const int mb = min(m0, ni - CH);
    return {(mb * ni + min(i, ni - 1)) * 8, m0 - mb, mb};
}

// r[q] = Xt[mb + q][ib] (zero for q < d): raw buffer loads with the
// thread's offset in one register and the uniform row stride q ni in the
// scalar offset, over the block's byte range
template <int CH>
__device__ __forceinline__ void load_chunk(double (&r)[CH], const double *__restrict__ Xt, int ni,
                                           const ChunkMap &cm) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(Xt), (short)0, ni * ni * 8, 0x00020000);
#pragma unroll
    for (int q = 0; q < CH; q++) {
        const double x = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, cm.voff, q * ni * 8, 0));
        r[q] = q >= cm.d ? x : 0.0;
    }
}

// layers in flight per thread: fewer for larger blocks (128 registers per thread at 1024 threads)
template <int NB>
struct SolveDepth {
    static constexpr int value = NB <= 4 ? 4 : (NB <= 5 ? 3 : (NB <= 8 ? 2 : 1));
};

template <int DEP, int CH>
__device__ __forceinline__ void rotate(double (&b)[DEP][CH]) {
#pragma unroll
    for (int d = 0; d + 1 < DEP; d++) {
#pragma unroll
        for (int q = 0; q < CH; q++) b[d][q] = b[d + 1][q];
    }
}

// y_i = sum_m Xt[m][i] v[m] from the registers of the row parts; the result
// is valid in threads h == 0 (i < ni).  One barrier (part is double-buffered
// by the callers).
template <int SP, int RP, int CH>
__device__ __forceinline__ double solve_matvec(const double (&cur)[CH], const double *v, double (*part)[RP], int i,
                                               int h, int m0) {
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < CH; q++) acc = fma(cur[q], v[m0 + q], acc);
    if (h < SP) part[h][i] = acc;
    __syncthreads();
    double y = part[0][i];
#pragma unroll
    for (int hh = 1; hh < SP; hh++) y += part[hh][i];
    return y;
}

template <int NB>
__global__ void __launch_bounds__(NTS) k_solve_fwd2(int ni, const int *__restrict__ off, const int *__restrict__ order,
                                                     const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                     const double *__restrict__ Usup, const double *in, double *tmp,
                                                     const vk_colstate *__restrict__ st) {
    constexpr int RP = SolveShape<NB>::RP, SP = SolveShape<NB>::SP, CH = SolveShape<NB>::CH;
    const int chain = order ? order[blockIdx.x] : (int)blockIdx.x;   // 2c + side
    const int c = chain >> 1, side = chain & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x, i = tid % RP, h = tid / RP;
    const ChunkMap cm = chunk_map<SP, CH>(ni, i, h);
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    // v and part double-buffered by step parity: two barriers per step; the
    // right-hand side and coupling of the next step are loaded a step ahead
    __shared__ double v[2][SP * CH];
    __shared__ double part[2][SP][RP];
    for (int q = tid; q < 2 * SP * CH; q += NTS) v[q / (SP * CH)][q % (SP * CH)] = 0.0;
    const int j0 = side == 0 ? 0 : nz - 1, dj = side == 0 ? 1 : -1;
    const int nsteps = side == 0 ? m : nz - 1 - m;
    const double *cpl = side == 0 ? Lsub : Usup;
    constexpr int DEP = SolveDepth<NB>::value;
    double buf[DEP][CH];
#pragma unroll
    for (int d = 0; d < DEP; d++)
        if (d < nsteps) load_chunk<CH>(buf[d], Dinv + (size_t)(l0 + j0 + d * dj) * nn, ni, cm);
    const bool row = h == 0 && i < ni;
    double zi = 0.0, in_n = 0.0, cp_n = 0.0;
    if (row && nsteps > 0) {
        in_n = in[(size_t)(l0 + j0) * ni + i];
        cp_n = cpl[(size_t)(l0 + j0) * ni + i];
    }
    __syncthreads();
    for (int s = 0; s < nsteps; s++) {
        const size_t l = (size_t)(l0 + j0 + dj * s);
        const int pb = s & 1;
        const double in_s = in_n, cp_s = cp_n;
        if (row && s + 1 < nsteps) {
            in_n = in[(l + dj) * ni + i];
            cp_n = cpl[(l + dj) * ni + i];
        }
        if (row) {
            double w = in_s;
            if (s > 0) w = w - cp_s * zi;
            v[pb][i] = w;
        }
        __syncthreads();
        const double z = solve_matvec<SP, RP, CH>(buf[0], v[pb], part[pb], i, h, cm.mb);
        if (row) {
            tmp[l * ni + i] = z;
            zi = z;
        }
        rotate<DEP, CH>(buf);
        if (s + DEP < nsteps) load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)(l0 + j0 + dj * (s + DEP)) * nn, ni, cm);
    }
    // park b_m in the unused slot m so the backward kernel never reads `in`
    // (which may alias `out`)
    if (side == 0 && h == 0 && i < ni) {
        const size_t lm = (size_t)l0 + m;
        tmp[lm * ni + i] = in[lm * ni + i];
    }
}

template <int NB>
__global__ void __launch_bounds__(NTS) k_solve_bwd2(int ni, const int *__restrict__ off, const int *__restrict__ order,
                                                     const double *__restrict__ Dinv, const double *__restrict__ Lsub,
                                                     const double *__restrict__ Usup, const double *tmp, double *out,
                                                     const vk_colstate *__restrict__ st, double *acc) {
    constexpr int RP = SolveShape<NB>::RP, SP = SolveShape<NB>::SP, CH = SolveShape<NB>::CH;
    const int chain = order ? order[blockIdx.x] : (int)blockIdx.x;   // 2c + side
    const int c = chain >> 1, side = chain & 1;
    if (st && st[c].done) return;
    const int tid = threadIdx.x, i = tid % RP, h = tid / RP;
    const ChunkMap cm = chunk_map<SP, CH>(ni, i, h);
    const int l0 = off[c], nz = off[c + 1] - l0, m = nz / 2;
    const size_t nn = (size_t)ni * ni;
    __shared__ double v[2][SP * CH];
    __shared__ double part[2][SP][RP];
    for (int q = tid; q < 2 * SP * CH; q += NTS) v[q / (SP * CH)][q % (SP * CH)] = 0.0;
    const size_t lm = (size_t)l0 + m;
    const int dj = side == 0 ? -1 : 1;
    const int nsteps = 1 + (side == 0 ? m : nz - 1 - m);     // the middle layer, then the chain outwards
    const double *cpl = side == 0 ? Usup : Lsub;
    constexpr int DEP = SolveDepth<NB>::value;
    double buf[DEP][CH];
#pragma unroll
    for (int d = 0; d < DEP; d++)
        if (d < nsteps) load_chunk<CH>(buf[d], Dinv + (size_t)((long)lm + (long)d * dj) * nn, ni, cm);
    const bool row = h == 0 && i < ni;
    // x_m = M^-1 (b_m - Lsub_m z_{m-1} - Usup_m w_{m+1}); both sides compute it
    if (row) {
        double w = tmp[lm * ni + i];
        if (m > 0) w = w - Lsub[lm * ni + i] * tmp[(lm - 1) * ni + i];
        if (m < nz - 1) w = w - Usup[lm * ni + i] * tmp[(lm + 1) * ni + i];
        v[0][i] = w;
    }
    // coupling and forward-sweep value of the next layer, a step ahead
    double cp_n = 0.0, t_n = 0.0;
    if (row && nsteps > 1) {
        cp_n = cpl[(lm + dj) * ni + i];
        t_n = tmp[(lm + dj) * ni + i];
    }
    __syncthreads();
    double xi = solve_matvec<SP, RP, CH>(buf[0], v[0], part[0], i, h, cm.mb);
    if (side == 0 && row) {
        out[lm * ni + i] = xi;
        if (acc) acc[lm * ni + i] = acc[lm * ni + i] + xi;
    }
    rotate<DEP, CH>(buf);
    if (DEP < nsteps) load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)((long)lm + (long)DEP * dj) * nn, ni, cm);
    for (int s = 1; s < nsteps; s++) {
        const size_t l = lm + (long)dj * s;
        const int pb = s & 1;
        const double cp_s = cp_n, t_s = t_n;
        if (row && s + 1 < nsteps) {
            cp_n = cpl[(l + dj) * ni + i];
            t_n = tmp[(l + dj) * ni + i];
        }
        if (row) v[pb][i] = cp_s * xi;
        __syncthreads();
        const double y = solve_matvec<SP, RP, CH>(buf[0], v[pb], part[pb], i, h, cm.mb);
        if (row) {
            xi = t_s - y;
            out[l * ni + i] = xi;
            if (acc) acc[l * ni + i] = acc[l * ni + i] + xi;
        }
        rotate<DEP, CH>(buf);
        if (s + DEP < nsteps)
            load_chunk<CH>(buf[DEP - 1], Dinv + (size_t)((long)lm + (long)dj * (s + DEP)) * nn, ni, cm);
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

// r = b - A x, 32 rows of one layer per workgroup: 8 threads share a row,
// each a double-double partial over an eighth of the columns (a contiguous
// run, read straight into registers: every load of the run is issued
// before the first product), and the partials, b and the two coupling
// terms are combined in double-double.
constexpr int RES_ROWS = 32, RES_PARTS = NT / RES_ROWS;

__device__ inline void dd_add(double &hi, double &lo, double a_hi, double a_lo) {
    const double s = hi + a_hi;
    const double bb = s - hi;
    const double e = (hi - (s - bb)) + (a_hi - bb);     // hi + a_hi = s + e exactly
    hi = s;
    lo = lo + (e + a_lo);
}

// CH: the longest run, 16 NB / RES_PARTS for blocks of up to 16 NB rows
template <int CH>
__global__ void __launch_bounds__(NT) k_residual(int ni, const int *__restrict__ off,
                                                  const int *__restrict__ layer_col, const double *__restrict__ D,
                                                  const double *__restrict__ Lsub, const double *__restrict__ Usup,
                                                  const double *__restrict__ x, const double *__restrict__ b,
                                                  double *__restrict__ r, const vk_colstate *__restrict__ st) {
    extern __shared__ double rsm[];
    double *xs = rsm;                        // ni
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
    const int rr = tid / RES_PARTS, pp = tid % RES_PARTS;
    const int chunk = (ni + RES_PARTS - 1) / RES_PARTS;
    const int m0 = pp * chunk, m1 = min(ni, m0 + chunk);
    // this thread's run of its row (indices clamped into the row; entries
    // past the run are loaded but not used)
    double dr[CH];
    const double *row = Dl + (size_t)min(rr, nrows - 1) * ni;
#pragma unroll
    for (int k = 0; k < CH; k++) dr[k] = row[min(m0 + k, ni - 1)];
    for (int m = tid; m < ni; m += NT) xs[m] = x[(size_t)l * ni + m];
    __syncthreads();
    double hi = 0.0, lo = 0.0;
    if (rr < nrows) {
#pragma unroll
        for (int k = 0; k < CH; k++)
            if (m0 + k < m1) dd_sub_prod(hi, lo, dr[k], xs[m0 + k]);
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


}  // namespace


// info_zeroed: the caller has cleared info in an earlier launch on s
int vk_launch_factor(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                     const double *Usup, double *Dinv, int *info, const vk_colstate *st, const vk_bcr_ws *ws,
                     hipStream_t s, bool info_zeroed = false) {
    const int ni = plan->ni;
    const int n = (ni + 15) / 16;
    const int nc = cols->ncols;
    const int v = plan->factor_variant;
    if (n < 1 || n > 10) return VK_ETOOBIG;
    if (info && !info_zeroed) hipLaunchKernelGGL(k_zero_int, dim3((nc + 255) / 256), dim3(256), 0, s, nc, info);
    int r;
    switch (v) {
    case VK_FACTOR_BCR:
        if (!ws || !ws->A) return VK_EINVAL;
        return vk_launch_bcr_factor(n, ni, cols, D, Lsub, Usup, Dinv, info, st, ws, s);
    case VK_FACTOR_BLOCKED:
        r = vk_launch_factor_blocked1(n, ni, nc, cols->d_off, cols->d_chain_order[0], D, Lsub, Usup, Dinv, info, st, s);
        break;
    case VK_FACTOR_BLOCKED_OCC2:
        r = vk_launch_factor_blocked2(n, ni, nc, cols->d_off, cols->d_chain_order[1], D, Lsub, Usup, Dinv, info, st, s);
        break;
    case VK_FACTOR_WAVE:
        if (n <= 7) {
            r = vk_launch_factor_wave(n, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s);
            break;
        }
        // fall through: the two-wave kernel holds at most 112 species
    default:
        r = vk_launch_factor_register(n, ni, nc, cols->d_off, D, Lsub, Usup, Dinv, info, st, s);
    }
    if (r) return r;
    VK_CHECK(hipGetLastError());
    return 0;
}

// out = lhs^-1 in (tmp: L*ni scratch, may not alias in/out; in may alias out)
namespace {
// order: the chain each workgroup runs (the pairing of the two-per-CU
// factorisation: long chains share a CU with short ones)
template <int NB>
void launch_solve_n(int ni, int ncols, const int *off, const int *order, const double *Dinv, const double *Lsub,
                    const double *Usup, const double *in, double *out, double *tmp, const vk_colstate *st, double *acc,
                    hipStream_t s) {
    hipLaunchKernelGGL((k_solve_fwd2<NB>), dim3(2 * ncols), dim3(NTS), 0, s, ni, off, order, Dinv, Lsub, Usup, in, tmp,
                       st);
    hipLaunchKernelGGL((k_solve_bwd2<NB>), dim3(2 * ncols), dim3(NTS), 0, s, ni, off, order, Dinv, Lsub, Usup, tmp, out,
                       st, acc);
}
}  // namespace

// acc (nullable): also add the solution into acc, row by row as it is written
int vk_launch_solve(const vk_plan *plan, const vk_columns *cols, const double *Dinv, const double *Lsub,
                    const double *Usup, const double *in, double *out, double *tmp, const vk_colstate *st,
                    const vk_bcr_ws *ws, hipStream_t s, double *acc = nullptr) {
    const int ni = plan->ni, nc = cols->ncols;
    const int *off = cols->d_off;
    if (plan->factor_variant == VK_FACTOR_BCR) {
        if (!ws || !ws->A) return VK_EINVAL;
        return vk_launch_bcr_solve(ni, cols, Dinv, Lsub, Usup, in, out, tmp, st, ws, acc, s);
    }
    switch ((ni + 15) / 16) {
    case 1: launch_solve_n<1>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 2: launch_solve_n<2>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 3: launch_solve_n<3>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 4: launch_solve_n<4>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 5: launch_solve_n<5>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 6: launch_solve_n<6>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 7: launch_solve_n<7>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 8: launch_solve_n<8>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 9: launch_solve_n<9>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    case 10: launch_solve_n<10>(ni, nc, off, cols->d_chain_order[1], Dinv, Lsub, Usup, in, out, tmp, st, acc, s); break;
    default: return VK_ETOOBIG;
    }
    VK_CHECK(hipGetLastError());
    return 0;
}

int vk_launch_residual(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Lsub,
                       const double *Usup, const double *x, const double *b, double *r, const vk_colstate *st,
                       hipStream_t s) {
    const int ni = plan->ni;
    const size_t smem = sizeof(double) * ((size_t)ni + 2 * NT);
    const dim3 grid(cols->L, (ni + RES_ROWS - 1) / RES_ROWS);
    // runs of ceil(ni / RES_PARTS) <= ceil(16 nb / RES_PARTS) entries
    switch ((ni + 15) / 16) {
#define VK_RES(NB)                                                                                                   \
    case NB:                                                                                                         \
        hipLaunchKernelGGL(k_residual<(16 * NB + RES_PARTS - 1) / RES_PARTS>, grid, dim3(NT), smem, s, ni, cols->d_off, \
                           cols->d_layer_col, D, Lsub, Usup, x, b, r, st);                                           \
        break;
        VK_RES(1) VK_RES(2) VK_RES(3) VK_RES(4) VK_RES(5) VK_RES(6) VK_RES(7) VK_RES(8) VK_RES(9) VK_RES(10)
#undef VK_RES
    default: return VK_ETOOBIG;
    }
    VK_CHECK(hipGetLastError());
    return 0;
}

// x = A^-1 b with `refine` steps of iterative refinement (work, tmp: L*ni scratch each)
int vk_launch_solve_refined(const vk_plan *plan, const vk_columns *cols, const double *D, const double *Dinv,
                            const double *Lsub, const double *Usup, const double *b, double *x, double *work,
                            double *tmp, int refine, const vk_colstate *st, const vk_bcr_ws *ws, hipStream_t s) {
    int r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, b, x, tmp, st, ws, s);
    if (r) return r;
    for (int q = 0; q < refine; q++) {
        r = vk_launch_residual(plan, cols, D, Lsub, Usup, x, b, work, st, s);
        if (r) return r;
        // the correction is added into x by the solve's backward sweep
        r = vk_launch_solve(plan, cols, Dinv, Lsub, Usup, work, work, tmp, st, ws, s, x);
        if (r) return r;
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
            hipMallocAsync((void **)&ws.Q, sizeof(double) * ns * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.XB, sizeof(double) * L * nn, s) != hipSuccess ||
            hipMallocAsync((void **)&ws.XC, sizeof(double) * L * nn, s) != hipSuccess)
            return VK_ENOMEM;
        return 0;
    }
    ~BcrTmp() {
        for (double *p : {ws.A, ws.B, ws.C, ws.P, ws.Q, ws.XB, ws.XC})
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

// Workspace of vk_block_solve_system_noalloc: one static array per device
// (each device holds its own copy until the process ends).  Host code never
// uses the array's name as an address: scratch_base() resolves it once per
// device with hipGetSymbolAddress, and only that pointer reaches kernels.
static __device__ __align__(256) double scratch[VK_NOALLOC_SCRATCH_BYTES / sizeof(double)];

static int scratch_base(double **out) {
    static std::atomic<double *> base[64];   // per device, zero-initialised
    int dev = 0;
    VK_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return VK_EINVAL;
    double *p = base[dev].load(std::memory_order_acquire);
    if (!p) {
        void *q = nullptr;
        VK_CHECK(hipGetSymbolAddress(&q, HIP_SYMBOL(scratch)));
        if (!q) return VK_ENOMEM;
        p = static_cast<double *>(q);
        base[dev].store(p, std::memory_order_release);
    }
    *out = p;
    return 0;
}

extern "C" int vk_block_solve_system_noalloc(const vk_plan *plan, const vk_columns *cols, const double *d_D,
                                             const double *d_Lsub, const double *d_Usup, const double *d_b,
                                             double *d_x, int refine, int *d_info, void *stream) {
    if (!plan || !cols || !d_D || !d_Lsub || !d_Usup || !d_b || !d_x || refine < 0 || d_b == d_x) return VK_EINVAL;
    hipStream_t s = (hipStream_t)stream;
    // carve the workspace of vk_block_solve_system out of the scratch, each
    // piece on a 256-byte boundary (32 doubles)
    const size_t nn = (size_t)plan->ni * plan->ni, L = (size_t)cols->L, ns = (size_t)cols->bcr.nsurv + 1;
    const bool bcr = plan->factor_variant == VK_FACTOR_BCR;
    auto up = [](size_t n) { return (n + 31) & ~(size_t)31; };
    const size_t sizes[10] = {L * nn, L * plan->ni, L * plan->ni,
                              bcr ? L * nn : 0, bcr ? L * nn : 0, bcr ? L * nn : 0, bcr ? ns * nn : 0,
                              bcr ? ns * nn : 0, bcr ? L * nn : 0, bcr ? L * nn : 0};
    size_t need = 0;
    for (size_t n : sizes) need += up(n);
    if (need > VK_NOALLOC_SCRATCH_BYTES / sizeof(double)) return VK_ETOOBIG;
    double *base = nullptr;
    int r = scratch_base(&base);
    if (r) return r;
    double *piece[10];
    size_t at = 0;
    for (int q = 0; q < 10; q++) {
        piece[q] = sizes[q] ? base + at : nullptr;
        at += up(sizes[q]);
    }
    vk_bcr_ws ws;
    ws.A = piece[3]; ws.B = piece[4]; ws.C = piece[5]; ws.P = piece[6]; ws.Q = piece[7];
    ws.XB = piece[8]; ws.XC = piece[9];
    r = vk_launch_factor(plan, cols, d_D, d_Lsub, d_Usup, piece[0], d_info, nullptr, &ws, s);
    if (!r)
        r = vk_launch_solve_refined(plan, cols, d_D, piece[0], d_Lsub, d_Usup, d_b, d_x, piece[1], piece[2], refine,
                                    nullptr, &ws, s);
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
__global__ void k_selftest_wave_max(const unsigned *in, unsigned *out) {
    out[threadIdx.x] = wave_max_u32(in[threadIdx.x]);
}
__global__ void k_selftest_recip(int n, const double *in, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = blk_recip(in[i]);
}
}  // namespace

extern "C" int vk_selftest_recip_f64(int n, const double *d_in, double *d_out, void *stream) {
    if (n < 0 || (n > 0 && (!d_in || !d_out))) return VK_EINVAL;
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_selftest_recip, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, d_in, d_out);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_selftest_wave_max_u32(const unsigned *d_in, unsigned *d_out, void *stream) {
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
