// Warm-start MLP kernels: BF16 MFMA GEMM with FP32 accumulation and the
// elementwise pieces of forward, backward and Adam.
//
// GEMM: C[M][N] (fp32) = A(m,k) * B(k,n) (+ bias[n]) with A, B BF16 at
// arbitrary (row, column) strides, so that X*W, dZ*W^T and X^T*dZ share one
// kernel.  A 64 x 64 C tile per 256-thread workgroup, each wave one 32 x 32
// sub-tile on v_mfma_f32_32x32x8_bf16 (k = 8 per instruction), K staged
// through LDS in steps of 32 with both operands stored k-contiguous.
// Optional fused epilogue: Y = SiLU(C) written as BF16 for the next layer.
#include "vk_internal.h"

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40);   // NaN stays NaN
    u += 0x7fffu + ((u >> 16) & 1u);                                                   // round to nearest even
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned int)h) << 16); }

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }

constexpr int TM = 64, TN = 64, TK = 32, KP = TK + 4;   // LDS row pitch (in bf16) avoids bank conflicts

__global__ void __launch_bounds__(256) k_gemm(int M, int N, int K, const unsigned short *__restrict__ A, long sam,
                                              long sak, const unsigned short *__restrict__ B, long sbk, long sbn,
                                              float *__restrict__ C, long ldc, const float *__restrict__ bias,
                                              int accumulate, unsigned short *__restrict__ Yact, long ldy) {
    __shared__ unsigned short As[TM][KP];
    __shared__ unsigned short Bs[TN][KP];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    f32x16 acc;
    for (int q = 0; q < 16; q++) acc[q] = 0.0f;
    const bool a_kfast = (sak == 1);
    const bool b_kfast = (sbk == 1);
    for (int k0 = 0; k0 < K; k0 += TK) {
        // A tile: 64 x 32 elements, 8 per thread; consecutive threads walk the unit-stride dimension
        for (int e = tid; e < TM * TK; e += 256) {
            int r, kk;
            if (a_kfast) { r = e / TK; kk = e % TK; } else { kk = e / TM; r = e % TM; }
            const int gm = m0 + r, gk = k0 + kk;
            As[r][kk] = (gm < M && gk < K) ? A[(long)gm * sam + (long)gk * sak] : (unsigned short)0;
        }
        for (int e = tid; e < TN * TK; e += 256) {
            int c, kk;
            if (b_kfast) { c = e / TK; kk = e % TK; } else { kk = e / TN; c = e % TN; }
            const int gn = n0 + c, gk = k0 + kk;
            Bs[c][kk] = (gn < N && gk < K) ? B[(long)gk * sbk + (long)gn * sbn] : (unsigned short)0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; kk += 8) {
            const int kq = kk + (lane >> 5) * 4;
            bf16x4 a, b;
            const unsigned short *pa = &As[wm + (lane & 31)][kq];
            const unsigned short *pb = &Bs[wn + (lane & 31)][kq];
            a[0] = pa[0]; a[1] = pa[1]; a[2] = pa[2]; a[3] = pa[3];
            b[0] = pb[0]; b[1] = pb[1]; b[2] = pb[2]; b[3] = pb[3];
            acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[q]: row = 8*(q/4) + 4*(lane/32) + q%4, column = lane%32 of the wave's 32 x 32 tile
    const int col = n0 + wn + (lane & 31);
    for (int q = 0; q < 16; q++) {
        const int row = m0 + wm + 8 * (q >> 2) + 4 * (lane >> 5) + (q & 3);
        if (row < M && col < N) {
            float v = acc[q];
            if (bias) v += bias[col];
            float *cp = C + (long)row * ldc + col;
            if (accumulate) v += *cp;
            *cp = v;
            if (Yact) Yact[(long)row * ldy + col] = f2bf(silu(v));
        }
    }
}

// dZ = dY * silu'(Z), written as fp32 and BF16; silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void k_silu_bwd(long n, const float *__restrict__ Z, const float *__restrict__ dY, float *__restrict__ dZ,
                           unsigned short *__restrict__ dZb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float z = Z[i];
    const float s = 1.0f / (1.0f + expf(-z));
    const float d = dY[i] * (s * (1.0f + z * (1.0f - s)));
    dZ[i] = d;
    dZb[i] = f2bf(d);
}

// MSE over masked entries: dY = 2 w (Y - T) / count (fp32 + BF16), loss partials per block
__global__ void k_mse(long n, const float *__restrict__ Y, const float *__restrict__ T,
                      const float *__restrict__ W, float inv_count, float *__restrict__ dY,
                      unsigned short *__restrict__ dYb, float *__restrict__ loss_part) {
    __shared__ float red[256];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (i < n) {
        const float w = W ? W[i] : 1.0f;
        const float d = Y[i] - T[i];
        const float g = 2.0f * w * d * inv_count;
        dY[i] = g;
        dYb[i] = f2bf(g);
        l = w * d * d * inv_count;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// column sums: out[n] = sum_m X[m][n]  (bias gradients), fixed order
__global__ void k_colsum(int M, int N, const float *__restrict__ X, float *__restrict__ out) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; m++) s += X[(long)m * N + n];
    out[n] = s;
}

__global__ void k_to_bf16(long n, const float *__restrict__ x, unsigned short *__restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Adam (Kingma & Ba 2015) on fp32 master weights; refreshes the BF16 copy
__global__ void k_adam(long n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                       float wd, unsigned short *__restrict__ wb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

extern "C" int vk_mlp_gemm(int M, int N, int K, const void *d_A, long sam, long sak, const void *d_B, long sbk,
                           long sbn, float *d_C, long ldc, const float *d_bias, int accumulate, void *d_Yact,
                           long ldy, void *stream) {
    if (M <= 0 || N <= 0 || K <= 0 || !d_A || !d_B || !d_C) return VK_EINVAL;
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM);
    hipLaunchKernelGGL(k_gemm, grid, dim3(256), 0, (hipStream_t)stream, M, N, K, (const unsigned short *)d_A, sam,
                       sak, (const unsigned short *)d_B, sbk, sbn, d_C, ldc, d_bias, accumulate,
                       (unsigned short *)d_Yact, ldy);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_mlp_silu_bwd(long n, const float *d_Z, const float *d_dY, float *d_dZ, void *d_dZb, void *stream) {
    if (n <= 0 || !d_Z || !d_dY || !d_dZ || !d_dZb) return VK_EINVAL;
    hipLaunchKernelGGL(k_silu_bwd, grid1(n), dim3(256), 0, (hipStream_t)stream, n, d_Z, d_dY, d_dZ,
                       (unsigned short *)d_dZb);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_mlp_mse(long n, const float *d_Y, const float *d_T, const float *d_W, float inv_count, float *d_dY,
                          void *d_dYb, float *d_loss_part, void *stream) {
    if (n <= 0 || !d_Y || !d_T || !d_dY || !d_dYb || !d_loss_part) return VK_EINVAL;
    hipLaunchKernelGGL(k_mse, grid1(n), dim3(256), 0, (hipStream_t)stream, n, d_Y, d_T, d_W, inv_count, d_dY,
                       (unsigned short *)d_dYb, d_loss_part);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_mlp_colsum(int M, int N, const float *d_X, float *d_out, void *stream) {
    if (M <= 0 || N <= 0 || !d_X || !d_out) return VK_EINVAL;
    hipLaunchKernelGGL(k_colsum, grid1(N), dim3(256), 0, (hipStream_t)stream, M, N, d_X, d_out);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_mlp_to_bf16(long n, const float *d_x, void *d_y, void *stream) {
    if (n <= 0 || !d_x || !d_y) return VK_EINVAL;
    hipLaunchKernelGGL(k_to_bf16, grid1(n), dim3(256), 0, (hipStream_t)stream, n, d_x, (unsigned short *)d_y);
    VK_CHECK(hipGetLastError());
    return 0;
}

extern "C" int vk_mlp_adam(long n, float *d_w, const float *d_g, float *d_m, float *d_v, float lr, float b1,
                           float b2, float eps, int step, float wd, void *d_wb, void *stream) {
    if (n <= 0 || !d_w || !d_g || !d_m || !d_v || step < 1) return VK_EINVAL;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, grid1(n), dim3(256), 0, (hipStream_t)stream, n, d_w, d_g, d_m, d_v, lr, b1, b2, eps,
                       c1, c2, wd, (unsigned short *)d_wb);
    VK_CHECK(hipGetLastError());
    return 0;
}
