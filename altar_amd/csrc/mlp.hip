// Surrogate MLP kernels: BF16 operands, FP32 accumulation (v_mfma_f32_32x32x16_bf16).
//
// A dense layer is Y = act(X W^T + b) with X: batch x in, W: out x in.  All
// products are "NT" GEMMs, Y[m, n] = sum_k A[m, k] B[n, k]:
//   forward            A = X,       B = W      -> bias + GELU epilogue
//   backward, data     A = dY,      B = W^T    -> times GELU'(pre) epilogue
//   backward, weights  A = dY^T,    B = X^T    -> FP32 gradient epilogue
// The transposed copies (W^T, dY^T, X^T) are produced by small transpose
// kernels and by the optimizer.  Tiles are 64 x 64 x 32 with 4 wavefronts
// (2 x 2) of one 32 x 32 MFMA tile each; every dimension is padded to the
// tile sizes by the host.
#include "common.hpp"
#include "kernels.hpp"
#include "chain_math.hpp"

#include <hip/hip_bf16.h>

namespace altar {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int MT = 64, NT = 64, KT = 32;
constexpr int MLDS = KT + 8;  // row pitch (bf16) of the LDS tiles

// gelu and gelu_grad (the tanh approximation, evaluated as x * logistic(2u)): chain_math.hpp

__device__ inline float bf2f(__bf16 v) { return (float)v; }
__device__ inline __bf16 f2bf(float v) { return (__bf16)v; }

template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const __bf16 *__restrict__ A, int lda,
                                                       const __bf16 *__restrict__ B, int ldb, int K,
                                                       MlpEpi ep)
{
    __shared__ __attribute__((aligned(16))) __bf16 As[MT * MLDS];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[NT * MLDS];
    const int m0 = blockIdx.y * MT, n0 = blockIdx.x * NT;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    // staging: 64 rows x 32 bf16 = 256 x 16 bytes: one 16-byte chunk per thread
    const int lr = tid >> 2, lc = (tid & 3) * 8;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    const int half = lane >> 5, l32 = lane & 31;
    // split-k: block z covers k in [z K / gridDim.z, (z + 1) K / gridDim.z)
    const int kspan = K / (int)gridDim.z;
    const int kbeg = (int)blockIdx.z * kspan;
    for (int k0 = kbeg; k0 < kbeg + kspan; k0 += KT) {
        *(bf16x8 *)(&As[lr * MLDS + lc]) = *(const bf16x8 *)(A + (size_t)(m0 + lr) * lda + k0 + lc);
        *(bf16x8 *)(&Bs[lr * MLDS + lc]) = *(const bf16x8 *)(B + (size_t)(n0 + lr) * ldb + k0 + lc);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
            bf16x8 a = *(const bf16x8 *)(&As[(wm * 32 + l32) * MLDS + half * 16 + kk * 8]);
            bf16x8 b = *(const bf16x8 *)(&Bs[(wn * 32 + l32) * MLDS + half * 16 + kk * 8]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    const int col = n0 + wn * 32 + l32;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        float y = acc[v];
        if constexpr (EPI == MLP_EPI_BIAS_ACT) {
            y += ep.bias[col];
            ep.pre[(size_t)row * ep.ldo + col] = y;
            float o = ep.act ? gelu(y) : y;
            ep.out16[(size_t)row * ep.ldo + col] = f2bf(o);
            if (ep.out32) ep.out32[(size_t)row * ep.ldo + col] = o;
        } else if constexpr (EPI == MLP_EPI_F32) {
            ep.out32[(size_t)blockIdx.z * ep.split_stride + (size_t)row * ep.ldo + col] = y;
        } else {  // MLP_EPI_DGELU
            float g = y * gelu_grad(ep.pre[(size_t)row * ep.ldo + col]);
            ep.out16[(size_t)row * ep.ldo + col] = f2bf(g);
        }
    }
}

// out[c, r] = in[r, c] (bf16), rows x cols input
__global__ void transpose_bf16_kernel(const __bf16 *__restrict__ in, int rows, int cols, int ldi,
                                      __bf16 *__restrict__ out, int ldo)
{
    __shared__ __bf16 tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;   // 32 x 8 threads
    for (int j = ty; j < 32; j += 8) {
        const int r = r0 + j, c = c0 + tx;
        tile[j][tx] = (r < rows && c < cols) ? in[(size_t)r * ldi + c] : f2bf(0.0f);
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) out[(size_t)c * ldo + r] = tile[tx][j];
    }
}

// output-layer gradient of the weighted squared error:
// loss = sum_b sum_k w_k (yhat - y)^2 / B  ->  dY = 2 w_k (yhat - y) / B (bf16);
// also accumulates the loss (FP32 atomics, for monitoring only)
__global__ void mse_grad_kernel(const float *__restrict__ yhat, const float *__restrict__ y, const float *__restrict__ w,
                                int rows, int cols, int ld, int valid_rows, __bf16 *__restrict__ dy,
                                float *__restrict__ loss)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float g = 0.0f, l = 0.0f;
    if (idx < (long)rows * cols) {
        const int r = (int)(idx / cols), c = (int)(idx % cols);
        if (r < valid_rows) {
            const float e = yhat[(size_t)r * ld + c] - y[(size_t)r * ld + c];
            g = 2.0f * w[c] * e / (float)valid_rows;
            l = w[c] * e * e / (float)valid_rows;
        }
        dy[(size_t)r * ld + c] = f2bf(g);
    }
    // one atomic per block (the loss is for monitoring only)
    __shared__ float part[4];
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0 && loss) {
        const float t = part[0] + part[1] + part[2] + part[3];
        if (t != 0.0f) atomicAdd(loss, t);
    }
}

// db[n] = sum_m dY[m, n] (FP32 out from a bf16 gradient), in two passes.
// Pass 1: block (x, y) sums rows [64 y, 64 y + 64) of columns
// [64 x, 64 x + 64): lane = column (coalesced rows), 4 waves of 16 rows each,
// wave partials added in order into part[y][column].  Pass 2 adds the slabs
// in order, so the result does not depend on scheduling.
constexpr int CS_ROWS = 64;              // rows per slab: 4 waves of 16
__global__ __launch_bounds__(256) void colsum_slab_kernel(const __bf16 *__restrict__ dy, int rows, int cols, int ld,
                                                          float *__restrict__ part)
{
    __shared__ float wp[4][64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int c = blockIdx.x * 64 + lane;
    const int r0 = blockIdx.y * CS_ROWS + wid * (CS_ROWS / 4);
    float s = 0.0f;
    if (c < cols) {
#pragma unroll 8
        for (int i = 0; i < CS_ROWS / 4; ++i) {
            const int r = r0 + i;
            if (r < rows) s += bf2f(dy[(size_t)r * ld + c]);
        }
    }
    wp[wid][lane] = s;
    __syncthreads();
    if (wid == 0 && c < cols) part[(size_t)blockIdx.y * cols + c] = ((wp[0][lane] + wp[1][lane]) + wp[2][lane]) + wp[3][lane];
}

__global__ void colsum_total_kernel(const float *__restrict__ part, int slabs, int cols, float *__restrict__ db)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float t = 0.0f;
    for (int y = 0; y < slabs; ++y) t += part[(size_t)y * cols + c];
    db[c] = t;
}

// out[i] = sum_z part[z stride + i], z in order (split-k partial products)
__global__ void splitk_sum_kernel(const float *__restrict__ part, int S, size_t stride, long n, float *__restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float t = 0.0f;
    for (int z = 0; z < S; ++z) t += part[(size_t)z * stride + i];
    out[i] = t;
}

// Adam (Kingma & Ba 2015) on FP32 master weights; refreshes the BF16 copy and
// its transpose.  W is rows x cols (ld = cols), WT is cols x rows.
__global__ void adam_kernel(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                            const float *__restrict__ g, int rows, int cols, float lr, float b1, float b2,
                            float eps, float c1, float c2, __bf16 *__restrict__ w16, __bf16 *__restrict__ wt16)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)rows * cols) return;
    const float gi = g[idx];
    const float mi = b1 * m[idx] + (1.0f - b1) * gi;
    const float vi = b2 * v[idx] + (1.0f - b2) * gi * gi;
    m[idx] = mi;
    v[idx] = vi;
    const float wi = w[idx] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[idx] = wi;
    if (w16) w16[idx] = f2bf(wi);
    if (wt16) {
        const int r = (int)(idx / cols), c = (int)(idx % cols);
        wt16[(size_t)c * rows + r] = f2bf(wi);
    }
}

// surrogate data log-likelihood from predicted normalized PCA coefficients:
// llk = norm - (cperp + sum_k (e_k + s_k z_k)^2) / 2, FP64 sum
__global__ void surrogate_llk_kernel(const float *__restrict__ z, int ldz, int C, int K, const double *__restrict__ e,
                                     const double *__restrict__ s, double cperp, double norm, double *__restrict__ llk)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    llk[c] = surrogate_llk_sum(z + (size_t)c * ldz, e, s, K, cperp, norm);
}

// theta (C x P FP64) -> normalized network input (bf16, C x ldx, zero padded):
// x' = 2 (x - lo) / (hi - lo) - 1
__global__ void mlp_input_kernel(const double *__restrict__ theta, int C, int P, const double *__restrict__ lo,
                                 const double *__restrict__ hi, __bf16 *__restrict__ x, int ldx)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * ldx) return;
    const int c = (int)(idx / ldx), j = (int)(idx % ldx);
    float v = 0.0f;
    if (j < P) v = (float)(2.0 * (theta[(size_t)c * P + j] - lo[j]) / (hi[j] - lo[j]) - 1.0);
    x[idx] = f2bf(v);
}

hipError_t launch_mlp_gemm(int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                           const MlpEpi &ep, hipStream_t st)
{
    dim3 grid(N / NT, M / MT, 1), block(256);
    const __bf16 *a = (const __bf16 *)A, *b = (const __bf16 *)B;
    if (epi == MLP_EPI_BIAS_ACT)
        hipLaunchKernelGGL((mlp_gemm_kernel<MLP_EPI_BIAS_ACT>), grid, block, 0, st, a, lda, b, ldb, K, ep);
    else if (epi == MLP_EPI_F32)
        hipLaunchKernelGGL((mlp_gemm_kernel<MLP_EPI_F32>), grid, block, 0, st, a, lda, b, ldb, K, ep);
    else
        hipLaunchKernelGGL((mlp_gemm_kernel<MLP_EPI_DGELU>), grid, block, 0, st, a, lda, b, ldb, K, ep);
    return hipGetLastError();
}

hipError_t launch_transpose_bf16(const void *in, int rows, int cols, int ldi, void *out, int ldo, hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32), block(256);
    hipLaunchKernelGGL(transpose_bf16_kernel, grid, block, 0, st, (const __bf16 *)in, rows, cols, ldi, (__bf16 *)out,
                       ldo);
    return hipGetLastError();
}

hipError_t launch_mse_grad(const float *yhat, const float *y, const float *w, int rows, int cols, int ld,
                           int valid_rows, void *dy, float *loss, hipStream_t st)
{
    long n = (long)rows * cols;
    hipLaunchKernelGGL(mse_grad_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, yhat, y, w, rows, cols,
                       ld, valid_rows, (__bf16 *)dy, loss);
    return hipGetLastError();
}

int colsum_slabs(int rows) { return (rows + CS_ROWS - 1) / CS_ROWS; }

hipError_t launch_colsum(const void *dy, int rows, int cols, int ld, float *part, float *db, hipStream_t st)
{
    const int slabs = colsum_slabs(rows);
    hipLaunchKernelGGL(colsum_slab_kernel, dim3((cols + 63) / 64, slabs), dim3(256), 0, st, (const __bf16 *)dy, rows,
                       cols, ld, part);
    hipLaunchKernelGGL(colsum_total_kernel, dim3((cols + 255) / 256), dim3(256), 0, st, part, slabs, cols, db);
    return hipGetLastError();
}

hipError_t launch_mlp_gemm_splitk(const void *A, int lda, const void *B, int ldb, int M, int N, int K, int S,
                                  float *part, float *out, hipStream_t st)
{
    MlpEpi ep{};
    ep.out32 = part;
    ep.ldo = N;
    ep.split_stride = (size_t)M * N;
    dim3 grid(N / NT, M / MT, S), block(256);
    hipLaunchKernelGGL((mlp_gemm_kernel<MLP_EPI_F32>), grid, block, 0, st, (const __bf16 *)A, lda, (const __bf16 *)B,
                       ldb, K, ep);
    const long n = (long)M * N;
    hipLaunchKernelGGL(splitk_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, part, S,
                       ep.split_stride, n, out);
    return hipGetLastError();
}

hipError_t launch_adam(float *w, float *m, float *v, const float *g, int rows, int cols, float lr, float b1,
                       float b2, float eps, float c1, float c2, void *w16, void *wt16, hipStream_t st)
{
    long n = (long)rows * cols;
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, w, m, v, g, rows, cols, lr,
                       b1, b2, eps, c1, c2, (__bf16 *)w16, (__bf16 *)wt16);
    return hipGetLastError();
}

hipError_t launch_surrogate_llk(const float *z, int ldz, int C, int K, const double *e, const double *s,
                                double cperp, double norm, double *llk, hipStream_t st)
{
    hipLaunchKernelGGL(surrogate_llk_kernel, dim3((C + 255) / 256), dim3(256), 0, st, z, ldz, C, K, e, s, cperp,
                       norm, llk);
    return hipGetLastError();
}

hipError_t launch_mlp_input(const double *theta, int C, int P, const double *lo, const double *hi, void *x, int ldx,
                            hipStream_t st)
{
    long n = (long)C * ldx;
    hipLaunchKernelGGL(mlp_input_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, theta, C, P, lo, hi,
                       (__bf16 *)x, ldx);
    return hipGetLastError();
}

}  // namespace altar
