// Surrogate MLP kernels: BF16 operands, FP32 accumulation (v_mfma_f32_32x32x16_bf16).
//
// A dense layer is Y = act(X W^T + b) with X: batch x in, W: out x in.  All
// products are "NT" GEMMs, Y[m, n] = sum_k A[m, k] B[n, k]:
//   forward            A = X,       B = W      -> bias + GELU epilogue
//   backward, data     A = dY,      B = W^T    -> times GELU'(pre) epilogue
//   backward, weights  A = dY^T,    B = X^T    -> FP32 gradient epilogue
// The transposed copies are written beside the originals by their producers:
// W^T by the optimizer, X^T by the input kernel and the forward epilogue, dY^T
// by the loss-gradient kernel and the backward-data epilogue (through a
// transposed tile in LDS, so both copies are written in whole rows).  Tiles
// are 64 x 64 with 4 wavefronts (2 x 2) of one 32 x 32 MFMA tile each, k in
// chunks of up to 256 staged whole in LDS (one barrier pair per chunk); every
// dimension is padded to the tile sizes by the host.
#include "common.hpp"
#include "kernels.hpp"
#include "chain_math.hpp"

#include <hip/hip_bf16.h>

namespace altar {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));

constexpr int MT = 64, NT = 64, KT = 32;
constexpr int KC = 256;       // k columns staged in LDS at once
constexpr int CP = KC + 8;    // row pitch (bf16) of the LDS staging tiles
constexpr int TP = 64 + 8;    // row pitch (bf16) of a transposed 64 x 64 output tile in LDS

// gelu and gelu_grad (the tanh approximation, evaluated as x * logistic(2u)): chain_math.hpp

__device__ inline float bf2f(__bf16 v) { return (float)v; }
__device__ inline __bf16 f2bf(float v) { return (__bf16)v; }

// From a 64 x 64 bf16 tile held transposed in LDS (T[c * TP + r] = tile row
// r, column c; NTH threads, 256 or more): writes the transpose to outT (row
// c0 + c at outT + (c0 + c) * ldT + r0), whole 128-byte rows, if outT is set;
// and if cpart is set, the tile's column sums to cpart[c0 + c], each the sum
// of the four 16-row groups ((g0 + g1) + g2) + g3, a group summed row by row
// from zero.  wp: 256 floats of LDS.  Ends with the block synchronized.
template <int NTH>
__device__ inline void transposed_tile_out(const __bf16 *T, float *wp, __bf16 *__restrict__ outT, int ldT, int c0,
                                           int r0, float *__restrict__ cpart)
{
    static_assert(NTH % 256 == 0, "256 threads or a multiple");
    const int tid = threadIdx.x;
    if (outT) {
#pragma unroll
        for (int j = tid; j < 512; j += NTH) {
            const int c = j >> 3, r8 = (j & 7) * 8;
            *(bf16x8 *)(outT + (size_t)(c0 + c) * ldT + r0 + r8) = *(const bf16x8 *)(&T[c * TP + r8]);
        }
    }
    if (cpart) {
        const int c = tid & 63, g = tid >> 6;
        if (tid < 256) {
            float sum = 0.0f;
#pragma unroll
            for (int i = 0; i < 16; ++i) sum += bf2f(T[c * TP + 16 * g + i]);
            wp[g * 64 + c] = sum;
        }
        __syncthreads();
        if (tid < 64) cpart[c0 + c] = ((wp[c] + wp[64 + c]) + wp[128 + c]) + wp[192 + c];
    }
    __syncthreads();
}

// The staging tiles (k chunks of up to KC); after the k loop, the
// transposed output tile (64 x TP bf16) and the column-sum partials (256
// floats) in the same bytes.
constexpr int GEMM_SMEM = (MT + NT) * CP;
static_assert(64 * TP * 2 + 256 * 4 <= GEMM_SMEM * 2, "epilogue tile exceeds the staging LDS");

// One 64 x 64 output tile (rows m0.., columns n0..) over k in [kbeg, kbeg +
// kspan): slice z of a split-k product (its partial at out32 + z
// split_stride), tile row ty (its column-sum partials at row ty of cpart).
template <int EPI>
__device__ inline void mlp_gemm_tile(const __bf16 *__restrict__ A, int lda, const __bf16 *__restrict__ B, int ldb,
                                     int kbeg, int kspan, int m0, int n0, int z, int ty, const MlpEpi &ep,
                                     __bf16 *smem)
{
    __bf16 *As = smem, *Bs = smem + MT * CP;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
    const int half = lane >> 5, l32 = lane & 31;
    const int kend = kbeg + kspan;
    for (int c0 = kbeg; c0 < kend; c0 += KC) {
        // a chunk of kc columns (a multiple of 32) of 64 rows of A and of B,
        // 16-byte pieces, consecutive threads along a row; every load of the
        // chunk is issued before the first LDS store
        const int kc = min(KC, kend - c0), pr = kc / 8, n = 64 * pr / 256;
        bf16x8 ra[KC / 32], rb[KC / 32];
#pragma unroll
        for (int i = 0; i < KC / 32; ++i) {
            if (i < n) {
                const int j = tid + 256 * i, r = j / pr, c = (j - r * pr) * 8;
                ra[i] = *(const bf16x8 *)(A + (size_t)(m0 + r) * lda + c0 + c);
                rb[i] = *(const bf16x8 *)(B + (size_t)(n0 + r) * ldb + c0 + c);
            }
        }
#pragma unroll
        for (int i = 0; i < KC / 32; ++i) {
            if (i < n) {
                const int j = tid + 256 * i, r = j / pr, c = (j - r * pr) * 8;
                *(bf16x8 *)(&As[r * CP + c]) = ra[i];
                *(bf16x8 *)(&Bs[r * CP + c]) = rb[i];
            }
        }
        __syncthreads();
        // the same MFMA sequence over k as one 32-column tile at a time
        for (int k0 = 0; k0 < kc; k0 += KT) {
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                bf16x8 a = *(const bf16x8 *)(&As[(wm * 32 + l32) * CP + k0 + half * 16 + kk * 8]);
                bf16x8 b = *(const bf16x8 *)(&Bs[(wn * 32 + l32) * CP + k0 + half * 16 + kk * 8]);
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
            }
        }
        __syncthreads();
    }
    const int col = n0 + wn * 32 + l32;
    __bf16 o16[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int row = m0 + wm * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        float y = acc[v];
        if constexpr (EPI == MLP_EPI_BIAS_ACT) {
            y += ep.bias[col];
            ep.pre[(size_t)row * ep.ldo + col] = y;
            float o = ep.act ? gelu(y) : y;
            o16[v] = f2bf(o);
            ep.out16[(size_t)row * ep.ldo + col] = o16[v];
            if (ep.out32) ep.out32[(size_t)row * ep.ldo + col] = o;
        } else if constexpr (EPI == MLP_EPI_F32) {
            ep.out32[(size_t)z * ep.split_stride + (size_t)row * ep.ldo + col] = y;
        } else {  // MLP_EPI_DGELU
            float g = y * gelu_grad(ep.pre[(size_t)row * ep.ldo + col]);
            o16[v] = f2bf(g);
            ep.out16[(size_t)row * ep.ldo + col] = o16[v];
        }
    }
    if constexpr (EPI != MLP_EPI_F32) {
        if (ep.out16T || ep.cpart) {   // the same for every block of the launch
            // the k loop ended with a barrier: the staging tiles are free
            __bf16 *T = smem;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                bf16x4 pk = {o16[4 * q], o16[4 * q + 1], o16[4 * q + 2], o16[4 * q + 3]};
                *(bf16x4 *)(&T[(wn * 32 + l32) * TP + wm * 32 + 8 * q + 4 * half]) = pk;
            }
            __syncthreads();
            transposed_tile_out<256>(T, (float *)(smem + 64 * TP), ep.out16T, ep.ldT, n0, m0,
                                ep.cpart ? ep.cpart + (size_t)ty * ep.ldo : nullptr);
        }
    }
}

// grid (N / 64, M / 64, S): block (x, y, z) is tile (y, x) of slice z, k in
// [z K / S, (z + 1) K / S)
template <int EPI>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const __bf16 *__restrict__ A, int lda,
                                                       const __bf16 *__restrict__ B, int ldb, int K,
                                                       MlpEpi ep)
{
    __shared__ __attribute__((aligned(16))) __bf16 smem[GEMM_SMEM];
    const int kspan = K / (int)gridDim.z;
    mlp_gemm_tile<EPI>(A, lda, B, ldb, (int)blockIdx.z * kspan, kspan, (int)blockIdx.y * MT, (int)blockIdx.x * NT,
                       (int)blockIdx.z, (int)blockIdx.y, ep, smem);
}

// The weight gradients of every layer in one launch: each layer is a
// segment of the table (blocks from first on), its blocks ordered as
// mlp_gemm_kernel's grid would be (x fastest, then y, then the slice).  Slice
// z of a layer takes k in [z K / S, (z + 1) K / S) and writes its partial at
// part + z M N; mlp_reduce_kernel adds the slices in order.
__global__ __launch_bounds__(256) void mlp_wgrad_kernel(const MlpWGradSeg *__restrict__ segs, int nseg, int K, int S,
                                                        int ld)
{
    __shared__ __attribute__((aligned(16))) __bf16 smem[GEMM_SMEM];
    int k = 0;
    for (int j = 1; j < nseg; ++j)
        if ((int)blockIdx.x >= segs[j].first) k = j;
    const MlpWGradSeg g = segs[k];
    const int ntn = g.N / NT, ntm = g.M / MT;
    int b = (int)blockIdx.x - g.first;
    const int x = b % ntn;
    b /= ntn;
    const int y = b % ntm, z = b / ntm;
    MlpEpi ep{};
    ep.out32 = g.part;
    ep.ldo = g.N;
    ep.split_stride = (size_t)g.M * g.N;
    const int kspan = K / S;
    mlp_gemm_tile<MLP_EPI_F32>(g.A, ld, g.B, ld, z * kspan, kspan, y * MT, x * NT, z, y, ep, smem);
}

// output-layer gradient of the weighted squared error:
// loss = sum_b sum_k w_k (yhat - y)^2 / B  ->  dY = 2 w_k (yhat - y) / B (bf16;
// rows >= valid_rows give 0), one 64 x 64 tile per block of TILE_THREADS
// (thread: column tid % 64, rows 4 (tid / 64) .. + 3).  Also writes dY^T
// (cols x ldT), the tile's column sums into part[blockIdx.y * cols + column]
// (the slab partials of the bias gradient) and, if loss is set, adds the
// tile's loss to it (FP32 atomics, for monitoring only).
constexpr int TILE_THREADS = 1024;
__global__ __launch_bounds__(TILE_THREADS) void mse_grad_kernel(const float *__restrict__ yhat,
                                                                const float *__restrict__ y,
                                                                const float *__restrict__ w, int cols, int ld,
                                                                int valid_rows, __bf16 *__restrict__ dy,
                                                                __bf16 *__restrict__ dyT, int ldT,
                                                                float *__restrict__ part, float *__restrict__ loss)
{
    __shared__ __attribute__((aligned(16))) __bf16 T[64 * TP];
    __shared__ float wp[256];
    __shared__ float lp[TILE_THREADS / 64];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const float wc = w[c0 + c];
    float l = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 4 * g + i;
        float gr = 0.0f;
        if (r < valid_rows) {
            const float e = yhat[(size_t)r * ld + c0 + c] - y[(size_t)r * ld + c0 + c];
            gr = 2.0f * wc * e / (float)valid_rows;
            l += wc * e * e / (float)valid_rows;
        }
        const __bf16 b = f2bf(gr);
        dy[(size_t)r * ld + c0 + c] = b;
        T[c * TP + 4 * g + i] = b;
    }
    l = wave_sum(l);
    if ((threadIdx.x & 63) == 0) lp[g] = l;
    __syncthreads();
    transposed_tile_out<TILE_THREADS>(T, wp, dyT, ldT, c0, r0, part + (size_t)blockIdx.y * cols);
    if (threadIdx.x == 0 && loss) {
        float t = 0.0f;
        for (int k = 0; k < TILE_THREADS / 64; ++k) t += lp[k];
        if (t != 0.0f) atomicAdd(loss, t);
    }
}

// db[n] = sum_m dY[m, n] (FP32 from the bf16 gradient), in two passes.  Pass
// 1 is the epilogue of dY's producer (mse_grad_kernel, the backward-data GEMM):
// per 64-row slab y and column n, part[y][n] adds the slab's four 16-row groups
// in order, each summed row by row.  Pass 2 (mlp_reduce_kernel) adds the
// slabs in order, so the result does not depend on scheduling.
constexpr int CS_ROWS = 64;              // rows per slab

// t + p[0] + p[stride] + ... + p[(n - 1) stride], added in that order.  The
// loads of a group of 16 are issued before its adds, so a thread waits for
// memory once per group rather than once per term.
__device__ inline float ordered_strided_sum(const float *__restrict__ p, int n, size_t stride)
{
    float t = 0.0f;
    int i = 0;
    for (; i + 16 <= n; i += 16) {
        float v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = p[(size_t)(i + j) * stride];
#pragma unroll
        for (int j = 0; j < 16; ++j) t += v[j];
    }
    for (; i < n; ++i) t += p[(size_t)i * stride];
    return t;
}

// Several ordered sums in one launch (the split-k and column sums of every
// layer): segment k owns the blocks from first_k on, dst[i] = src[i] +
// src[stride + i] + ... (count terms, in order) for i < n.
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const MlpReduceSeg *__restrict__ segs, int nseg)
{
    int k = 0;
    for (int j = 1; j < nseg; ++j)
        if ((int)blockIdx.x >= segs[j].first) k = j;
    const MlpReduceSeg g = segs[k];
    const long i = (long)((int)blockIdx.x - g.first) * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    g.dst[i] = ordered_strided_sum(g.src + i, g.count, g.stride);
}

// Adam (Kingma & Ba 2015) on FP32 master weights, every tensor of the
// network in one launch: segment k (a table in device memory) owns the blocks
// [first_k, first_{k+1}).  Refreshes the BF16 copy and its transpose where the
// segment has them.  W is rows x cols (ld = cols), WT is cols x rows.
__global__ __launch_bounds__(256) void adam_kernel(const AdamSeg *__restrict__ segs, int nseg, float lr, float b1,
                                                   float b2, float eps, float c1, float c2)
{
    int k = 0;
    for (int j = 1; j < nseg; ++j)
        if ((int)blockIdx.x >= segs[j].first) k = j;
    const AdamSeg sg = segs[k];
    const long idx = (long)((int)blockIdx.x - sg.first) * blockDim.x + threadIdx.x;
    if (idx >= (long)sg.rows * sg.cols) return;
    const float gi = sg.g[idx];
    const float mi = b1 * sg.m[idx] + (1.0f - b1) * gi;
    const float vi = b2 * sg.v[idx] + (1.0f - b2) * gi * gi;
    sg.m[idx] = mi;
    sg.v[idx] = vi;
    const float wi = sg.w[idx] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    sg.w[idx] = wi;
    if (sg.w16) sg.w16[idx] = f2bf(wi);
    if (sg.wt16) {
        const int r = (int)(idx / sg.cols), c = (int)(idx % sg.cols);
        sg.wt16[(size_t)c * sg.rows + r] = f2bf(wi);
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

// theta (C x P FP64) -> normalized network input x (bf16, rows x ldx, both
// multiples of 64): x' = 2 (x - lo) / (hi - lo) - 1 for row < C and column
// < P, zero elsewhere; and, if xT is set, its transpose (ldx x ldT).  One
// 64 x 64 tile per block of TILE_THREADS (4 rows per thread), staged
// transposed in LDS.
__global__ __launch_bounds__(TILE_THREADS) void mlp_input_kernel(const double *__restrict__ theta, int C, int P,
                                                                 const double *__restrict__ lo,
                                                                 const double *__restrict__ hi,
                                                                 __bf16 *__restrict__ x, int ldx,
                                                                 __bf16 *__restrict__ xT, int ldT)
{
    __shared__ __attribute__((aligned(16))) __bf16 T[64 * TP];
    __shared__ float wp[256];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int j = c0 + c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 4 * g + i;
        float v = 0.0f;
        if (r < C && j < P) v = (float)(2.0 * (theta[(size_t)r * P + j] - lo[j]) / (hi[j] - lo[j]) - 1.0);
        const __bf16 b = f2bf(v);
        x[(size_t)r * ldx + j] = b;
        T[c * TP + 4 * g + i] = b;
    }
    if (xT) {   // the same for every block of the launch
        __syncthreads();
        transposed_tile_out<TILE_THREADS>(T, wp, xT, ldT, c0, r0, nullptr);
    }
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

hipError_t launch_mse_grad(const float *yhat, const float *y, const float *w, int rows, int cols, int ld,
                           int valid_rows, void *dy, void *dyT, int ldT, float *part, float *loss, hipStream_t st)
{
    if (rows % 64 || cols % 64 || ld < cols || ldT < rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mse_grad_kernel, dim3(cols / 64, rows / 64), dim3(TILE_THREADS), 0, st, yhat, y, w, cols, ld,
                       valid_rows, (__bf16 *)dy, (__bf16 *)dyT, ldT, part, loss);
    return hipGetLastError();
}

int colsum_slabs(int rows) { return (rows + CS_ROWS - 1) / CS_ROWS; }

hipError_t launch_mlp_wgrad(const MlpWGradSeg *segs_dev, int nseg, int blocks, int K, int S, int ld, hipStream_t st)
{
    if (S < 1 || K % (S * KT) || ld < K) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3(blocks), dim3(256), 0, st, segs_dev, nseg, K, S, ld);
    return hipGetLastError();
}

hipError_t launch_mlp_reduce(const MlpReduceSeg *segs_dev, int nseg, int blocks, hipStream_t st)
{
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3(blocks), dim3(256), 0, st, segs_dev, nseg);
    return hipGetLastError();
}

hipError_t launch_adam(const AdamSeg *segs_dev, int nseg, int blocks, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(256), 0, st, segs_dev, nseg, lr, b1, b2, eps, c1, c2);
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
                            int rows, void *xT, int ldT, hipStream_t st)
{
    if (rows % 64 || ldx % 64 || rows < C || ldx < P || (xT && ldT < rows)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(mlp_input_kernel, dim3(ldx / 64, rows / 64), dim3(TILE_THREADS), 0, st, theta, C, P, lo, hi,
                       (__bf16 *)x, ldx, (__bf16 *)xT, ldT);
    return hipGetLastError();
}

}  // namespace altar
