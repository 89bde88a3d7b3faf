// Whitening and linear-model GEMMs on the CDNA4 matrix cores.
//
// Y[m, n] = sum_k A[m, k] B[n, k]   (A: M x K row-major, B: N x K row-major)
//
//  * whitening: A = residuals R (chains x observations), B = W, the lower
//    triangular whitener (|W r|^2 = r^T Cd^-1 r); the epilogue squares Y
//    and reduces each row over the tile, so W r never reaches HBM.  Only the
//    k <= n part of a lower-triangular B is visited.
//  * linear forward: A = theta (chains x parameters), B = G (observations x
//    parameters); the epilogue writes Y - d (the residual matrix) or, when B
//    is the pre-whitened W G, reduces (Y - W d)^2 like the whitening.
//
// FP32 path: v_mfma_f32_32x32x2_f32.  Split-BF16 path: operands are stored
// as hi + lo BF16 pairs and each tile accumulates hi*hi + hi*lo + lo*hi with
// v_mfma_f32_32x32x16_bf16 in FP32 — one BF16 product plus one correction
// (refinement) step; with plain=1 only hi*hi is formed.  FP64 path (the
// reference's precision, whitening only): v_mfma_f64_16x16x4_f64 with the
// sum-of-squares epilogue in FP64.
//
// Tiling: 256 threads = 4 wavefronts in a 2 x 2 grid; block tile 128 x 128;
// each wavefront owns 64 x 64 = 2 x 2 MFMA 32x32 tiles.  Operands are
// staged through LDS, double-buffered, one barrier per K step.  Inside a K
// step the lanes of the upper half-wave take the upper half of the K slice
// (a permutation of the summation order), so every lane reads its A and B
// fragments with 16-byte LDS loads.  Operands must be zero-padded to
// multiples of the tile sizes (the host guarantees it).
#include "common.hpp"
#include "kernels.hpp"

#include <hip/hip_bf16.h>

#include <type_traits>

namespace altar {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int BM = 128, BN = 128;

// ---------------------------------------------------------------- epilogue
// 32x32 MFMA accumulator layout: register v of lane l holds
// row i = 8 (v / 4) + 4 (l / 32) + (v % 4), column j = l % 32.
__device__ inline int acc_row(int v, int lane) { return 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3); }

template <int EPI>
__device__ inline void gemm_epilogue(floatx16 (&acc)[2][2], const GemmEpi &ep, int m0, int n0, int tile,
                                     int wm, int wn, float *red /* [2][BM] */)
{
    const int lane = threadIdx.x & 63;
    if constexpr (EPI == EPI_STORE) {
#pragma unroll
        for (int mi = 0; mi < 2; ++mi)
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int col = n0 + wn * 64 + ni * 32 + (lane & 31);
                const float bias = ep.bias ? ep.bias[col] : 0.0f;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int row = m0 + wm * 64 + mi * 32 + acc_row(v, lane);
                    ep.out[(size_t)row * ep.ldo + col] = acc[mi][ni][v] - bias;
                }
            }
    } else {
        if constexpr (EPI == EPI_STORE_SUMSQ) {
#pragma unroll
            for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                for (int ni = 0; ni < 2; ++ni) {
                    const int col = n0 + wn * 64 + ni * 32 + (lane & 31);
#pragma unroll
                    for (int v = 0; v < 16; ++v) {
                        const int row = m0 + wm * 64 + mi * 32 + acc_row(v, lane);
                        ep.out[(size_t)row * ep.ldo + col] = acc[mi][ni][v];
                    }
                }
        }
        // sum over this block's 128 columns of (Y - bias)^2, per row
#pragma unroll
        for (int mi = 0; mi < 2; ++mi) {
            float s[16];
#pragma unroll
            for (int v = 0; v < 16; ++v) s[v] = 0.0f;
#pragma unroll
            for (int ni = 0; ni < 2; ++ni) {
                const int col = n0 + wn * 64 + ni * 32 + (lane & 31);
                const float bias = (EPI == EPI_SUMSQ_BIAS) ? ep.bias[col] : 0.0f;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    float y = acc[mi][ni][v] - bias;
                    s[v] += y * y;
                }
            }
            // reduce across the 32 lanes that share rows (xor within a half-wave)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
#pragma unroll
                for (int off = 1; off < 32; off <<= 1) s[v] += __shfl_xor(s[v], off, 64);
            }
            if ((lane & 31) == 0) {
#pragma unroll
                for (int v = 0; v < 16; ++v) red[wn * BM + wm * 64 + mi * 32 + acc_row(v, lane)] = s[v];
            }
        }
        __syncthreads();
        if (threadIdx.x < BM) {
            const int row = m0 + threadIdx.x;
            ep.partial[(size_t)row * ep.ntiles + tile] = red[threadIdx.x] + red[BM + threadIdx.x];
        }
    }
}

// Block -> (n tile, m tile) on a 1-D grid, m fastest.  Consecutive blocks
// go round-robin to the 8 XCDs, so with the m index fastest each XCD gets
// the same mix of n tiles (when ntm is a multiple of 8, XCD x owns the m
// tiles congruent to x mod 8, for every n).  For a lower-triangular B the cost of an n tile grows with
// its index: the expensive tiles go first (longest-first order), which keeps
// the tail short, and every XCD carries the same share of every cost class.
__device__ inline void tile_coords(int tri, int ntn, int &tn, int &tm)
{
    const int ntm = (int)gridDim.x / ntn;
    const int cls = (int)blockIdx.x / ntm;
    tm = (int)blockIdx.x - cls * ntm;
    tn = tri ? ntn - 1 - cls : cls;
}

// LDS rows hold one K slice of 4 chunks of 16 bytes.  Chunk c of row r is
// stored at chunk position c ^ swz(r), swz(r) = (r / 4) % 4, so that both
// the staging writes (4 lanes per row, 4 rows per 16 lanes) and the fragment
// reads (16 rows, one chunk each) hit 16 distinct 4-bank groups.
__device__ inline int swz(int r) { return (r >> 2) & 3; }

// ---------------------------------------------------------------- FP32
constexpr int BK32 = 16;
constexpr int LDS32 = BK32;  // row pitch (floats), unpadded: the chunk swizzle removes bank conflicts

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_f32(const float *__restrict__ A, size_t lda,
                                                   const float *__restrict__ B, size_t ldb, int Kp,
                                                   int ntn, int tri, GemmEpi ep)
{
    __shared__ __attribute__((aligned(16))) float As[2][BM * LDS32];
    __shared__ __attribute__((aligned(16))) float Bs[2][BN * LDS32];
    __shared__ float red[2 * BM];

    int tn, tm;
    tile_coords(tri, ntn, tn, tm);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int kend = tri ? min(Kp, n0 + BN) : Kp;
    const int nk = kend / BK32;

    // global -> register staging: thread t moves rows t/4 and t/4 + 64,
    // columns 4 (t % 4) .. +3 of each operand tile (LDS chunk (t % 4) ^ swz)
    const int lr = tid >> 2, lc = (tid & 3) * 4;
    const int ls = 4 * ((tid & 3) ^ swz(lr));  // swz(lr + 64) == swz(lr)
    const float *Ag0 = A + (size_t)(m0 + lr) * lda + lc;
    const float *Ag1 = A + (size_t)(m0 + lr + 64) * lda + lc;
    const float *Bg0 = B + (size_t)(n0 + lr) * ldb + lc;
    const float *Bg1 = B + (size_t)(n0 + lr + 64) * ldb + lc;

    floatx4 ra0, ra1, rb0, rb1;
    auto gload = [&](int kt) {
        const int k = kt * BK32;
        ra0 = *(const floatx4 *)(Ag0 + k);
        ra1 = *(const floatx4 *)(Ag1 + k);
        rb0 = *(const floatx4 *)(Bg0 + k);
        rb1 = *(const floatx4 *)(Bg1 + k);
    };
    auto sstore = [&](int buf) {
        *(floatx4 *)(&As[buf][lr * LDS32 + ls]) = ra0;
        *(floatx4 *)(&As[buf][(lr + 64) * LDS32 + ls]) = ra1;
        *(floatx4 *)(&Bs[buf][lr * LDS32 + ls]) = rb0;
        *(floatx4 *)(&Bs[buf][(lr + 64) * LDS32 + ls]) = rb1;
    };

    floatx16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0.0f;

    const int half = lane >> 5, l32 = lane & 31;
    // this lane's rows are 32-aligned bases + l32: the swizzle depends on l32 only
    const int c0 = 4 * ((2 * half) ^ swz(l32)), c1 = 4 * ((2 * half + 1) ^ swz(l32));
    if (nk > 0) {
        gload(0);
        sstore(0);
    }
    __syncthreads();
    // NI: how many of this wave's two 32-column MFMA tiles hold a column
    // below ep.nvalid.  A tile of padding columns only would multiply zero
    // rows of B; its MFMAs are skipped (its accumulator stays zero) while the
    // wave still stages operands and meets every barrier.
    auto kloop = [&](auto NI) {
        for (int kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) gload(kt + 1);
            // fragments: lane (half, l32) holds k = 8 half .. 8 half + 7 of row l32
            floatx4 a[2][2], b[2][2];
#pragma unroll
            for (int mi = 0; mi < 2; ++mi) {
                const float *p = &As[cur][(wm * 64 + mi * 32 + l32) * LDS32];
                a[mi][0] = *(const floatx4 *)(p + c0);
                a[mi][1] = *(const floatx4 *)(p + c1);
            }
#pragma unroll
            for (int ni = 0; ni < decltype(NI)::value; ++ni) {
                const float *p = &Bs[cur][(wn * 64 + ni * 32 + l32) * LDS32];
                b[ni][0] = *(const floatx4 *)(p + c0);
                b[ni][1] = *(const floatx4 *)(p + c1);
            }
#pragma unroll
            for (int kk = 0; kk < 8; ++kk) {
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < decltype(NI)::value; ++ni)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x2f32(
                            a[mi][kk >> 2][kk & 3], b[ni][kk >> 2][kk & 3], acc[mi][ni], 0, 0, 0);
            }
            if (kt + 1 < nk) sstore(cur ^ 1);
            __syncthreads();
        }
    };
    const int col0 = n0 + wn * 64;
    if (ep.nvalid <= 0 || col0 + 32 < ep.nvalid)
        kloop(std::integral_constant<int, 2>{});
    else if (col0 < ep.nvalid)
        kloop(std::integral_constant<int, 1>{});
    else
        kloop(std::integral_constant<int, 0>{});
    gemm_epilogue<EPI>(acc, ep, m0, n0, tn, wm, wn, red);
}

// ---------------------------------------------------------------- split BF16
constexpr int BK16 = 32;
constexpr int LDS16 = BK16;  // row pitch (bf16): 64 bytes, chunk-swizzled like the FP32 tiles

template <int EPI>
__global__ __launch_bounds__(256) void gemm_nt_bf16x3(const __bf16 *__restrict__ Ahi, const __bf16 *__restrict__ Alo,
                                                      size_t lda, const __bf16 *__restrict__ Bhi,
                                                      const __bf16 *__restrict__ Blo, size_t ldb, int Kp,
                                                      int ntn, int tri, int plain, GemmEpi ep)
{
    __shared__ __attribute__((aligned(16))) __bf16 Ash[2][BM * LDS16];
    __shared__ __attribute__((aligned(16))) __bf16 Asl[2][BM * LDS16];
    __shared__ __attribute__((aligned(16))) __bf16 Bsh[2][BN * LDS16];
    __shared__ __attribute__((aligned(16))) __bf16 Bsl[2][BN * LDS16];
    __shared__ float red[2 * BM];

    int tn, tm;
    tile_coords(tri, ntn, tn, tm);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int kend = tri ? min(Kp, n0 + BN) : Kp;
    const int nk = kend / BK16;

    // tile = 128 rows x 32 bf16 = 128 x 64 bytes: thread t moves 16 bytes at
    // rows t/4 and t/4 + 64, columns 8 (t % 4) .. +7 (LDS chunk (t % 4) ^ swz)
    const int lr = tid >> 2, lc = (tid & 3) * 8;
    const int ls = 8 * ((tid & 3) ^ swz(lr));
    const size_t a0 = (size_t)(m0 + lr) * lda + lc, a1 = (size_t)(m0 + lr + 64) * lda + lc;
    const size_t b0 = (size_t)(n0 + lr) * ldb + lc, b1 = (size_t)(n0 + lr + 64) * ldb + lc;

    bf16x8 rah0, rah1, ral0, ral1, rbh0, rbh1, rbl0, rbl1;
    auto gload = [&](int kt) {
        const int k = kt * BK16;
        rah0 = *(const bf16x8 *)(Ahi + a0 + k);
        rah1 = *(const bf16x8 *)(Ahi + a1 + k);
        rbh0 = *(const bf16x8 *)(Bhi + b0 + k);
        rbh1 = *(const bf16x8 *)(Bhi + b1 + k);
        if (!plain) {
            ral0 = *(const bf16x8 *)(Alo + a0 + k);
            ral1 = *(const bf16x8 *)(Alo + a1 + k);
            rbl0 = *(const bf16x8 *)(Blo + b0 + k);
            rbl1 = *(const bf16x8 *)(Blo + b1 + k);
        }
    };
    auto sstore = [&](int buf) {
        *(bf16x8 *)(&Ash[buf][lr * LDS16 + ls]) = rah0;
        *(bf16x8 *)(&Ash[buf][(lr + 64) * LDS16 + ls]) = rah1;
        *(bf16x8 *)(&Bsh[buf][lr * LDS16 + ls]) = rbh0;
        *(bf16x8 *)(&Bsh[buf][(lr + 64) * LDS16 + ls]) = rbh1;
        if (!plain) {
            *(bf16x8 *)(&Asl[buf][lr * LDS16 + ls]) = ral0;
            *(bf16x8 *)(&Asl[buf][(lr + 64) * LDS16 + ls]) = ral1;
            *(bf16x8 *)(&Bsl[buf][lr * LDS16 + ls]) = rbl0;
            *(bf16x8 *)(&Bsl[buf][(lr + 64) * LDS16 + ls]) = rbl1;
        }
    };

    floatx16 acc[2][2];
#pragma unroll
    for (int mi = 0; mi < 2; ++mi)
#pragma unroll
        for (int ni = 0; ni < 2; ++ni)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[mi][ni][v] = 0.0f;

    const int half = lane >> 5, l32 = lane & 31;
    if (nk > 0) {
        gload(0);
        sstore(0);
    }
    __syncthreads();
    // NI: as in gemm_nt_f32, the 32-column tiles of this wave below ep.nvalid
    auto kloop = [&](auto NI) {
        for (int kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) gload(kt + 1);
            // lane (half, l32) takes k = 16 half .. 16 half + 15 of row l32 as two
            // groups of 8 (one per MFMA)
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                bf16x8 ah[2], bh[2], al[2], bl[2];
                const int ck = 8 * ((2 * half + kk) ^ swz(l32));
#pragma unroll
                for (int mi = 0; mi < 2; ++mi) {
                    const int off = (wm * 64 + mi * 32 + l32) * LDS16 + ck;
                    ah[mi] = *(const bf16x8 *)(&Ash[cur][off]);
                    if (!plain) al[mi] = *(const bf16x8 *)(&Asl[cur][off]);
                }
#pragma unroll
                for (int ni = 0; ni < decltype(NI)::value; ++ni) {
                    const int off = (wn * 64 + ni * 32 + l32) * LDS16 + ck;
                    bh[ni] = *(const bf16x8 *)(&Bsh[cur][off]);
                    if (!plain) bl[ni] = *(const bf16x8 *)(&Bsl[cur][off]);
                }
#pragma unroll
                for (int mi = 0; mi < 2; ++mi)
#pragma unroll
                    for (int ni = 0; ni < decltype(NI)::value; ++ni) {
                        if (!plain) {
                            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                            acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mi], bl[ni], acc[mi][ni], 0, 0, 0);
                        }
                        acc[mi][ni] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[mi], bh[ni], acc[mi][ni], 0, 0, 0);
                    }
            }
            if (kt + 1 < nk) sstore(cur ^ 1);
            __syncthreads();
        }
    };
    const int col0 = n0 + wn * 64;
    if (ep.nvalid <= 0 || col0 + 32 < ep.nvalid)
        kloop(std::integral_constant<int, 2>{});
    else if (col0 < ep.nvalid)
        kloop(std::integral_constant<int, 1>{});
    else
        kloop(std::integral_constant<int, 0>{});
    gemm_epilogue<EPI>(acc, ep, m0, n0, tn, wm, wn, red);
}

// ---------------------------------------------------------------- FP64
// The reference-precision whitening: partial[m, tile] = sum over the tile's
// 128 columns n of (sum_k A[m, k] B[n, k])^2, every product and sum in FP64
// (v_mfma_f64_16x16x4_f64).  Block tile and wavefront grid as in the FP32
// path; each wavefront owns 64 x 64 = 4 x 4 MFMA 16x16 tiles (4 FP64
// accumulators per lane each).  A K step is 8 doubles = two MFMAs: the lanes
// of group g = lane / 16 read the 16-byte chunk g of their row, k = 2g and
// 2g + 1, and feed k = 2g to the first MFMA and 2g + 1 to the second (a
// permutation of the summation order, the same for A and B).  Operand
// layout of the FP64 16x16x4 MFMA: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator register v of lane l holds row l / 16 + 4 v,
// column l % 16 (rows interleaved, unlike the FP32 16x16 layout; measured).  Two waves per SIMD: the 64 FP64 accumulators per lane
// (128 registers) and the fragments fit 256 registers without spilling.
constexpr int BK64 = 8;
constexpr int LDS64 = BK64;  // row pitch (doubles): 64 bytes, chunk-swizzled like the FP32 tiles

typedef double doublex2 __attribute__((ext_vector_type(2)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void gemm_nt_f64_sumsq(const double *__restrict__ A, size_t lda,
                                                         const double *__restrict__ B, size_t ldb, int Kp,
                                                         int ntn, int tri, int nvalid,
                                                         double *__restrict__ partial, int ntiles)
{
    __shared__ __attribute__((aligned(16))) double As[2][BM * LDS64];
    __shared__ __attribute__((aligned(16))) double Bs[2][BN * LDS64];
    __shared__ double red[2 * BM];

    int tn, tm;
    tile_coords(tri, ntn, tn, tm);
    const int m0 = tm * BM, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int wm = wid >> 1, wn = wid & 1;
    const int kend = tri ? min(Kp, n0 + BN) : Kp;
    const int nk = kend / BK64;

    // global -> register staging: thread t moves chunk t % 4 (doubles
    // 2 (t % 4), +1) of rows t/4 and t/4 + 64 of each operand tile
    const int lr = tid >> 2, lc = (tid & 3) * 2;
    const int ls = 2 * ((tid & 3) ^ swz(lr));  // swz(lr + 64) == swz(lr)
    const double *Ag0 = A + (size_t)(m0 + lr) * lda + lc;
    const double *Ag1 = A + (size_t)(m0 + lr + 64) * lda + lc;
    const double *Bg0 = B + (size_t)(n0 + lr) * ldb + lc;
    const double *Bg1 = B + (size_t)(n0 + lr + 64) * ldb + lc;

    doublex2 ra0, ra1, rb0, rb1;
    auto gload = [&](int kt) {
        const int k = kt * BK64;
        ra0 = *(const doublex2 *)(Ag0 + k);
        ra1 = *(const doublex2 *)(Ag1 + k);
        rb0 = *(const doublex2 *)(Bg0 + k);
        rb1 = *(const doublex2 *)(Bg1 + k);
    };
    auto sstore = [&](int buf) {
        *(doublex2 *)(&As[buf][lr * LDS64 + ls]) = ra0;
        *(doublex2 *)(&As[buf][(lr + 64) * LDS64 + ls]) = ra1;
        *(doublex2 *)(&Bs[buf][lr * LDS64 + ls]) = rb0;
        *(doublex2 *)(&Bs[buf][(lr + 64) * LDS64 + ls]) = rb1;
    };

    doublex4 acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
        for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = doublex4{0.0, 0.0, 0.0, 0.0};

    const int g = lane >> 4, l16 = lane & 15;
    // fragment rows are 16-aligned bases + l16: the swizzle depends on l16 only
    const int cg = 2 * (g ^ swz(l16));
    if (nk > 0) {
        gload(0);
        sstore(0);
    }
    __syncthreads();
    // NI: how many of this wave's four 16-column MFMA tiles hold a column
    // below nvalid (the others multiply zero rows of B and are skipped)
    auto kloop = [&](auto NI) {
        for (int kt = 0; kt < nk; ++kt) {
            const int cur = kt & 1;
            if (kt + 1 < nk) gload(kt + 1);
            doublex2 a[4], b[4];
#pragma unroll
            for (int mi = 0; mi < 4; ++mi)
                a[mi] = *(const doublex2 *)(&As[cur][(wm * 64 + mi * 16 + l16) * LDS64 + cg]);
#pragma unroll
            for (int ni = 0; ni < decltype(NI)::value; ++ni)
                b[ni] = *(const doublex2 *)(&Bs[cur][(wn * 64 + ni * 16 + l16) * LDS64 + cg]);
#pragma unroll
            for (int kk = 0; kk < 2; ++kk)
#pragma unroll
                for (int mi = 0; mi < 4; ++mi)
#pragma unroll
                    for (int ni = 0; ni < decltype(NI)::value; ++ni)
                        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[mi][kk], b[ni][kk], acc[mi][ni], 0, 0, 0);
            if (kt + 1 < nk) sstore(cur ^ 1);
            __syncthreads();
        }
    };
    const int col0 = n0 + wn * 64;
    const int nt = nvalid <= 0 ? 4 : max(0, min(4, (nvalid - col0 + 15) / 16));
    switch (nt) {
    case 4: kloop(std::integral_constant<int, 4>{}); break;
    case 3: kloop(std::integral_constant<int, 3>{}); break;
    case 2: kloop(std::integral_constant<int, 2>{}); break;
    case 1: kloop(std::integral_constant<int, 1>{}); break;
    default: kloop(std::integral_constant<int, 0>{}); break;
    }

    // sum of squares over this block's 128 columns, per row (the skipped
    // tiles' accumulators are zero)
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ni = 0; ni < 4; ++ni)
#pragma unroll
            for (int v = 0; v < 4; ++v) s[v] += acc[mi][ni][v] * acc[mi][ni][v];
        // the 16 lanes of a group share rows: reduce over l16
#pragma unroll
        for (int v = 0; v < 4; ++v) {
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) s[v] += __shfl_xor(s[v], off, 64);
        }
        if (l16 == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) red[wn * BM + wm * 64 + mi * 16 + g + 4 * v] = s[v];
        }
    }
    __syncthreads();
    if (tid < BM) partial[(size_t)(m0 + tid) * ntiles + tn] = red[tid] + red[BM + tid];
}

// ---------------------------------------------------------------- reductions
// llk[c] = normalization - sum_t partial[c, t] / 2, summed in FP64 in a fixed
// order (deterministic).
template <typename T>
__global__ void llk_reduce_kernel(const T *__restrict__ partial, int ntiles, int C, double normalization,
                                  double *__restrict__ llk)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const T *p = partial + (size_t)c * ntiles;
    double s = 0.0;
    for (int t = 0; t < ntiles; ++t) s += (double)p[t];
    llk[c] = normalization - 0.5 * s;
}

// ---------------------------------------------------------------- launchers
hipError_t launch_gemm_f32(const float *A, size_t lda, const float *B, size_t ldb, int Mp, int Np, int Kp,
                           int tri, int epi, const GemmEpi &ep, hipStream_t st)
{
    const int ntn = Np / BN;
    dim3 grid(ntn * (Mp / BM)), block(256);
    if (epi == EPI_STORE)
        hipLaunchKernelGGL((gemm_nt_f32<EPI_STORE>), grid, block, 0, st, A, lda, B, ldb, Kp, ntn, tri, ep);
    else if (epi == EPI_STORE_SUMSQ)
        hipLaunchKernelGGL((gemm_nt_f32<EPI_STORE_SUMSQ>), grid, block, 0, st, A, lda, B, ldb, Kp, ntn, tri, ep);
    else if (epi == EPI_SUMSQ)
        hipLaunchKernelGGL((gemm_nt_f32<EPI_SUMSQ>), grid, block, 0, st, A, lda, B, ldb, Kp, ntn, tri, ep);
    else
        hipLaunchKernelGGL((gemm_nt_f32<EPI_SUMSQ_BIAS>), grid, block, 0, st, A, lda, B, ldb, Kp, ntn, tri, ep);
    return hipGetLastError();
}

hipError_t launch_gemm_bf16x3(const void *Ahi, const void *Alo, size_t lda, const void *Bhi, const void *Blo,
                              size_t ldb, int Mp, int Np, int Kp, int tri, int plain, int epi,
                              const GemmEpi &ep, hipStream_t st)
{
    const int ntn = Np / BN;
    dim3 grid(ntn * (Mp / BM)), block(256);
    const __bf16 *ah = (const __bf16 *)Ahi, *al = (const __bf16 *)Alo;
    const __bf16 *bh = (const __bf16 *)Bhi, *bl = (const __bf16 *)Blo;
    if (epi == EPI_STORE)
        hipLaunchKernelGGL((gemm_nt_bf16x3<EPI_STORE>), grid, block, 0, st, ah, al, lda, bh, bl, ldb, Kp, ntn, tri, plain, ep);
    else if (epi == EPI_STORE_SUMSQ)
        hipLaunchKernelGGL((gemm_nt_bf16x3<EPI_STORE_SUMSQ>), grid, block, 0, st, ah, al, lda, bh, bl, ldb, Kp, ntn, tri, plain, ep);
    else if (epi == EPI_SUMSQ)
        hipLaunchKernelGGL((gemm_nt_bf16x3<EPI_SUMSQ>), grid, block, 0, st, ah, al, lda, bh, bl, ldb, Kp, ntn, tri, plain, ep);
    else
        hipLaunchKernelGGL((gemm_nt_bf16x3<EPI_SUMSQ_BIAS>), grid, block, 0, st, ah, al, lda, bh, bl, ldb, Kp, ntn, tri, plain, ep);
    return hipGetLastError();
}

hipError_t launch_gemm_f64_sumsq(const double *A, size_t lda, const double *B, size_t ldb, int Mp, int Np, int Kp,
                                 int tri, int nvalid, double *partial, int ntiles, hipStream_t st)
{
    if (Mp % BM || Np % BN || Kp % BK64 || lda < (size_t)Kp || ldb < (size_t)Kp || lda % 2 || ldb % 2)
        return hipErrorInvalidValue;
    const int ntn = Np / BN;
    if (ntiles < ntn) return hipErrorInvalidValue;
    dim3 grid(ntn * (Mp / BM)), block(256);
    hipLaunchKernelGGL(gemm_nt_f64_sumsq, grid, block, 0, st, A, lda, B, ldb, Kp, ntn, tri, nvalid, partial, ntiles);
    return hipGetLastError();
}

__global__ void gather_rows_f32_kernel(const float *__restrict__ src, size_t lds, const int *__restrict__ idx, int n,
                                       int cols, float *__restrict__ dst, size_t ldd)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * cols) return;
    const int r = (int)(i / cols), c = (int)(i % cols);
    dst[(size_t)r * ldd + c] = src[(size_t)idx[r] * lds + c];
}

hipError_t launch_gather_rows_f32(const float *src, size_t lds, const int *idx, int n, int cols, float *dst,
                                  size_t ldd, hipStream_t st)
{
    long tot = (long)n * cols;
    if (tot <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_f32_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, src, lds, idx,
                       n, cols, dst, ldd);
    return hipGetLastError();
}

// dst[c, r] = src[r, c] for r < rows, c < cols: 64 x 64 tiles through LDS
// (pitch 65 floats, so the column reads hit distinct banks)
__global__ __launch_bounds__(256) void transpose_f32_kernel(const float *__restrict__ src, size_t lds, int rows,
                                                            int cols, float *__restrict__ dst, size_t ldd)
{
    __shared__ float tile[64][65];
    const int c0 = blockIdx.x * 64, r0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;   // 64 x 4 threads
    for (int j = ty; j < 64; j += 4) {
        const int r = r0 + j, c = c0 + tx;
        tile[j][tx] = (r < rows && c < cols) ? src[(size_t)r * lds + c] : 0.0f;
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const int c = c0 + j, r = r0 + tx;
        if (c < cols && r < rows) dst[(size_t)c * ldd + r] = tile[tx][j];
    }
}

hipError_t launch_transpose_f32(const float *src, size_t lds, int rows, int cols, float *dst, size_t ldd,
                                hipStream_t st)
{
    if (rows <= 0 || cols <= 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_f32_kernel, dim3((cols + 63) / 64, (rows + 63) / 64), dim3(256), 0, st, src, lds, rows,
                       cols, dst, ldd);
    return hipGetLastError();
}

hipError_t launch_llk_reduce(const void *partial, int fp64, int ntiles, int C, double normalization,
                             double *llk, hipStream_t st)
{
    dim3 block(256), grid((C + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL((llk_reduce_kernel<double>), grid, block, 0, st, (const double *)partial, ntiles, C,
                           normalization, llk);
    else
        hipLaunchKernelGGL((llk_reduce_kernel<float>), grid, block, 0, st, (const float *)partial, ntiles, C,
                           normalization, llk);
    return hipGetLastError();
}

}  // namespace altar
