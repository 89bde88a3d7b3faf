// Surrogate trajectories in one launch (Liu 2001, the surrogate transition
// method; no reference counterpart).
//
// A workgroup owns 64 chains and runs their whole trajectory: the surrogate
// posterior of the starting state, then `substeps` Metropolis sub-steps of
// pi* = prior + beta llk_s, each one
//   normals -> proposal phi + L xi -> prior support and model check ->
//   prior llk -> MLP (BF16 MFMA, FP32 accumulation) -> surrogate llk ->
//   accept or keep phi,
// with the chain state, the activations and the outputs in LDS.  Only the
// weights come from memory (L2-resident; 300 KB for 10 -> 256 x 3 -> 64).
//
// Every step repeats the arithmetic of the separate kernels it replaces,
// in the same order: normals_kernel, propose_kernel, prior_verify_kernel,
// cdm_verify_kernel, restore_kernel, prior_llk_kernel, mlp_input_kernel,
// mlp_gemm_kernel (the same MFMA sequence over k), surrogate_llk_kernel,
// posterior_kernel and da_traj_step_kernel.  The two paths therefore give
// the same trajectories bit for bit, and the surrogate posterior is one
// function throughout, which the exact stage's ratio relies on.
#include "common.hpp"
#include "kernels.hpp"
#include "chain_math.hpp"

#include <hip/hip_bf16.h>

namespace altar {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int TT = 64;             // chains per workgroup
constexpr int THREADS = 512;       // 8 wavefronts, two per SIMD
constexpr int HID = 256;           // hidden width
constexpr int KIN = 64;            // padded input width
constexpr int KOUT = 64;           // padded output width
constexpr int APITCH = HID + 8;    // bf16 row pitch of the activation tile
constexpr int XPITCH = KIN + 8;    // bf16 row pitch of the network-input tile
constexpr int PP = TRAJ_PMAX + 1;  // row pitch (doubles) of the chain tiles

// One dense layer on the 64-row tile: in (64 x DIN bf16, row pitch ipitch)
// times W^T (W: DOUT x DIN bf16, row-major) plus bias; GELU into act (hidden
// layers, after every wave has read its input, which may be act itself) or
// FP32 into zt (the last layer).  For
// DOUT = 256 each of the 8 waves computes 64 rows x 32 columns (2 MFMA
// tiles), for DOUT = 64 waves 0-3 one 32 x 32 tile each.  The k order is
// mlp_gemm_kernel's: chunks of 32, within a chunk the lanes' k-sets
// half * 16 + kk * 8 for kk = 0, 1.
template <int DIN, int DOUT, bool LAST, int CHUNK = DIN / 16>
__device__ void tile_layer(const __bf16 *in, int ipitch, __bf16 *act, float *zt, const __bf16 *__restrict__ W,
                           int wpitch, const float *__restrict__ bias, int lane, int wid)
{
    constexpr int RT = DOUT == HID ? 2 : 1;
    const int half = lane >> 5, l32 = lane & 31;
    const bool active = DOUT == HID || wid < 4;
    const int row0 = DOUT == HID ? 0 : 32 * ((wid >> 1) & 1);
    const int col0 = DOUT == HID ? 32 * wid : 32 * (wid & 1);
    floatx16 acc[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[rt][v] = 0.0f;
    // the wave's weight fragments of this layer in flight CHUNK at a time
    // (16-byte loads; the whole layer at once by default, at most 64
    // registers), each chunk followed by its MFMAs in k order
    constexpr int KS = DIN / 16;
    static_assert(KS % CHUNK == 0, "chunks must tile the layer");
    if (active) {
        const __bf16 *wrow = W + (size_t)(col0 + l32) * wpitch + half * 16;
#pragma unroll
        for (int k0 = 0; k0 < KS; k0 += CHUNK) {
            bf16x8 b[CHUNK];
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const int ks = k0 + j;
                b[j] = *(const bf16x8 *)(wrow + (ks >> 1) * 32 + (ks & 1) * 8);
            }
#pragma unroll
            for (int j = 0; j < CHUNK; ++j) {
                const int ks = k0 + j;
                const int k = (ks >> 1) * 32 + half * 16 + (ks & 1) * 8;
                bf16x8 a[RT];
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) a[rt] = *(const bf16x8 *)(in + (row0 + rt * 32 + l32) * ipitch + k);
#pragma unroll
                for (int rt = 0; rt < RT; ++rt)
                    acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[rt], b[j], acc[rt], 0, 0, 0);
            }
        }
    }
    __syncthreads();  // every wave has read its input
    if (active) {
        const int col = col0 + l32;
        const float bc = bias[col];
        // one base address per lane, constant offsets per element
        float *zb = zt + (row0 + 4 * half) * (KOUT + 1) + col;
        __bf16 *ab = act + (row0 + 4 * half) * APITCH + col;
#pragma unroll
        for (int rt = 0; rt < RT; ++rt)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const int drow = rt * 32 + 8 * (v >> 2) + (v & 3);
                const float y = acc[rt][v] + bc;
                if constexpr (LAST)
                    zb[drow * (KOUT + 1)] = y;
                else
                    ab[drow * APITCH] = (__bf16)gelu(y);
            }
    }
    __syncthreads();
}

// The hidden 256 x 256 layers with the wave's weight fragments held in
// registers for the whole launch (loaded once by load_fragments): the same
// MFMA sequence, k order and epilogue as tile_layer<HID, HID, false>, so the
// same results bit for bit, without reading the weights again each sub-step.
constexpr int HKS = HID / 16;   // 16-byte fragments per lane of one 256-input layer

__device__ void load_fragments(const __bf16 *__restrict__ W, bf16x8 (&b)[HKS], int lane, int wid)
{
    const int half = lane >> 5, l32 = lane & 31;
    const __bf16 *wrow = W + (size_t)(32 * wid + l32) * HID + half * 16;
#pragma unroll
    for (int ks = 0; ks < HKS; ++ks) b[ks] = *(const bf16x8 *)(wrow + (ks >> 1) * 32 + (ks & 1) * 8);
}

__device__ void hidden_layer_resident(__bf16 *act, const bf16x8 (&b)[HKS], const float *__restrict__ bias, int lane,
                                      int wid)
{
    const int half = lane >> 5, l32 = lane & 31;
    floatx16 acc[2];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[rt][v] = 0.0f;
    const __bf16 *arow = act + l32 * APITCH;
#pragma unroll
    for (int ks = 0; ks < HKS; ++ks) {
        const int k = (ks >> 1) * 32 + half * 16 + (ks & 1) * 8;
        bf16x8 a[2];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) a[rt] = *(const bf16x8 *)(arow + rt * 32 * APITCH + k);
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[rt], b[ks], acc[rt], 0, 0, 0);
    }
    __syncthreads();  // every wave has read its input (the activation tile it overwrites)
    const int col = 32 * wid + l32;
    const float bc = bias[col];
    __bf16 *ab = act + 4 * half * APITCH + col;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 16; ++v) ab[(rt * 32 + 8 * (v >> 2) + (v & 3)) * APITCH] = (__bf16)gelu(acc[rt][v] + bc);
    __syncthreads();
}

struct TrajShared {
    __bf16 act[TT * APITCH];
    __bf16 xin[TT * XPITCH];   // network input; columns P.. and rows past C stay zero
    float zt[TT * (KOUT + 1)];
    double phi[TT * PP], cand[TT * PP], xi[TT * PP];
    double L[TRAJ_PMAX * TRAJ_PMAX];
    double lo[TRAJ_PMAX], hi[TRAJ_PMAX], pa[TRAJ_PMAX], pb[TRAJ_PMAX], pconst[TRAJ_PMAX];
    int kind[TRAJ_PMAX];
    double e[KOUT], s[KOUT];
    double prior_phi[TT], post_phi[TT], cprior[TT], sllk[TT], logu[TT];
    int valid[TT];
};

// network input of the tile's rows from a chain tile: the first P columns
// of the rows below C (the rest of xin is zeroed once per launch)
__device__ void write_input(TrajShared &sh, const double *rows, int c0, const TrajArgs &a)
{
    for (int it = threadIdx.x; it < TT * a.P; it += THREADS) {
        const int r = it / a.P, j = it % a.P;
        if (c0 + r < a.C)
            sh.xin[r * XPITCH + j] = (__bf16)(float)(2.0 * (rows[r * PP + j] - sh.lo[j]) / (sh.hi[j] - sh.lo[j]) - 1.0);
    }
}

__device__ void network(TrajShared &sh, const TrajArgs &a, int lane, int wid)
{
    __syncthreads();
    // layer 0 over k < 32 only: the inputs past P <= 32 are zero, and zero
    // products leave the MFMA sums unchanged (mlp_gemm_kernel runs all 64)
    tile_layer<TRAJ_PMAX, HID, false>(sh.xin, XPITCH, sh.act, sh.zt, a.W[0], KIN, a.bias[0], lane, wid);
    for (int l = 1; l < a.nlayers - 1; ++l)
        tile_layer<HID, HID, false>(sh.act, APITCH, sh.act, sh.zt, a.W[l], HID, a.bias[l], lane, wid);
    tile_layer<HID, KOUT, true>(sh.act, APITCH, sh.act, sh.zt, a.W[a.nlayers - 1], HID, a.bias[a.nlayers - 1], lane,
                                wid);
}

// network() for a network of four layers, the two hidden 256 x 256 layers
// from the register fragments w1, w2
__device__ void network_resident(TrajShared &sh, const TrajArgs &a, const bf16x8 (&w2)[HKS], int lane, int wid)
{
    __syncthreads();
    tile_layer<TRAJ_PMAX, HID, false>(sh.xin, XPITCH, sh.act, sh.zt, a.W[0], KIN, a.bias[0], lane, wid);
    tile_layer<HID, HID, false, 4>(sh.act, APITCH, sh.act, sh.zt, a.W[1], HID, a.bias[1], lane, wid);
    hidden_layer_resident(sh.act, w2, a.bias[2], lane, wid);
    tile_layer<HID, KOUT, true, 4>(sh.act, APITCH, sh.act, sh.zt, a.W[3], HID, a.bias[3], lane, wid);
}

// surrogate_llk_kernel's sum for every tile row into sh.sllk: threads
// 4 r .. 4 r + 3 take the four partial sums of row r, the first adds them in order
__device__ void tile_llk(TrajShared &sh, const TrajArgs &a)
{
    const int tid = threadIdx.x, lane = tid & 63;
    double part = 0.0;
    if (tid < SLLK_PARTS * TT)
        part = surrogate_llk_part(sh.zt + (tid >> 2) * (KOUT + 1), sh.e, sh.s, a.K, tid & 3);
    const int src = lane < 61 ? lane : 60;  // lanes 61-63 take no part in a combination
    const double p1 = __shfl(part, src + 1, 64), p2 = __shfl(part, src + 2, 64), p3 = __shfl(part, src + 3, 64);
    if (tid < SLLK_PARTS * TT && (tid & 3) == 0) sh.sllk[tid >> 2] = surrogate_llk_combine(part, p1, p2, p3, a.cperp, a.norm);
    __syncthreads();
}

}  // namespace

__global__ __launch_bounds__(THREADS) void traj_fused_kernel(TrajArgs a)
{
    __shared__ TrajShared sh;
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int c0 = blockIdx.x * TT;
    const int P = a.P;
    for (int i = tid; i < P * P; i += THREADS) sh.L[i] = a.L[i];
    if (tid < P) {
        const int j = tid;
        sh.lo[j] = a.lo[j];
        sh.hi[j] = a.hi[j];
        sh.kind[j] = a.kind[j];
        sh.pa[j] = a.pa[j];
        sh.pb[j] = a.pb[j];
        // the data-independent term of prior_llk_kernel's density
        sh.pconst[j] = a.kind[j] == 0 ? log(1.0 / (a.pb[j] - a.pa[j])) : log(a.pb[j] * sqrt(2.0 * PI));
    }
    for (int k = tid; k < a.K; k += THREADS) {
        sh.e[k] = a.e[k];
        sh.s[k] = a.s[k];
    }
    for (int it = tid; it < TT * P; it += THREADS) {
        const int r = it / P, j = it % P;
        if (c0 + r < a.C) sh.phi[r * PP + j] = a.theta[(size_t)(c0 + r) * P + j];
    }
    for (int it = tid; it < TT * XPITCH; it += THREADS) sh.xin[it] = (__bf16)0.0f;
    __syncthreads();
    // the surrogate posterior of the starting state
    write_input(sh, sh.phi, c0, a);
    network(sh, a, lane, wid);
    tile_llk(sh, a);
    const int r = tid;  // the chain of a thread of wavefront 0
    const int c = c0 + r;
    const bool mine = tid < TT && c < a.C;
    int moved = 0;
    unsigned n_acc = 0, n_rej = 0, n_unl = 0;
    if (mine) {
        const double llk = sh.sllk[r];
        const double post = tempered(a.prior_theta[c], a.beta, llk);
        sh.prior_phi[r] = a.prior_theta[c];
        sh.post_phi[r] = post;
        a.post_s[c] = post;
    }
    const int pairs = (P + 1) / 2;
    // four-layer networks (the surrogate's 10 -> 256 -> 256 -> 256 -> 64) keep
    // their hidden layers' weights in registers for the whole trajectory
    const bool resident = a.nlayers == 4;
    bf16x8 w2[HKS];
    if (resident) load_fragments(a.W[2], w2, lane, wid);
    for (int sub = 0; sub < a.substeps; ++sub) {
        const uint32_t step2 = a.step2 + (uint32_t)sub;
        // normals (normals_kernel), and the log of each chain's sub-step dice
        // (da_traj_step_kernel's), which do not depend on the sub-step's outcome
        for (int it = tid; it < TT * (pairs + 1); it += THREADS) {
            if (it >= TT * pairs) {
                const int rr = it - TT * pairs;
                if (c0 + rr < a.C) {
                    u32x4 rn = draw(a.seed, a.stage, step2, (uint32_t)(c0 + rr), TAG_TRAJ_ACCEPT, 0u);
                    sh.logu[rr] = log(u53(rn.x, rn.y));
                }
                continue;
            }
            const int rr = it / pairs, w = it % pairs;
            if (c0 + rr >= a.C) continue;
            u32x4 rn = draw(a.seed, a.stage, step2, (uint32_t)(c0 + rr), TAG_TRAJ_PROPOSAL, (uint32_t)w);
            double u1 = u53(rn.x, rn.y), u2 = u53(rn.z, rn.w);
            double rad = sqrt(-2.0 * log(u1));
            double sn, co;
            sincos(2.0 * PI * u2, &sn, &co);
            sh.xi[rr * PP + 2 * w] = rad * co;
            if (2 * w + 1 < P) sh.xi[rr * PP + 2 * w + 1] = rad * sn;
        }
        __syncthreads();
        // proposal (propose_kernel)
        for (int it = tid; it < TT * P; it += THREADS) {
            const int rr = it / P, i = it % P;
            if (c0 + rr >= a.C) continue;
            sh.cand[rr * PP + i] = propose_elem(sh.phi[rr * PP + i], sh.L + i * P, sh.xi + rr * PP, i);
        }
        __syncthreads();
        // support and model check, restore, prior (prior_verify_kernel,
        // cdm_verify_kernel, restore_kernel, prior_llk_kernel)
        if (mine) {
            double *t = sh.cand + r * PP;
            int bad = 0;
            for (int j = 0; j < P; ++j)
                if (sh.kind[j] == 0 && !(sh.pa[j] <= t[j] && t[j] <= sh.pb[j])) bad = 1;
            if (a.has_cdm) {
                const CdmLayout &cl = a.cdm;
                if (cdm::above_free_surface(t[cl.dIdx], t[cl.aXIdx], t[cl.aXIdx + 1], t[cl.aXIdx + 2],
                                            t[cl.omegaXIdx], t[cl.omegaXIdx + 1]))
                    bad = 1;
            }
            sh.valid[r] = !bad;
            if (bad)
                for (int j = 0; j < P; ++j) t[j] = sh.phi[r * PP + j];
            double acc = 0.0;
            for (int j = 0; j < P; ++j) {
                if (sh.kind[j] == 0) {
                    acc += sh.pconst[j];
                } else {
                    acc += gauss_log_term(t[j], sh.pa[j], sh.pb[j], sh.pconst[j]);
                }
            }
            sh.cprior[r] = acc;
        }
        __syncthreads();
        write_input(sh, sh.cand, c0, a);
        if (resident)
            network_resident(sh, a, w2, lane, wid);
        else
            network(sh, a, lane, wid);
        tile_llk(sh, a);
        // sub-step acceptance (da_traj_step_kernel)
        if (mine) {
            const double cdata_s = sh.sllk[r];
            const double cp = tempered(sh.cprior[r], a.beta, cdata_s);
            if (!sh.valid[r]) {
                ++n_rej;
            } else {
                if (sh.logu[r] <= cp - sh.post_phi[r]) {
                    ++n_acc;
                    moved = 1;
                    for (int j = 0; j < P; ++j) sh.phi[r * PP + j] = sh.cand[r * PP + j];
                    sh.prior_phi[r] = sh.cprior[r];
                    sh.post_phi[r] = cp;
                } else {
                    ++n_unl;
                }
            }
        }
        __syncthreads();
    }
    for (int it = tid; it < TT * P; it += THREADS) {
        const int rr = it / P, j = it % P;
        if (c0 + rr < a.C) a.phi[(size_t)(c0 + rr) * P + j] = sh.phi[rr * PP + j];
    }
    if (mine) {
        a.prior_phi[c] = sh.prior_phi[r];
        a.post_phi[c] = sh.post_phi[r];
        a.moved[c] = moved;
    }
    if (wid == 0) {
        unsigned long long s0 = n_acc, s1 = n_rej, s2 = n_unl;
        for (int off = 32; off > 0; off >>= 1) {
            s0 += __shfl_xor(s0, off, 64);
            s1 += __shfl_xor(s1, off, 64);
            s2 += __shfl_xor(s2, off, 64);
        }
        if (lane == 0) {
            if (s0) atomicAdd(&a.counters[0], s0);
            if (s1) atomicAdd(&a.counters[1], s1);
            if (s2) atomicAdd(&a.counters[2], s2);
        }
    }
}

hipError_t launch_traj_fused(const TrajArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(traj_fused_kernel, dim3((a.C + TT - 1) / TT), dim3(THREADS), 0, st, a);
    return hipGetLastError();
}

}  // namespace altar
