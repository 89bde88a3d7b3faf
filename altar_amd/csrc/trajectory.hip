// Surrogate trajectories in one launch (Liu 2001, the surrogate transition
// method; no reference counterpart).
//
// A workgroup owns 64 chains and runs their whole trajectory: the surrogate
// posterior of the starting state, then `substeps` Metropolis sub-steps of
// pi* = prior + beta llk_s, each one
//   normals -> proposal phi + L xi, with each element's support test and
//   network input -> MLP (BF16 MFMA, FP32 accumulation) -> the chain's
//   model check and prior (beside the last layer) -> surrogate llk and
//   accept or keep phi,
// with the chain state, the activations and the outputs in LDS.  Only the
// weights come from memory (L2-resident; 300 KB for 10 -> 256 x 3 -> 64);
// for the four-layer network the weights of layers 0-2 stay in registers
// for the whole launch and layer 3's in LDS.
//
// Every step repeats the arithmetic of the separate kernels it replaces,
// in the same order: normals_kernel, propose_kernel, prior_verify_kernel,
// cdm_verify_kernel, prior_llk_kernel, mlp_input_kernel, mlp_gemm_kernel
// (the same MFMA sequence over k), surrogate_llk_kernel, posterior_kernel
// and da_traj_step_kernel.  A candidate outside the prior support or above
// the free surface is rejected, as there; the separate path evaluates the
// network on the restored state (restore_kernel) for such a chain, which is
// never read, and here it sees the proposal instead, equally unread.  The
// two paths therefore give the same trajectories bit for bit, and the
// surrogate posterior is one function throughout, which the exact stage's
// ratio relies on.
#include "common.hpp"
#include "kernels.hpp"
#include "chain_math.hpp"

#include <hip/hip_bf16.h>

namespace altar {

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

constexpr int TT = 64;             // chains per workgroup
constexpr int THREADS = 512;       // 8 wavefronts, two per SIMD
constexpr int HID = 256;           // hidden width
constexpr int KIN = 64;            // padded input width
constexpr int KOUT = 64;           // padded output width
constexpr int APITCH = HID + 8;    // bf16 row pitch of the activation tile
constexpr int XPITCH = KIN + 8;    // bf16 row pitch of the network-input tile
constexpr int PP = TRAJ_PMAX + 1;  // row pitch (doubles) of the chain tiles
constexpr int W3PITCH = HID + 8;   // bf16 row pitch of the output layer's weights in LDS

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

// The four-layer network's own path: the same MFMA sequences, k orders and
// epilogues as tile_layer, so the same results bit for bit.  Layers 0, 1
// and 2 use weight fragments held in registers for the whole launch
// (load_chunk, load_fragments), layer 3 a copy of its weights in LDS.
constexpr int HKS = HID / 16;   // 16-byte fragments per lane of one 256-input layer

// fragments ks0 .. ks0 + N - 1 of a lane's weight row (the k-sets of
// tile_layer: 32-wide chunks, within a chunk half * 16 + kk * 8)
template <int N>
__device__ inline void load_chunk(const __bf16 *__restrict__ wrow, int ks0, bf16x8 (&b)[N])
{
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int ks = ks0 + j;
        b[j] = *(const bf16x8 *)(wrow + (ks >> 1) * 32 + (ks & 1) * 8);
    }
}

// the lane's weight row of a layer: output column col, k offset half * 16
__device__ inline const __bf16 *weight_row(const __bf16 *W, int wpitch, int col, int lane)
{
    return W + (size_t)col * wpitch + (lane >> 5) * 16;
}

__device__ void load_fragments(const __bf16 *__restrict__ W, bf16x8 (&b)[HKS], int lane, int wid)
{
    load_chunk(weight_row(W, HID, 32 * wid + (lane & 31), lane), 0, b);
}

__device__ inline void zero_acc(floatx16 (&acc)[2])
{
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[rt][v] = 0.0f;
}

// MFMAs of fragments ks0 .. ks0 + N - 1 over both 32-row tiles of the input
template <int N>
__device__ inline void mfma_chunk(const __bf16 *arow, int apitch, const bf16x8 (&b)[N], int ks0, int half,
                                  floatx16 (&acc)[2])
{
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const int ks = ks0 + j;
        const int k = (ks >> 1) * 32 + half * 16 + (ks & 1) * 8;
        bf16x8 a[2];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) a[rt] = *(const bf16x8 *)(arow + rt * 32 * apitch + k);
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) acc[rt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[rt], b[j], acc[rt], 0, 0, 0);
    }
}

// two accumulator elements in BF16 with one conversion (round to nearest
// even, as the scalar conversion)
__device__ inline bf16x2 to_bf16x2(float a, float b)
{
    return __builtin_convertvector((floatx2){a, b}, bf16x2);
}

// the epilogue of one 32-row tile: bias, GELU, BF16 into rows 8 (v / 4) +
// (v % 4) of ob (elements v and v + 1 are adjacent rows of one column)
__device__ inline void gelu_rows(__bf16 *ob, const floatx16 &acc, float bc)
{
#pragma unroll
    for (int v = 0; v < 16; v += 2) {
        const bf16x2 g = to_bf16x2(gelu(acc[v] + bc), gelu(acc[v + 1] + bc));
        ob[(8 * (v >> 2) + (v & 3)) * APITCH] = g[0];
        ob[(8 * (v >> 2) + (v & 3) + 1) * APITCH] = g[1];
    }
}

// hidden-layer epilogue: bias, GELU, BF16 into the activation tile
__device__ inline void gelu_store(__bf16 *act, const floatx16 (&acc)[2], const float *__restrict__ bias, int lane,
                                  int wid)
{
    const int col = 32 * wid + (lane & 31);
    const float bc = bias[col];
    __bf16 *ab = act + 4 * (lane >> 5) * APITCH + col;
    gelu_rows(ab, acc[0], bc);
    gelu_rows(ab + 32 * APITCH, acc[1], bc);
}

// layer 0 (inputs k < 32) from the register fragments w0; its input tile is
// not its output tile, so no wait between the MFMAs and the epilogue
__device__ void input_layer_resident(const __bf16 *xin, __bf16 *act, const bf16x8 (&w0)[2],
                                     const float *__restrict__ bias, int lane, int wid)
{
    floatx16 acc[2];
    zero_acc(acc);
    mfma_chunk(xin + (lane & 31) * XPITCH, XPITCH, w0, 0, lane >> 5, acc);
    gelu_store(act, acc, bias, lane, wid);
    __syncthreads();
}

// a 256 -> 256 hidden layer from the register fragments b
__device__ void hidden_layer_resident(__bf16 *act, const bf16x8 (&b)[HKS], const float *__restrict__ bias, int lane,
                                      int wid)
{
    const int half = lane >> 5, l32 = lane & 31;
    const int col = 32 * wid + l32;
    const float bc = bias[col];
    const __bf16 *arow = act + l32 * APITCH + half * 16;
    floatx16 acc0, acc1;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc0[v] = acc1[v] = 0.0f;
    // row tile 0, then row tile 1 with row tile 0's epilogue arithmetic
    // between its MFMAs (each accumulator sees the same MFMA sequence over k)
#pragma unroll
    for (int ks = 0; ks < HKS; ++ks) {
        const bf16x8 a0 = *(const bf16x8 *)(arow + (ks >> 1) * 32 + (ks & 1) * 8);
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b[ks], acc0, 0, 0, 0);
    }
    bf16x2 g0[8];
    float gprev = 0.0f;
#pragma unroll
    for (int ks = 0; ks < HKS; ++ks) {
        const bf16x8 a1 = *(const bf16x8 *)(arow + 32 * APITCH + (ks >> 1) * 32 + (ks & 1) * 8);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b[ks], acc1, 0, 0, 0);
        const float g = gelu(acc0[ks] + bc);
        if (ks & 1)
            g0[ks >> 1] = to_bf16x2(gprev, g);
        else
            gprev = g;
    }
    __syncthreads();  // every wave has read its input (the activation tile it overwrites)
    __bf16 *ob = act + 4 * half * APITCH + col;
#pragma unroll
    for (int v = 0; v < 16; v += 2) {
        ob[(8 * (v >> 2) + (v & 3)) * APITCH] = g0[v >> 1][0];
        ob[(8 * (v >> 2) + (v & 3) + 1) * APITCH] = g0[v >> 1][1];
    }
    gelu_rows(ob + 32 * APITCH, acc1, bc);
    __syncthreads();
}

// the 256 -> 64 output layer on waves 0-3, one 32 x 32 tile each (rows
// 32 ((wid >> 1) & 1), columns 32 (wid & 1)), FP32 into zt, its weights
// from the copy w3 in LDS (row pitch W3PITCH).  The caller waits afterwards.
__device__ void output_layer_lds(const __bf16 *act, float *zt, const __bf16 *w3, const float *__restrict__ bias,
                                 int lane, int wid)
{
    const int half = lane >> 5, l32 = lane & 31;
    const int row0 = 32 * ((wid >> 1) & 1), col = 32 * (wid & 1) + l32;
    const __bf16 *wrow = w3 + col * W3PITCH + half * 16;
    const __bf16 *arow = act + (row0 + l32) * APITCH + half * 16;
    floatx16 acc;
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[v] = 0.0f;
#pragma unroll
    for (int ks = 0; ks < HKS; ++ks) {
        const int k = (ks >> 1) * 32 + (ks & 1) * 8;
        const bf16x8 a0 = *(const bf16x8 *)(arow + k);
        const bf16x8 b0 = *(const bf16x8 *)(wrow + k);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    }
    const float bc = bias[col];
    float *zb = zt + (row0 + 4 * half) * (KOUT + 1) + col;
#pragma unroll
    for (int v = 0; v < 16; ++v) zb[(8 * (v >> 2) + (v & 3)) * (KOUT + 1)] = acc[v] + bc;
}

// The small per-parameter and per-row arrays come first, below 64 KB, so
// that their addresses fold into the LDS instructions' offset field.
struct TrajShared {
    double lo[TRAJ_PMAX], hi[TRAJ_PMAX], pa[TRAJ_PMAX], pb[TRAJ_PMAX], pconst[TRAJ_PMAX];
    int kind[TRAJ_PMAX];
    double e[KOUT], s[KOUT];
    double prior_phi[TT], post_phi[TT], cprior[TT], sllk[TT], logu[TT];
    int valid[TT];
    int cur[TT];
    int ebad[TT];              // some element of the candidate is outside the prior support
    double L[TRAJ_PMAX * TRAJ_PMAX];
    float bias[3][HID], bias3[KOUT];  // the four-layer network's biases
    // two chain tiles: row r's state phi is tile cur[r], its candidate the
    // other one, so that an acceptance flips cur[r] instead of copying
    double chain[2][TT * PP], xi[TT * PP];
    float zt[TT * (KOUT + 1)];
    __bf16 xin[TT * XPITCH];   // network input; columns P.. and rows past C stay zero
    __bf16 act[TT * APITCH];
    __bf16 w3[KOUT * W3PITCH];  // the four-layer network's output-layer weights
};

// x as a value the compiler cannot treat as loop-invariant: the address
// arithmetic derived from it inside the sub-step loop is redone each
// sub-step instead of being held in registers across the loop, which leaves
// them to the weights
__device__ inline int opaque(int x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// network input of one parameter value (mlp_input_kernel's scaling)
__device__ inline __bf16 input_value(double x, double lo, double hi)
{
    return (__bf16)(float)(2.0 * (x - lo) / (hi - lo) - 1.0);
}

// network input of the tile's rows from a chain tile: the first P columns
// of the rows below C (the rest of xin is zeroed once per launch)
__device__ void write_input(TrajShared &sh, const double *rows, int c0, const TrajArgs &a, int tid)
{
    for (int it = tid; it < TT * a.P; it += THREADS) {
        const int r = it / a.P, j = it % a.P;
        if (c0 + r < a.C) sh.xin[r * XPITCH + j] = input_value(rows[r * PP + j], sh.lo[j], sh.hi[j]);
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

// surrogate_llk_kernel's eight partial sums of tile row tid / 8 on threads
// 8 r .. 8 r + 7, gathered on the row's first thread (the others' result is
// unused)
__device__ inline double row_llk(const TrajShared &sh, const TrajArgs &a, int tid)
{
    const int lane = tid & 63;
    double p[SLLK_PARTS];
    p[0] = surrogate_llk_part_upto<KOUT>(sh.zt + (tid >> 3) * (KOUT + 1), sh.e, sh.s, a.K, tid & 7);
    const int src = lane < 64 - SLLK_PARTS ? lane : 64 - SLLK_PARTS;  // the last lanes take no part in a combination
#pragma unroll
    for (int q = 1; q < SLLK_PARTS; ++q) p[q] = __shfl(p[0], src + q, 64);
    return surrogate_llk_combine(p, a.cperp, a.norm);
}

// every tile row's surrogate llk into sh.sllk
__device__ void tile_llk(TrajShared &sh, const TrajArgs &a, int tid)
{
    const double llk = row_llk(sh, a, tid);
    if ((tid & 7) == 0) sh.sllk[tid >> 3] = llk;
    __syncthreads();
}

// proposal (propose_kernel) of every element of the tile, with its support
// test (prior_verify_kernel: a failure marks the row in ebad) and its
// network input (mlp_input_kernel)
__device__ void propose_tile(TrajShared &sh, const TrajArgs &a, int c0, int tid)
{
    const int P = a.P;
    for (int it = tid; it < TT * P; it += THREADS) {
        const int rr = it / P, i = it % P;
        if (c0 + rr >= a.C) continue;
        const int cr = sh.cur[rr];
        const double v = propose_elem(sh.chain[cr][rr * PP + i], sh.L + i * P, sh.xi + rr * PP, i);
        sh.chain[cr ^ 1][rr * PP + i] = v;
        if (sh.kind[i] == 0 && !(sh.pa[i] <= v && v <= sh.pb[i])) sh.ebad[rr] = 1;
        sh.xin[rr * XPITCH + i] = input_value(v, sh.lo[i], sh.hi[i]);
    }
}

// the rest of tile row r's checks (cdm_verify_kernel): sh.valid[r]
__device__ void check_row(TrajShared &sh, const TrajArgs &a, int r, int c0)
{
    if (c0 + r >= a.C) return;
    const double *t = sh.chain[sh.cur[r] ^ 1] + r * PP;
    int bad = sh.ebad[r];
    sh.ebad[r] = 0;  // for the next sub-step
    if (a.has_cdm) {
        const CdmLayout &cl = a.cdm;
        if (cdm::above_free_surface(t[cl.dIdx], t[cl.aXIdx], t[cl.aXIdx + 1], t[cl.aXIdx + 2], t[cl.omegaXIdx],
                                    t[cl.omegaXIdx + 1]))
            bad = 1;
    }
    sh.valid[r] = !bad;
}

// the prior of tile row r's candidate (prior_llk_kernel's sum, in order):
// sh.cprior[r], used only if the candidate is valid
__device__ void prior_row(TrajShared &sh, const TrajArgs &a, int r, int c0)
{
    if (c0 + r >= a.C) return;
    const double *t = sh.chain[sh.cur[r] ^ 1] + r * PP;
    double acc = 0.0;
#pragma unroll 4
    for (int j = 0; j < a.P; ++j)
        acc += sh.kind[j] == 0 ? sh.pconst[j] : gauss_log_term(t[j], sh.pa[j], sh.pb[j], sh.pconst[j]);
    sh.cprior[r] = acc;
}

// surrogate llk of every tile row and the sub-step's acceptance
// (da_traj_step_kernel) on the row's first thread, the row's owner
__device__ void llk_accept(TrajShared &sh, const TrajArgs &a, int c0, int tid, unsigned &n_acc, unsigned &n_rej,
                           unsigned &n_unl, int &moved)
{
    const double cdata_s = row_llk(sh, a, tid);
    const int r = tid >> 3;
    if ((tid & 7) == 0 && c0 + r < a.C) {
        if (!sh.valid[r]) {
            ++n_rej;
        } else {
            const double cp = tempered(sh.cprior[r], a.beta, cdata_s);
            if (sh.logu[r] <= cp - sh.post_phi[r]) {
                ++n_acc;
                moved = 1;
                sh.cur[r] ^= 1;  // the candidate's tile holds the state from now on
                sh.prior_phi[r] = sh.cprior[r];
                sh.post_phi[r] = cp;
            } else {
                ++n_unl;
            }
        }
    }
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
    if (tid < TT) {
        sh.ebad[tid] = 0;
        sh.cur[tid] = 0;
    }
    for (int k = tid; k < a.K; k += THREADS) {
        sh.e[k] = a.e[k];
        sh.s[k] = a.s[k];
    }
    for (int it = tid; it < TT * P; it += THREADS) {
        const int r = it / P, j = it % P;
        if (c0 + r < a.C) sh.chain[0][r * PP + j] = a.theta[(size_t)(c0 + r) * P + j];
    }
    for (int it = tid; it < TT * XPITCH; it += THREADS) sh.xin[it] = (__bf16)0.0f;
    const bool four = a.nlayers == 4;
    if (four) {  // layer 3's weights, 16 bytes a thread at a time, and the biases
        for (int it = tid; it < KOUT * HID / 8; it += THREADS) {
            const int row = it / (HID / 8), c8 = it % (HID / 8);
            *(bf16x8 *)(sh.w3 + row * W3PITCH + 8 * c8) = *(const bf16x8 *)(a.W[3] + (size_t)row * HID + 8 * c8);
        }
        for (int it = tid; it < 3 * HID; it += THREADS) sh.bias[it / HID][it % HID] = a.bias[it / HID][it % HID];
        if (tid < KOUT) sh.bias3[tid] = a.bias[3][tid];
    }
    __syncthreads();
    // the surrogate posterior of the starting state
    write_input(sh, sh.chain[0], c0, a, tid);
    network(sh, a, lane, wid);
    tile_llk(sh, a, tid);
    if (tid < TT && c0 + tid < a.C) {
        const int r = tid, c = c0 + tid;
        const double post = tempered(a.prior_theta[c], a.beta, sh.sllk[r]);
        sh.prior_phi[r] = a.prior_theta[c];
        sh.post_phi[r] = post;
        a.post_s[c] = post;
    }
    // the chain of tile row r is owned (accepted, counted) by thread 8 r
    int moved = 0;
    unsigned n_acc = 0, n_rej = 0, n_unl = 0;
    const int pairs = (P + 1) / 2;
    // four-layer networks (the surrogate's 10 -> 256 -> 256 -> 256 -> 64) keep
    // the weights of layers 0-2 in registers for the whole trajectory
    bf16x8 w0[2], w1[HKS], w2[HKS];
    if (four) {
        load_fragments(a.W[1], w1, lane, wid);
        load_chunk(weight_row(a.W[0], KIN, 32 * wid + (lane & 31), lane), 0, w0);
        load_fragments(a.W[2], w2, lane, wid);
    }
    for (int sub = 0; sub < a.substeps; ++sub) {
        const uint32_t step2 = a.step2 + (uint32_t)sub;
        // the thread's indices, opaque to loop-invariant code motion
        const int tid = opaque(threadIdx.x), lane = tid & 63, wid = tid >> 6;
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
        propose_tile(sh, a, c0, tid);
        __syncthreads();
        if (four) {
            input_layer_resident(sh.xin, sh.act, w0, sh.bias[0], lane, wid);
            hidden_layer_resident(sh.act, w1, sh.bias[1], lane, wid);
            hidden_layer_resident(sh.act, w2, sh.bias[2], lane, wid);
            if (wid < 4)
                output_layer_lds(sh.act, sh.zt, sh.w3, sh.bias3, lane, wid);
            else if (wid == 4)
                check_row(sh, a, lane, c0);  // beside the output layer
            else if (wid == 5)
                prior_row(sh, a, lane, c0);
            __syncthreads();
        } else {
            if (wid == 4) check_row(sh, a, lane, c0);
            if (wid == 5) prior_row(sh, a, lane, c0);
            network(sh, a, lane, wid);
        }
        llk_accept(sh, a, c0, tid, n_acc, n_rej, n_unl, moved);
    }
    for (int it = tid; it < TT * P; it += THREADS) {
        const int rr = it / P, j = it % P;
        if (c0 + rr < a.C) a.phi[(size_t)(c0 + rr) * P + j] = sh.chain[sh.cur[rr]][rr * PP + j];
    }
    if ((tid & 7) == 0 && c0 + (tid >> 3) < a.C) {
        const int r = tid >> 3, c = c0 + r;
        a.prior_phi[c] = sh.prior_phi[r];
        a.post_phi[c] = sh.post_phi[r];
        a.moved[c] = moved;
    }
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

hipError_t launch_traj_fused(const TrajArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(traj_fused_kernel, dim3((a.C + TT - 1) / TT), dim3(THREADS), 0, st, a);
    return hipGetLastError();
}

}  // namespace altar
