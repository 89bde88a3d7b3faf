// Forward models (Mogi point source, compound dislocation model) fused with
// the residual r = LOS.u - shift - d and one of three epilogues:
//   RES_F32   write r (FP32) into a padded C x ldr matrix, the A operand of
//             the whitening GEMM;
//   RES_BF16  write r as a BF16 hi/lo pair (r = hi + lo to ~2^-16) for the
//             split-precision whitening GEMM;
//   DIAG      whiten by a diagonal (1/sigma_i), square and reduce: one
//             partial sum per (chain, observation tile).
//
// Launch shape: one thread per (chain, observation).  A block is 256
// threads = 4 wavefronts over 256 consecutive observations of ONE chain, so
// every per-chain constant is wave-uniform: the Mogi kernel reads them with
// scalar loads, the CDM kernel from a copy of the chain's geometry in LDS.
// The grid is (C, ceil(N / 256)).
#include "common.hpp"
#include "kernels.hpp"
#include "cdm_math.hpp"

#include <hip/hip_bf16.h>

#include <algorithm>

namespace altar {

constexpr int OBS_TILE = 256;

// ------------------------------------------------------------------ epilogue
template <int MODE, typename T>
struct Epilogue;

template <typename T>
__device__ inline void block_partial(T v, T *partial, int chain, int tile, int ntiles)
{
    __shared__ T red[OBS_TILE / 64];
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T s = red[0] + red[1] + red[2] + red[3];
        partial[(size_t)chain * ntiles + tile] = s;
    }
}

// ------------------------------------------------------------------ Mogi
// models/mogi/lib/libmogi/Source.cc:65-118: u = (x nx + y ny - d nz) (nu-1)
// 10^s / (pi R^3) - shift with x = x_src - x_obs; residual u - d_obs.
// Per-chain constants (prepared in FP64): x_src, y_src, d_src, (nu-1)10^s/pi,
// and the dataset shifts.
__global__ void mogi_prepare_kernel(const double *__restrict__ theta, int C, int P, int xIdx,
                                    int dIdx, int sIdx, int offsetIdx, int nsets, double nu,
                                    double *__restrict__ prm /* C x MOGI_PRM */)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double *t = theta + (size_t)c * P;
    double *o = prm + (size_t)c * MOGI_PRM;
    o[0] = t[xIdx];
    o[1] = t[xIdx + 1];
    o[2] = t[dIdx];
    o[3] = (nu - 1.0) * pow(10.0, t[sIdx]) / PI;
    for (int k = 0; k < MAX_SETS; ++k) o[4 + k] = (offsetIdx >= 0 && k < nsets) ? t[offsetIdx + k] : 0.0;
}

template <typename T>
__device__ inline T select_shift(const double *__restrict__ o, int nsets, int id)
{
    T shift = T(0);
    for (int k = 0; k < nsets; ++k) shift = (id == k) ? (T)o[4 + k] : shift;
    return shift;
}

template <typename T, int MODE>
__global__ __launch_bounds__(OBS_TILE) void mogi_kernel(ObsView obs, const double *__restrict__ prm,
                                                        int nsets, int C, Out out)
{
    const int chain = blockIdx.x;
    const int n = blockIdx.y * OBS_TILE + threadIdx.x;
    const double *o = prm + (size_t)chain * MOGI_PRM;
    T val = T(0);
    if (n < obs.N) {
        T xs = (T)o[0], ys = (T)o[1], ds = (T)o[2], cc = (T)o[3];
        T x, y, nx, ny, nz, d;
        if constexpr (sizeof(T) == 8) {
            x = xs - obs.xd[n]; y = ys - obs.yd[n];
            nx = obs.lxd[n]; ny = obs.lyd[n]; nz = obs.lzd[n]; d = obs.dd[n];
        } else {
            x = xs - obs.x[n]; y = ys - obs.y[n];
            nx = obs.lx[n]; ny = obs.ly[n]; nz = obs.lz[n]; d = obs.d[n];
        }
        T R2 = x * x + y * y + ds * ds;
        T R = sqrt(R2);
        T cr3 = cc / (R2 * R);
        T u = (x * nx + y * ny - ds * nz) * cr3;
        u -= select_shift<T>(o, nsets, obs.oid[n]);
        val = u - d;
    }
    if constexpr (MODE == MODE_DIAG) {
        T w = n < obs.N ? (sizeof(T) == 8 ? (T)obs.wdd[n] : (T)obs.wd[n]) : T(0);
        T z = val * w;
        block_partial<T>(z * z, (T *)out.partial, chain, blockIdx.y, gridDim.y);
    } else if constexpr (MODE == MODE_RES_F32) {
        if (n < obs.N) out.r32[(size_t)chain * out.ld + n] = (float)val;
    } else if constexpr (MODE == MODE_RES_F64) {
        if (n < obs.N) out.r64[(size_t)chain * out.ld + n] = (double)val;
    } else {
        if (n < obs.N) {
            float v = (float)val;
            __hip_bfloat16 hi = __float2bfloat16(v);
            __hip_bfloat16 lo = __float2bfloat16(v - __bfloat162float(hi));
            out.rhi[(size_t)chain * out.ld + n] = hi;
            out.rlo[(size_t)chain * out.ld + n] = lo;
        }
    }
}

// ------------------------------------------------------------------ CDM
// Geometry preparation, FP64 (libcdm.py:163-273), one thread per (chain,
// side): 16 chains x 12 sides per block.  Each thread builds the vertices
// (make_vertices) and its side (make_side_of, as make_geometry does for all
// twelve), and writes that side in FP64 and FP32; the chain's flags are
// OR-ed in LDS, and the side-0 thread writes the vertices, the dataset
// shifts and the flags.
constexpr int PREP_CHAINS = 16;
__global__ __launch_bounds__(PREP_CHAINS * 12) void cdm_prepare_sides_kernel(
    const double *__restrict__ theta, int C, int P, CdmLayout L, double nu, double fp64_sin,
    cdm::Geom<float> *__restrict__ g32, cdm::Geom<double> *__restrict__ g64, unsigned *__restrict__ flags,
    float *__restrict__ shifts)
{
    __shared__ unsigned sflags[PREP_CHAINS];
    const int lc = threadIdx.x / 12, s = threadIdx.x % 12;
    const int c = blockIdx.x * PREP_CHAINS + lc;
    if (s == 0) sflags[lc] = 0u;
    __syncthreads();
    if (c < C) {
        const double *t = theta + (size_t)c * P;
        const double ax = t[L.aXIdx], ay = t[L.aXIdx + 1], az = t[L.aXIdx + 2], opening = t[L.openingIdx];
        double V[12][3];
        unsigned f = 0;
        if (cdm::make_vertices(t[L.xIdx], t[L.xIdx + 1], t[L.dIdx], ax, ay, az, t[L.omegaXIdx],
                               t[L.omegaXIdx + 1], t[L.omegaXIdx + 2], V))
            f |= cdm::FLAG_INVALID;
        cdm::Side<double> sd;
        f |= cdm::make_side_of(V, ax, ay, az, opening, nu, fp64_sin, s, sd);
        g64[c].side[s] = sd;
        cdm::Side<float> sf;
        cdm::convert_side(sd, sf);
        g32[c].side[s] = sf;
        atomicOr(&sflags[lc], f);
        if (s == 0) {
            for (int v = 0; v < 12; ++v) {
                const double va = -V[v][2];
                const double ka2 = 0.5 * (1.0 - 2.0 * nu) * va;
                g64[c].vx[v] = V[v][0]; g64[c].vy[v] = V[v][1]; g64[c].va[v] = va; g64[c].ka2[v] = ka2;
                g64[c].va2[v] = va * va;
                g32[c].va2[v] = (float)(va * va);
                g32[c].vx[v] = (float)V[v][0]; g32[c].vy[v] = (float)V[v][1];
                g32[c].va[v] = (float)va; g32[c].ka2[v] = (float)ka2;
            }
            for (int k = 0; k < MAX_SETS; ++k)
                shifts[(size_t)c * MAX_SETS + k] =
                    (L.offsetIdx >= 0 && k < L.nsets) ? (float)t[L.offsetIdx + k] : 0.0f;
        }
    }
    __syncthreads();
    if (s == 0 && c < C) flags[c] = sflags[lc];
}

// INLINE64: the FP64 sides are evaluated here; otherwise they are skipped and
// cdm_fp64_kernel adds them afterwards (residual modes), which keeps this
// kernel at about half the registers and twice the waves per SIMD.
template <int MODE, bool INLINE64>
__global__ __launch_bounds__(OBS_TILE) void cdm_kernel(ObsView obs, const cdm::Geom<float> *__restrict__ g32,
                                                       const cdm::Geom<double> *__restrict__ g64,
                                                       const unsigned *__restrict__ flags,
                                                       const float *__restrict__ shifts, int nsets,
                                                       float nu, int exact64, Out out)
{
    const int chain = blockIdx.x;
    const int n = blockIdx.y * OBS_TILE + threadIdx.x;
    // The chain's FP32 geometry is copied into LDS, so the side constants
    // reach the arithmetic as vector registers (LDS broadcast reads): on
    // gfx950, back-to-back VALU instructions that read a scalar register
    // issue at about half the rate of all-VGPR ones (a rate probe: 4.1 against
    // 2.2 cycles per wave64 multiply).  With the constants read by scalar
    // loads instead, the FP32 pass at 16,384 chains x 5,000 LOS takes 10% longer
    // (same results bit for bit).
    __shared__ cdm::Geom<float> sG;
    {
        static_assert(sizeof(cdm::Geom<float>) % 16 == 0, "FP32 geometry is copied in 16-byte pieces");
        constexpr int NV4 = (int)(sizeof(cdm::Geom<float>) / 16);
        static_assert(NV4 <= OBS_TILE, "one piece per thread");
        const float4 *src = reinterpret_cast<const float4 *>(g32 + chain);
        if (threadIdx.x < NV4) reinterpret_cast<float4 *>(&sG)[threadIdx.x] = src[threadIdx.x];
        __syncthreads();
    }
    const cdm::Geom<float> &G = sG;
    const unsigned fl = flags[chain];
    float val = 0.0f;
    double val64 = 0.0;   // MODE_RES_F64: the residual formed in FP64
    if (n < obs.N) {
        if (fl & (cdm::FLAG_INVALID | cdm::FLAG_NAN)) {
            val = __builtin_nanf("");
            val64 = __builtin_nan("");
        } else {
            const float X = obs.x[n], Y = obs.y[n];
            const float k = 1.0f - 2.0f * nu;
            float ue = 0.0f, un = 0.0f, uv = 0.0f;
            double due = 0.0, dun = 0.0, duv = 0.0;
            // rectangle by rectangle: the four sides of rectangle q join its
            // vertices 4q .. 4q + 3 only, so the per-vertex terms live for one
            // rectangle (fewer registers, more waves per SIMD)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                if ((fl & (0xfu << (4 * q))) == (0xfu << (4 * q))) continue;  // rectangle switched off
                float dx[4], dy[4], r[4], ir[4], irpa[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int v = 4 * q + j;
                    dx[j] = X - G.vx[v];
                    dy[j] = Y - G.vy[v];
                    float a = G.va[v];
                    const float r2 = dx[j] * dx[j] + (dy[j] * dy[j] + G.va2[v]);
                    ir[j] = cdm::f_rsq(r2);
                    r[j] = r2 * ir[j];
                    irpa[j] = cdm::f_rcp(r[j] + a);
                }
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int s = 4 * q + j;
                    if (fl & cdm::skip_bit(s)) continue;
                    const int a = cdm::side_vA(s), b = cdm::side_vB(s);
                    const int ja = a - 4 * q, jb = b - 4 * q;
                    if (!INLINE64 && (fl & cdm::fp64_bit(s))) continue;
                    if (INLINE64 && (fl & cdm::fp64_bit(s))) {
                        // near-vertical side: the pair is evaluated in FP64
                        const cdm::Geom<double> &H = g64[chain];
                        double Xd = obs.xd[n], Yd = obs.yd[n];
                        double dxa = Xd - H.vx[a], dya = Yd - H.vy[a], aa = H.va[a];
                        double dxb = Xd - H.vx[b], dyb = Yd - H.vy[b], ab = H.va[b];
                        const double ra2 = dxa * dxa + dya * dya + aa * aa;
                        const double rb2 = dxb * dxb + dyb * dyb + ab * ab;
                        const double ira = cdm::f_rsq(ra2), irb = cdm::f_rsq(rb2);
                        const double ra = ra2 * ira, rb = rb2 * irb;
                        cdm::side_disp<double>(H.side[s], (double)nu, 1.0 - 2.0 * (double)nu, dxa, dya, aa,
                                               H.ka2[a], ra, ira, cdm::f_rcp(ra + aa), ab, H.ka2[b], rb,
                                               irb, cdm::f_rcp(rb + ab), due, dun, duv, exact64 != 0);
                    } else {
                        cdm::side_disp<float>(G.side[s], nu, k, dx[ja], dy[ja], G.va[a], G.ka2[a], r[ja], ir[ja],
                                              irpa[ja], G.va[b], G.ka2[b], r[jb], ir[jb], irpa[jb], ue, un, uv);
                    }
                }
            }
            float fe = ue + (float)due, fn = un + (float)dun, fv = uv + (float)duv;
            float u = fe * obs.lx[n] + fn * obs.ly[n] + fv * obs.lz[n];
            if constexpr (MODE == MODE_ACC_F32 || MODE == MODE_ACC_F64) {
                val = u;
            } else {
                const float *sh = shifts + (size_t)chain * MAX_SETS;
                float shift = 0.0f;
                const int id = obs.oid[n];
                for (int q = 0; q < nsets; ++q) shift = (id == q) ? sh[q] : shift;
                if constexpr (MODE == MODE_RES_F64)
                    val64 = (double)u - (double)shift - obs.dd[n];
                else
                    val = u - shift - obs.d[n];
            }
        }
    }
    if constexpr (MODE == MODE_ACC_F32) {
        if (n < obs.N) out.r32[(size_t)chain * out.ld + n] += val;
    } else if constexpr (MODE == MODE_ACC_F64) {
        if (n < obs.N) out.r64[(size_t)chain * out.ld + n] += (double)val;
    } else if constexpr (MODE == MODE_DIAG) {
        float w = n < obs.N ? obs.wd[n] : 0.0f;
        float z = val * w;
        block_partial<float>(z * z, (float *)out.partial, chain, blockIdx.y, gridDim.y);
    } else if constexpr (MODE == MODE_RES_F32) {
        if (n < obs.N) out.r32[(size_t)chain * out.ld + n] = val;
    } else if constexpr (MODE == MODE_RES_F64) {
        if (n < obs.N) out.r64[(size_t)chain * out.ld + n] = val64;
    } else {
        if (n < obs.N) {
            __hip_bfloat16 hi = __float2bfloat16(val);
            __hip_bfloat16 lo = __float2bfloat16(val - __bfloat162float(hi));
            out.rhi[(size_t)chain * out.ld + n] = hi;
            out.rlo[(size_t)chain * out.ld + n] = lo;
        }
    }
}

// ------------------------------------------------------------------ launchers
hipError_t launch_mogi(const MogiArgs &a, int mode, int fp64, hipStream_t st)
{
    dim3 pb(256), pg((a.C + 255) / 256);
    hipLaunchKernelGGL(mogi_prepare_kernel, pg, pb, 0, st, a.theta, a.C, a.P, a.xIdx, a.dIdx,
                       a.sIdx, a.offsetIdx, a.nsets, a.nu, a.prm);
    dim3 grid(a.C, (a.obs.N + OBS_TILE - 1) / OBS_TILE), block(OBS_TILE);
#define MOGI_LAUNCH(T, M) hipLaunchKernelGGL((mogi_kernel<T, M>), grid, block, 0, st, a.obs, a.prm, a.nsets, a.C, a.out)
    if (fp64) {
        if (mode == MODE_DIAG) MOGI_LAUNCH(double, MODE_DIAG);
        else if (mode == MODE_RES_F64) MOGI_LAUNCH(double, MODE_RES_F64);
        else if (mode == MODE_RES_F32) MOGI_LAUNCH(double, MODE_RES_F32);
        else MOGI_LAUNCH(double, MODE_RES_BF16);
    } else {
        if (mode == MODE_DIAG) MOGI_LAUNCH(float, MODE_DIAG);
        else if (mode == MODE_RES_F64) MOGI_LAUNCH(float, MODE_RES_F64);
        else if (mode == MODE_RES_F32) MOGI_LAUNCH(float, MODE_RES_F32);
        else MOGI_LAUNCH(float, MODE_RES_BF16);
    }
#undef MOGI_LAUNCH
    return hipGetLastError();
}

hipError_t launch_cdm_prepare(const CdmArgs &a, hipStream_t st)
{
    dim3 pb(PREP_CHAINS * 12), pg((a.C + PREP_CHAINS - 1) / PREP_CHAINS);
    hipLaunchKernelGGL(cdm_prepare_sides_kernel, pg, pb, 0, st, a.theta, a.C, a.P, a.L, a.nu, a.fp64_sin,
                       a.g32, a.g64, a.flags, a.shifts);
    return hipGetLastError();
}

hipError_t launch_cdm(const CdmArgs &a, int mode, hipStream_t st)
{
    hipError_t e = launch_cdm_prepare(a, st);
    if (e != hipSuccess) return e;
    return launch_cdm_main(a, mode, st);
}

// The FP64 sides of the chains that have any, added to the residual matrix
// (FP32 or FP64).  Block (tile, g) visits chains g, g + gridDim.y, ... and
// skips those without FP64 sides (a wave-uniform test of the chain's flags);
// the loop ends for every block at C.  EXACT: the Burgers angle from the
// FP64 atan2 (every side in FP64, the reference's precision).  Each chain's
// FP64 geometry is copied into LDS, and the side loop is not unrolled (a
// chain has two or four FP64 sides of twelve): 114 instead of 136 VGPRs, and
// 2% off the whole forward at 16,384 chains x 5,000 LOS.
template <int MODE, bool EXACT>
__global__ __launch_bounds__(OBS_TILE) void cdm_fp64_kernel(ObsView obs, const cdm::Geom<double> *__restrict__ g64,
                                                            const unsigned *__restrict__ flags, int C, double nu,
                                                            Out out)
{
    const int n = blockIdx.x * OBS_TILE + threadIdx.x;
    double Xd = 0.0, Yd = 0.0, lx = 0.0, ly = 0.0, lz = 0.0;
    if (n < obs.N) {
        Xd = obs.xd[n]; Yd = obs.yd[n];
        lx = obs.lxd[n]; ly = obs.lyd[n]; lz = obs.lzd[n];
    }
    const double k = 1.0 - 2.0 * nu;
    for (int chain = blockIdx.y; chain < C; chain += gridDim.y) {
        const unsigned fl = flags[chain];
        if ((fl & (cdm::FLAG_INVALID | cdm::FLAG_NAN)) || !((fl >> 12) & ~fl & 0xfffu)) continue;
        __shared__ cdm::Geom<double> sH;
        {
            static_assert(sizeof(cdm::Geom<double>) % 16 == 0, "FP64 geometry is copied in 16-byte pieces");
            constexpr int NV4 = (int)(sizeof(cdm::Geom<double>) / 16);
            static_assert(NV4 <= OBS_TILE, "one piece per thread");
            const float4 *src = reinterpret_cast<const float4 *>(g64 + chain);
            __syncthreads();
            if (threadIdx.x < NV4) reinterpret_cast<float4 *>(&sH)[threadIdx.x] = src[threadIdx.x];
            __syncthreads();
        }
        const cdm::Geom<double> &H = sH;
        double due = 0.0, dun = 0.0, duv = 0.0;
#pragma unroll 1
        for (int s = 0; s < 12; ++s) {
            if (!(fl & cdm::fp64_bit(s)) || (fl & cdm::skip_bit(s))) continue;
            const int a = cdm::side_vA(s), b = cdm::side_vB(s);
            double dxa = Xd - H.vx[a], dya = Yd - H.vy[a], aa = H.va[a];
            double dxb = Xd - H.vx[b], dyb = Yd - H.vy[b], ab = H.va[b];
            const double ra2 = dxa * dxa + dya * dya + aa * aa;
            const double rb2 = dxb * dxb + dyb * dyb + ab * ab;
            const double ira = cdm::f_rsq(ra2), irb = cdm::f_rsq(rb2);
            const double ra = ra2 * ira, rb = rb2 * irb;
            cdm::side_disp<double>(H.side[s], nu, k, dxa, dya, aa, H.ka2[a], ra, ira, cdm::f_rcp(ra + aa),
                                   ab, H.ka2[b], rb, irb, cdm::f_rcp(rb + ab), due, dun, duv, EXACT);
        }
        if (n < obs.N) {
            const double u = due * lx + dun * ly + duv * lz;
            if constexpr (MODE == MODE_ACC_F32)
                out.r32[(size_t)chain * out.ld + n] += (float)u;
            else
                out.r64[(size_t)chain * out.ld + n] += u;
        }
    }
}

hipError_t launch_cdm_main(const CdmArgs &a, int mode, hipStream_t st)
{
    dim3 grid(a.C, (a.obs.N + OBS_TILE - 1) / OBS_TILE), block(OBS_TILE);
    float nu = (float)a.nu;
    // every side in FP64 (threshold above 1): the reference's precision,
    // with the FP64 atan2 for the Burgers angle
    const int exact64 = a.fp64_sin > 1.0 ? 1 : 0;
    // residual modes take the two-pass path (FP32 sides, then the FP64 sides
    // of the chains that have any); the partial-sum and BF16 epilogues need
    // the complete residual in one kernel
    const bool split = mode == MODE_RES_F32 || mode == MODE_RES_F64 || mode == MODE_ACC_F32 || mode == MODE_ACC_F64;
#define CDM_LAUNCH(M, I) hipLaunchKernelGGL((cdm_kernel<M, I>), grid, block, 0, st, a.obs, a.g32, a.g64, a.flags, a.shifts, a.L.nsets, nu, exact64, a.out)
    if (!split) {
        if (mode == MODE_DIAG) CDM_LAUNCH(MODE_DIAG, true);
        else if (mode == MODE_RES_F32) CDM_LAUNCH(MODE_RES_F32, true);
        else if (mode == MODE_RES_F64) CDM_LAUNCH(MODE_RES_F64, true);
        else if (mode == MODE_ACC_F32) CDM_LAUNCH(MODE_ACC_F32, true);
        else if (mode == MODE_ACC_F64) CDM_LAUNCH(MODE_ACC_F64, true);
        else CDM_LAUNCH(MODE_RES_BF16, true);
        return hipGetLastError();
    }
    if (mode == MODE_RES_F32) CDM_LAUNCH(MODE_RES_F32, false);
    else if (mode == MODE_RES_F64) CDM_LAUNCH(MODE_RES_F64, false);
    else if (mode == MODE_ACC_F32) CDM_LAUNCH(MODE_ACC_F32, false);
    else CDM_LAUNCH(MODE_ACC_F64, false);
#undef CDM_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    dim3 g2((a.obs.N + OBS_TILE - 1) / OBS_TILE, (unsigned)std::min(a.C, 2048));
#define FP64_LAUNCH(M, E) hipLaunchKernelGGL((cdm_fp64_kernel<M, E>), g2, block, 0, st, a.obs, a.g64, a.flags, a.C, a.nu, a.out)
    if (mode == MODE_RES_F32 || mode == MODE_ACC_F32) {
        if (exact64) FP64_LAUNCH(MODE_ACC_F32, true); else FP64_LAUNCH(MODE_ACC_F32, false);
    } else {
        if (exact64) FP64_LAUNCH(MODE_ACC_F64, true); else FP64_LAUNCH(MODE_ACC_F64, false);
    }
#undef FP64_LAUNCH
    return hipGetLastError();
}

}  // namespace altar

namespace altar {

// Model-side verify of the CDM: a source whose rectangles reach above the
// free surface is rejected (the reference raises ValueError there,
// models/cdm/cdm/libcdm.py:219-224).  mask[c] += 1.
__global__ void cdm_verify_kernel(const double *__restrict__ theta, int C, int P, CdmLayout L,
                                  int *__restrict__ mask)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double *t = theta + (size_t)c * P;
    if (cdm::above_free_surface(t[L.dIdx], t[L.aXIdx], t[L.aXIdx + 1], t[L.aXIdx + 2], t[L.omegaXIdx],
                                t[L.omegaXIdx + 1]))
        mask[c] += 1;
}

hipError_t launch_cdm_verify(const double *theta, int C, int P, CdmLayout L, int *mask, hipStream_t st)
{
    hipLaunchKernelGGL(cdm_verify_kernel, dim3((C + 63) / 64), dim3(64), 0, st, theta, C, P, L, mask);
    return hipGetLastError();
}

// How the sides of the prepared chains are evaluated (one workgroup; the
// flags word of each chain as built by make_geometry).
__global__ __launch_bounds__(256) void side_census_kernel(const unsigned *__restrict__ flags, int C,
                                                          unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long acc[4];
    if (threadIdx.x < 4) acc[threadIdx.x] = 0ull;
    __syncthreads();
    unsigned long long f32 = 0, f64 = 0, skip = 0, bad = 0;
    for (int c = threadIdx.x; c < C; c += blockDim.x) {
        const unsigned f = flags[c];
        if (f & (cdm::FLAG_INVALID | cdm::FLAG_NAN)) {
            ++bad;
            continue;
        }
        const unsigned sk = f & 0xfffu, d = (f >> 12) & 0xfffu & ~sk;
        skip += __popc(sk);
        f64 += __popc(d);
        f32 += 12 - __popc(sk) - __popc(d);
    }
    atomicAdd(&acc[0], f32);
    atomicAdd(&acc[1], f64);
    atomicAdd(&acc[2], skip);
    atomicAdd(&acc[3], bad);
    __syncthreads();
    if (threadIdx.x < 4) out[threadIdx.x] = acc[threadIdx.x];
}

hipError_t launch_side_census(const unsigned *flags, int C, unsigned long long *out, hipStream_t st)
{
    hipLaunchKernelGGL(side_census_kernel, dim3(1), dim3(256), 0, st, flags, C, out);
    return hipGetLastError();
}

}  // namespace altar
