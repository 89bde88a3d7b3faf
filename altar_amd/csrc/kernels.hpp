// Internal kernel interfaces of the altar_amd HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <stdint.h>
#include <stddef.h>

#include "cdm_math.hpp"

namespace altar {

constexpr int MAX_SETS = 8;   // datasets (offset parameters) per source
constexpr int MOGI_PRM = 4 + MAX_SETS;

// forward-kernel epilogues
enum { MODE_DIAG = 0, MODE_RES_F32 = 1, MODE_RES_F64 = 2, MODE_RES_BF16 = 3 };
// GEMM epilogues
enum { EPI_STORE = 0, EPI_SUMSQ = 1, EPI_SUMSQ_BIAS = 2 };

// observations, structure of arrays, FP32 and FP64 copies
struct ObsView {
    int N;
    const float *x, *y, *lx, *ly, *lz, *d, *wd;
    const double *xd, *yd, *lxd, *lyd, *lzd, *dd, *wdd;
    const int *oid;
};

struct Out {
    size_t ld;            // row pitch (elements) of the residual matrices
    float *r32;
    double *r64;
    __hip_bfloat16 *rhi, *rlo;
    void *partial;        // DIAG: C x ntiles partial sums (float or double)
};

struct MogiArgs {
    const double *theta;
    int C, P, xIdx, dIdx, sIdx, offsetIdx, nsets;
    double nu;
    double *prm;          // C x MOGI_PRM workspace
    ObsView obs;
    Out out;
};

struct CdmLayout {
    int xIdx, dIdx, openingIdx, aXIdx, omegaXIdx, offsetIdx, nsets;
};

struct CdmArgs {
    const double *theta;
    int C, P;
    CdmLayout L;
    double nu, fp64_sin;
    cdm::Geom<float> *g32;
    cdm::Geom<double> *g64;
    unsigned *flags;
    float *shifts;        // C x MAX_SETS
    ObsView obs;
    Out out;
};

struct GemmEpi {
    const float *bias;    // per column (EPI_STORE: subtracted; EPI_SUMSQ_BIAS: subtracted)
    float *out;           // EPI_STORE output
    size_t ldo;
    float *partial;       // SUMSQ: Mp x ntiles
    int ntiles;
};

struct AcceptArgs {
    double *theta, *prior, *data, *post;
    const double *ctheta, *cprior, *cdata;
    double *cpost;
    const int *mask;
    double beta;
    int C, P;
    uint64_t seed;
    uint32_t stage, step;
    const double *dice;
    int *decision;
    unsigned long long *counters;
};

// sources.hip
hipError_t launch_mogi(const MogiArgs &a, int mode, int fp64, hipStream_t st);
hipError_t launch_cdm_prepare(const CdmArgs &a, hipStream_t st);
hipError_t launch_cdm(const CdmArgs &a, int mode, hipStream_t st);
hipError_t launch_cdm_main(const CdmArgs &a, int mode, hipStream_t st);

// gemm.hip
hipError_t launch_gemm_f32(const float *A, size_t lda, const float *B, size_t ldb, int Mp, int Np, int Kp,
                           int tri, int epi, const GemmEpi &ep, hipStream_t st);
hipError_t launch_gemm_bf16x3(const void *Ahi, const void *Alo, size_t lda, const void *Bhi, const void *Blo,
                              size_t ldb, int Mp, int Np, int Kp, int tri, int plain, int epi,
                              const GemmEpi &ep, hipStream_t st);
hipError_t launch_llk_reduce(const void *partial, int fp64, int ntiles, int C, double normalization,
                             double *llk, hipStream_t st);
hipError_t launch_whiten_f64(const double *R, size_t ldr, const double *W, size_t ldw, int N, int C,
                             double normalization, double *llk, hipStream_t st);

// sampler.hip
hipError_t launch_normals(double *xi, int C, int P, uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                          hipStream_t st);
hipError_t launch_propose(const double *theta, const double *L, const double *xi, int C, int P, double *cand,
                          hipStream_t st);
hipError_t launch_prior_verify(const double *theta, int C, int P, const int *kind, const double *a,
                               const double *b, int *mask, hipStream_t st);
hipError_t launch_prior_llk(const double *theta, int C, int P, const int *kind, const double *a, const double *b,
                            double *llk, hipStream_t st);
hipError_t launch_prior_sample(double *theta, int C, int P, const int *kind, const double *a, const double *b,
                               uint64_t seed, uint32_t stage, uint32_t attempt, const int *only, hipStream_t st);
hipError_t launch_restore(double *cand, const double *theta, const int *mask, int C, int P, hipStream_t st);
hipError_t launch_accept(const AcceptArgs &a, hipStream_t st);
hipError_t launch_posterior(const double *prior, const double *data, double beta, int C, double *post,
                            hipStream_t st);
hipError_t launch_wcov(const double *theta, const double *w, int C, int P, double *mean, double *sigma,
                       hipStream_t st);
hipError_t launch_gather(const double *theta_old, const double *prior_old, const double *data_old,
                         const double *post_old, const int *src, int C, int P, double *theta, double *prior,
                         double *data, double *post, hipStream_t st);
hipError_t launch_cdm_verify(const double *theta, int C, int P, CdmLayout L, int *mask, hipStream_t st);

}  // namespace altar
