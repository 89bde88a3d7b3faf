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
// CDM only: add the predicted LOS displacement (no shift, no data) to an
// existing residual matrix, FP32 or FP64 (the second source of a joint model)
enum { MODE_ACC_F32 = 4, MODE_ACC_F64 = 5 };
// GEMM epilogues
enum { EPI_STORE = 0, EPI_SUMSQ = 1, EPI_SUMSQ_BIAS = 2, EPI_STORE_SUMSQ = 3 };

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
    int nvalid;           // columns from nvalid on are zero padding of B; their MFMAs are
                          // skipped (0: every column is computed)
};

// surrogate MLP GEMM epilogues
enum { MLP_EPI_BIAS_ACT = 0, MLP_EPI_F32 = 1, MLP_EPI_DGELU = 2 };

struct MlpEpi {
    const float *bias;
    float *pre;           // BIAS_ACT: pre-activation out; DGELU: pre-activation in
    float *out32;
    __bf16 *out16;
    int ldo;
    int act;              // BIAS_ACT: apply GELU
    size_t split_stride;  // F32 with split-k: slice z is written at out32 + z * split_stride
    __bf16 *out16T;       // BIAS_ACT, DGELU (optional): the transpose of out16, row pitch ldT
    int ldT;
    float *cpart;         // DGELU (optional): column-sum partials of out16 per 64-row slab, pitch ldo
};

// one layer's weight gradient in mlp_wgrad_kernel: the split-k NT product
// A (M x K) B^T (N x K) into part (S slices of M x N); blocks from first on
struct MlpWGradSeg {
    const __bf16 *A, *B;
    float *part;
    int M, N, first;
};
// one ordered sum of mlp_reduce_kernel: dst[i] = sum over c < count of
// src[c stride + i], i < n; blocks from first on
struct MlpReduceSeg {
    const float *src;
    float *dst;
    long n;
    size_t stride;
    int count, first;
};

// one tensor of an Adam step (adam_kernel): w, m, v, g are rows x cols FP32;
// w16 (rows x cols) and wt16 (cols x rows) are optional BF16 copies of w; the
// segment owns the launch's blocks from first on
struct AdamSeg {
    float *w, *m, *v;
    const float *g;
    __bf16 *w16, *wt16;
    int rows, cols, first;
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
// partial[m, t] (FP64, Mp x ntiles) = |row m of A B^T over column tile t|^2,
// all in FP64; operands zero-padded like the FP32 path's
hipError_t launch_gemm_f64_sumsq(const double *A, size_t lda, const double *B, size_t ldb, int Mp, int Np, int Kp,
                                 int tri, int nvalid, double *partial, int ntiles, hipStream_t st);
hipError_t launch_gather_rows_f32(const float *src, size_t lds, const int *idx, int n, int cols, float *dst,
                                  size_t ldd, hipStream_t st);
hipError_t launch_transpose_f32(const float *src, size_t lds, int rows, int cols, float *dst, size_t ldd,
                                hipStream_t st);
hipError_t launch_llk_reduce(const void *partial, int fp64, int ntiles, int C, double normalization,
                             double *llk, hipStream_t st);

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
// mlp.hip
hipError_t launch_mlp_gemm(int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                           const MlpEpi &ep, hipStream_t st);
// dY of the weighted squared error (rows x cols tiles of 64), its transpose
// dYT (cols x ldT), the 64-row slab partials of its column sums (part,
// colsum_slabs(rows) x cols) and, if loss is set, the loss added to *loss
hipError_t launch_mse_grad(const float *yhat, const float *y, const float *w, int rows, int cols, int ld,
                           int valid_rows, void *dy, void *dyT, int ldT, float *part, float *loss, hipStream_t st);
// 64-row slabs of a batch of rows (the column-sum partials per layer)
int colsum_slabs(int rows);
// every layer's weight gradient (segment table in device memory; K the
// batch rows, the row pitch of both operands, split into S slices)
hipError_t launch_mlp_wgrad(const MlpWGradSeg *segs_dev, int nseg, int blocks, int K, int S, int ld, hipStream_t st);
// several ordered sums in one launch (segment table in device memory)
hipError_t launch_mlp_reduce(const MlpReduceSeg *segs_dev, int nseg, int blocks, hipStream_t st);
// one Adam step over nseg tensors (table in device memory) in one launch of blocks blocks
hipError_t launch_adam(const AdamSeg *segs_dev, int nseg, int blocks, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t st);
hipError_t launch_surrogate_llk(const float *z, int ldz, int C, int K, const double *e, const double *s,
                                double cperp, double norm, double *llk, hipStream_t st);
// network input of C rows into x (rows x ldx, rows >= C, zero padded) and,
// if xT is set, its transpose (ldx x ldT)
hipError_t launch_mlp_input(const double *theta, int C, int P, const double *lo, const double *hi, void *x, int ldx,
                            int rows, void *xT, int ldT, hipStream_t st);

// delayed acceptance (sampler.hip)
hipError_t launch_da_screen(const double *cprior, const double *cdata_s, const double *post_s, const int *mask,
                            double beta, int C, uint64_t seed, uint32_t stage, uint32_t step, double *cpost_s,
                            int *index, int *count, unsigned long long *counters, hipStream_t st);
hipError_t launch_da_pack(const double *cand, const int *index, int n, int P, double *packed, hipStream_t st);
hipError_t launch_da_accept(double *theta, double *prior, double *data, double *post, double *post_s,
                            const double *cand, const double *cprior, const double *cdata_packed,
                            const double *cpost_s, const int *index, int n, double beta, int P, uint64_t seed,
                            uint32_t stage, uint32_t step, int *decision, unsigned long long *counters,
                            hipStream_t st);
// surrogate trajectory (sampler.hip); step2 = step * 1024 + sub-step
hipError_t launch_da_traj_propose(const double *phi, int C, int P, const double *L, uint64_t seed, uint32_t stage,
                                  uint32_t step2, double *xi, double *cand, hipStream_t st);
hipError_t launch_da_traj_step(double *phi, double *prior_phi, double *post_phi, const double *cand,
                               const double *cprior, const double *cdata_s, double *cpost_s, const int *mask,
                               double beta, int C, int P, uint64_t seed, uint32_t stage, uint32_t step2, int first,
                               int *index, unsigned long long *counters, hipStream_t st);
hipError_t launch_da_traj_list(int *index, int C, int *count, hipStream_t st);

// the whole surrogate trajectory in one launch (trajectory.hip)
constexpr int TRAJ_PMAX = 32;        // parameters per sample
constexpr int TRAJ_MAX_LAYERS = 8;
struct TrajArgs {
    const double *theta, *prior_theta;   // C x P state and its prior llk
    double *post_s;                      // out: surrogate posterior of the state
    double *phi, *prior_phi, *post_phi;  // out: trajectory end points
    const double *L;                     // P x P lower (proposal)
    const int *kind;                     // prior: kind, a, b per parameter
    const double *pa, *pb;
    int has_cdm;                         // model check: CDM source above the free surface
    CdmLayout cdm;
    const __bf16 *W[TRAJ_MAX_LAYERS];    // layer l: dout x din (din 64 first, 256 after; dout 256, 64 last)
    const float *bias[TRAJ_MAX_LAYERS];
    int nlayers;
    const double *lo, *hi;               // network input scaling
    const double *e, *s;                 // surrogate llk: norm - (cperp + sum_k (e_k + s_k z_k)^2) / 2
    double cperp, norm;
    int K;
    double beta;
    int C, P;
    uint64_t seed;
    uint32_t stage, step2;               // sub-step j draws at step2 + j
    int substeps;
    int *moved;                          // C: 1 if any sub-step moved the chain
    unsigned long long *counters;        // [0] accepted, [1] rejected by the model, [2] unlikely
};
hipError_t launch_traj_fused(const TrajArgs &a, hipStream_t st);

// cov.hip
hipError_t launch_cov_update(const double *llk, int C, double beta, double tolerance, double target, double *w,
                             double *out, hipStream_t st);
hipError_t launch_cov_resample_src(const double *w, int C, uint64_t seed, uint32_t stage, double *edges, int *count,
                                   int *offset, int *src, hipStream_t st);

hipError_t launch_cdm_verify(const double *theta, int C, int P, CdmLayout L, int *mask, hipStream_t st);
// out[0..3] = FP32 sides, FP64 sides, skipped sides of the valid chains, invalid chains
hipError_t launch_side_census(const unsigned *flags, int C, unsigned long long *out, hipStream_t st);

}  // namespace altar
