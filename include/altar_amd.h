/*
 * altar_amd — C ABI of the MI355X (gfx950) CATMIP inner loop.
 *
 * Plain pointers and sizes only.  "dev" pointers are device memory on the
 * current HIP device; "host" pointers are host memory.  `stream` is a
 * hipStream_t passed as void* (NULL = the null stream).  Every entry point
 * returns 0 on success and a nonzero ALTAR_E* code on failure, with a
 * message available from altar_last_error() (thread local).  No C++
 * exception crosses this boundary.  Handles are not thread safe: one host
 * thread per handle.  Calls are asynchronous on `stream` unless stated.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   models/mogi/ext/mogi/source.cc:44-571       libmogi newSource/data/locations/los/oid/layout/
 *                                                displacements/residuals (CPython capsules)
 *   models/mogi/ext/cudamogi/source.cc:54-488   libcudamogi (same verbs, fused residuals)
 *   models/cdm/ext/cdm/source.cc:44-571         libcdm (same verbs, CDM layout)
 *   models/cdm/ext/cudacdm/source.cc            libcudacdm
 *   models/linear/linear/Linear.py:151-185      Linear.dataLikelihood (dgemm + L2 loop)
 *   altar/altar/norms/L2.py:39-52               L2.withCovariance (dtrmv + dnrm2), per sample
 *   altar/altar/bayesian/Metropolis.py:93-281   prepareSamplingPDF/displace/walkChains accept loop
 *   altar/altar/distributions/Uniform.py:44-72  Uniform.verify
 *   altar/altar/distributions/Base.py:47-79     initializeSample / priorLikelihood
 *   altar/ext/dbeta.cc:39-199                   libaltar cov / dbeta_grid (COV.cc:222-368)
 *   altar/altar/bayesian/COV.py:105-251         computeCovariance / rank
 */
#ifndef ALTAR_AMD_H
#define ALTAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ALTAR_OK 0
#define ALTAR_EINVAL 1   /* bad argument (the reference raises TypeError, ext/mogi/source.cc:95-107) */
#define ALTAR_EHIP 2     /* HIP runtime failure */
#define ALTAR_ENOMEM 3   /* allocation failure */
#define ALTAR_ESTATE 4   /* call out of order (e.g. llk before observations) */

/* forward models */
#define ALTAR_MOGI 0
#define ALTAR_CDM 1
#define ALTAR_LINEAR 2
#define ALTAR_JOINT 3   /* Mogi + CDM sources superposed (no reference counterpart) */

/* whitening operators (|W r|^2 with W lower triangular) */
#define ALTAR_WHITEN_DIAG 1
#define ALTAR_WHITEN_LOWER 2
#define ALTAR_WHITEN_BLOCKS 3   /* block-diagonal, one lower-triangular block per dataset */

/* arithmetic of the forward model + whitening */
#define ALTAR_FP64 0          /* FP64 everywhere (small problems, error studies) */
#define ALTAR_FP32 1          /* FP32 forward, FP32 MFMA whitening */
#define ALTAR_BF16 2          /* FP32 forward, BF16 MFMA whitening (FP32 accumulate) */
#define ALTAR_BF16_REFINED 3  /* as ALTAR_BF16 plus one split-precision correction step */

typedef struct altar_source altar_source;
typedef struct altar_prior altar_prior;

const char *altar_last_error(void);
int altar_version(void);

/* ---- sources: one handle per forward model (replaces newSource, ext/mogi/source.cc:44-63) */

/* model = ALTAR_MOGI | ALTAR_CDM | ALTAR_LINEAR | ALTAR_JOINT; P parameters,
 * N observations, nu the Poisson ratio (ignored for ALTAR_LINEAR).
 * ALTAR_JOINT has no working reference counterpart (altar/altar/models/
 * Ensemble.py:17-98 cannot run): the prediction is the sum of a Mogi source
 * and a CDM source at every observation, minus the dataset shift. */
int altar_source_create(altar_source **out, int model, int P, int N, double nu);
int altar_source_destroy(altar_source *h);

/* Parameter layout (replaces layout(), ext/mogi/source.cc:373 "O!kkkk" and
 * ext/cdm/source.cc:373 "O!kkkkkk"):
 *   Mogi: idx = {xIdx, dIdx, sIdx, offsetIdx, nsets}           (y = xIdx + 1; s = log10 dV)
 *   CDM:  idx = {xIdx, dIdx, openingIdx, aXIdx, omegaXIdx, offsetIdx, nsets}
 *   joint: idx = {Mogi xIdx, dIdx, sIdx, CDM xIdx, dIdx, openingIdx, aXIdx, omegaXIdx,
 *                 offsetIdx, nsets}
 * offsetIdx = -1 with nsets = 0 means no per-dataset shifts; nsets <= 8. */
int altar_source_layout(altar_source *h, const int *idx, int nidx);

/* Observations, host FP64 (replaces locations/los/oid/data):
 * x, y: N; los: N x 3 unit vectors; d: N observed displacements; oid: N
 * dataset ids in [0, nsets).  Copied to the device (synchronous). */
int altar_source_observations(altar_source *h, const double *x, const double *y, const double *los,
                              const double *d, const int *oid);

/* Linear model operator, host FP64: G is N x P row-major, d is N
 * (replaces Linear.loadInputs/initializeResiduals, Linear.py:231-343).
 * prewhitened = 1 declares that G and d were already multiplied by the
 * whitener W (so |G theta - d|^2 is the whitened norm and no whitening
 * operator is set). */
int altar_source_green(altar_source *h, const double *G, const double *d, int prewhitened);

/* Whitening operator and normalization, host FP64 (replaces cd_inv and
 * normalization, models/mogi/mogi/Mogi.py:327-356):
 * kind = ALTAR_WHITEN_DIAG: W is N values; kind = ALTAR_WHITEN_LOWER: W is
 * N x N row-major lower triangular (entries above the diagonal ignored).
 * llk = normalization - |W r|^2 / 2. */
int altar_source_whitening(altar_source *h, int kind, const double *W, double normalization);

/* Block-diagonal whitening (replaces a per-dataset cd_inv; the reference
 * whitens one dataset per model, Mogi.py:327-356): block b covers the
 * observations [offsets[b], offsets[b] + sizes[b]), offsets multiples of 128,
 * blocks ordered and disjoint; W holds the blocks' lower-triangular
 * whiteners one after the other, each sizes[b] x sizes[b] row-major.
 * Observations outside every block do not enter the llk (padding).
 * llk = normalization - sum_b |W_b r_b|^2 / 2.  FP32 and BF16 precisions. */
int altar_source_whitening_blocks(altar_source *h, int nblocks, const int *offsets, const int *sizes,
                                  const double *W, double normalization);

/* forwardModelBatched: predicted LOS displacements minus the dataset shift
 * (Source::displacements, libmogi/Source.cc:65-118), FP64 out, C x ldo rows.
 * precision ALTAR_FP64 or ALTAR_FP32 (the arithmetic used). */
int altar_source_forward(altar_source *h, const double *theta_dev, int C, double *pred_dev, size_t ldo,
                         int precision, void *stream);

/* residuals = forward - d (Source::residuals, libmogi/Source.cc:121-154;
 * libcudamogi fuses both, libcudamogi/Source.cc:458-474), FP64 out. */
int altar_source_residuals(altar_source *h, const double *theta_dev, int C, double *res_dev, size_t ldr,
                           int precision, void *stream);

/* Data log-likelihood of C samples (theta_dev: C x P FP64 row-major):
 * llk_dev[c] = normalization - |W (forward(theta_c) - d)|^2 / 2
 * (Native/Fast strategies, models/mogi/mogi/Fast.py:60-94). */
int altar_source_data_llk(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision,
                          void *stream);

/* As altar_source_data_llk, and also stores the whitened residuals W r
 * (C x ldw FP32, ldw >= N rounded up to 128, C a multiple of 128): the
 * training fields of the surrogate, produced by the same evaluation.
 * Any whitener: diagonal (w r per element, FP32), lower triangular, the
 * pre-whitened residual of a linear source, or a block whitener whose blocks'
 * padded ranges tile [0, N rounded up to 128) (block b's columns hold
 * W_b r_b, its padding columns zero).  Precision FP32 or BF16 (not FP64). */
int altar_source_data_llk_fields(altar_source *h, const double *theta_dev, int C, double *llk_dev, int precision,
                                 float *wr_dev, size_t ldw, void *stream);

/* Model-specific constraints: mask_dev[c] += 1 for invalid samples (CDM: a
 * source above the free surface, where libcdm.py:219-224 raises).  No-op
 * for Mogi and linear. */
int altar_source_verify(altar_source *h, const double *theta_dev, int C, int *mask_dev, void *stream);

/* CDM precision policy (no reference counterpart: the reference evaluates
 * every side in FP64, models/cdm/cdm/libcdm.py:375-430).  Sides whose
 * |sin(beta)| lies below threshold (default 0.2) are evaluated in FP64, the
 * rest in FP32; 0 turns the FP64 sides off and any value above 1 (at most
 * 2) evaluates every side in FP64. */
int altar_source_fp64_sides(altar_source *h, double threshold);
/* Census of the sides of the C chains prepared by the last CDM call:
 * counts[0..3] = FP32 sides, FP64 sides, skipped (degenerate or switched-off
 * rectangle) sides of the valid chains, and chains rejected as invalid. */
int altar_source_side_census(altar_source *h, int C, long long *counts, void *stream);

/* Per-kernel timing with HIP events recorded on the launch stream (no
 * reference counterpart; the reference's Profiler.py times wall clock).
 * Slots: 0 forward-model kernel, 1 whitening / linear GEMM.  altar_source_timers
 * waits for the recorded events, returns the summed milliseconds and launch
 * counts per slot, and clears them. */
int altar_source_timing(altar_source *h, int enable);
int altar_source_timers(altar_source *h, double *ms, long long *count, int nslots);

/* ---- priors (Contiguous psets of Uniform / Gaussian, altar/altar/models/Contiguous.py) */

/* kind[j] = 0: uniform on [a[j], b[j]]; kind[j] = 1: gaussian, mean a[j], sigma b[j] */
int altar_prior_create(altar_prior **out, int P, const int *kind, const double *a, const double *b);
int altar_prior_destroy(altar_prior *h);
/* initializeSample (Base.py:47-57): theta_dev[c, :] drawn from the prep
 * distribution with the Philox stream (seed, stage, attempt); only rows
 * with only_dev[c] != 0 when only_dev is non-NULL */
int altar_prior_sample(altar_prior *h, double *theta_dev, int C, uint64_t seed, uint32_t stage,
                       uint32_t attempt, const int *only_dev, void *stream);
/* priorLikelihood (Base.py:60-79): llk_dev[c] = sum_j log pdf_j(theta[c, j]) */
int altar_prior_llk(altar_prior *h, const double *theta_dev, int C, double *llk_dev, void *stream);
/* verify (Uniform.py:44-72): mask_dev[c] += 1 if a parameter leaves its support */
int altar_prior_verify(altar_prior *h, const double *theta_dev, int C, int *mask_dev, void *stream);

/* ---- Metropolis (Metropolis.py:114-259) */

/* candidate = theta + (L xi)^T with xi ~ N(0, 1) from the Philox stream
 * (seed, stage, step); L_dev is P x P lower triangular (the Cholesky factor
 * of scaling^2 Sigma).  xi_dev (C x P) receives the normals (workspace). */
int altar_mh_propose(const double *theta_dev, int C, int P, const double *L_dev, uint64_t seed,
                     uint32_t stage, uint32_t step, double *xi_dev, double *cand_dev, void *stream);
/* rows with mask_dev[c] != 0 revert to theta (Metropolis.py:177-182) */
int altar_mh_restore(double *cand_dev, const double *theta_dev, const int *mask_dev, int C, int P,
                     void *stream);
/* Accept/reject (Metropolis.py:189-223) with the posterior fused
 * (cpost = cprior + beta cdata, Bayesian.py:104-115).  Uniforms from the
 * Philox stream (seed, stage, step) unless dice_dev (C values in (0,1]) is
 * given.  decision_dev (optional): 0 accepted, 1 rejected, 2 unlikely.
 * counters_dev[0..2] += (accepted, rejected, unlikely).  cpost_dev optional. */
int altar_mh_accept(double *theta_dev, double *prior_dev, double *data_dev, double *post_dev,
                    const double *ctheta_dev, const double *cprior_dev, const double *cdata_dev,
                    double *cpost_dev, const int *mask_dev, double beta, int C, int P, uint64_t seed,
                    uint32_t stage, uint32_t step, const double *dice_dev, int *decision_dev,
                    unsigned long long *counters_dev, void *stream);
/* Metropolis.adjustCovarianceScaling (Metropolis.py:262-281) */
double altar_mh_scaling(double acceptance_weight, double rejection_weight, unsigned long long accepted,
                        unsigned long long rejected, unsigned long long unlikely);

/* ---- COV scheduler (COV.cc:222-368, COV.py:65-251) */

/* dbeta grid search on host FP64 data (libaltar dbeta_grid): given llk
 * (C), its median and the current beta, writes the normalized weights w
 * (C) and returns the new beta and the attained COV. */
int altar_cov_dbeta_grid(const double *llk_host, int C, double median, double beta, double tolerance,
                         double target, double *w_host, double *beta_out, double *cov_out);
/* dbeta by Brent's method on host FP64 data (libaltar dbeta_brent,
 * COV.cc:58-219, Brent.py:19-63; the reference uses GSL's Brent minimizer):
 * minimizes (COV - target)^2 over [0, 1 - beta] from 5e-5, stopping when the
 * last COV is within tolerance of the target or after maxiter iterations
 * (the reference's default 1000); the skip to beta = 1 as in the grid solver.
 * Writes w (C) and returns the new beta and the attained COV. */
int altar_cov_dbeta_brent(const double *llk_host, int C, double median, double beta, double tolerance,
                          double target, int maxiter, double *w_host, double *beta_out, double *cov_out);
/* weighted mean (P) and covariance Sigma (P x P) of theta_dev with weights w_dev */
int altar_cov_covariance(const double *theta_dev, const double *w_dev, int C, int P, double *mean_dev,
                         double *sigma_dev, void *stream);
/* multinomial multiplicities (COV.py:222-251): C uniforms from the Philox
 * stream (seed, stage) binned on the edges {0, cumsum(w)} (last edge forced
 * to 1); count_host (C).  Host, synchronous. */
int altar_cov_multiplicities(const double *w_host, int C, uint64_t seed, uint32_t stage, int *count_host);
/* The same COV update with the data llk resident on the device: median,
 * grid search and weights in one workgroup (fixed-order FP64 reductions);
 * out_dev[0] = new beta, out_dev[1] = COV, out_dev[2] = median. */
int altar_cov_update_dev(const double *llk_dev, int C, double beta, double tolerance, double target, double *w_dev,
                         double *out_dev, void *stream);
/* multinomial resampling on the device with the same Philox stream as
 * altar_cov_multiplicities: count_dev (C) the multiplicities, src_dev (C) the
 * old row of each new row (original-sample order); edges_dev (C + 1
 * doubles) and offset_dev (C ints) are scratch */
int altar_cov_resample_dev(const double *w_dev, int C, uint64_t seed, uint32_t stage, double *edges_dev,
                           int *count_dev, int *offset_dev, int *src_dev, void *stream);
/* As altar_cov_resample_dev, with the scratch (bin edges and count offsets)
 * in a fixed device array of the library instead of caller buffers.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid
 * hipMalloc).  It calls no allocation routine.  C <= ALTAR_NOALLOC_MAX_CHAINS
 * (larger C: ALTAR_EINVAL); the array holds 3 MiB per device for the life of
 * the process.  All calls on one device share the array: calls that overlap
 * on one device (different streams or host threads) race. */
#define ALTAR_NOALLOC_MAX_CHAINS 262144
int altar_cov_resample_dev_noalloc(const double *w_dev, int C, uint64_t seed, uint32_t stage, int *count_dev,
                                   int *src_dev, void *stream);
/* resampled state: row c <- old row src_dev[c] (theta, prior, data, posterior) */
int altar_resample_gather(const double *theta_old, const double *prior_old, const double *data_old,
                          const double *post_old, const int *src_dev, int C, int P, double *theta,
                          double *prior, double *data, double *post, void *stream);
/* post = prior + beta data (CoolingStep.computePosterior) */
int altar_posterior(const double *prior_dev, const double *data_dev, double beta, int C, double *post_dev,
                    void *stream);

/* ---- dense algebra helpers (FP32 MFMA GEMM, row gather) */

/* out = A B^T - bias (per column, optional): A M x K, B N x K, row-major;
 * M, N multiples of 128, K a multiple of 16 */
int altar_gemm_f32(const float *A_dev, size_t lda, const float *B_dev, size_t ldb, int M, int N, int K,
                   const float *bias_dev, float *out_dev, size_t ldo, void *stream);
/* dst[c, r] = src[r, c] for r < rows, c < cols (FP32, row pitches lds, ldd) */
int altar_transpose_f32(const float *src_dev, size_t lds, int rows, int cols, float *dst_dev, size_t ldd,
                        void *stream);
/* dst[r, 0:cols] = src[idx[r], 0:cols], r < n */
int altar_gather_rows_f32(const float *src_dev, size_t lds, const int *idx_dev, int n, int cols, float *dst_dev,
                          size_t ldd, void *stream);

/* ---- learned surrogate (no reference counterpart: the delayed-acceptance
 * screen; an addition to CATMIP).  A dense network nin -> nhid x
 * (nlayers - 1) -> nout with GELU hidden activations, BF16 weights with FP32
 * master copies and Adam, FP32 accumulation on the matrix cores.  Inputs are
 * parameter vectors scaled to [-1, 1] by altar_mlp_input_scaling. */
typedef struct altar_mlp altar_mlp;
int altar_mlp_create(altar_mlp **out, int nin, int nhid, int nlayers, int nout, int max_batch);
int altar_mlp_destroy(altar_mlp *h);
int altar_mlp_input_scaling(altar_mlp *h, const double *lo_host, const double *hi_host);
/* layer l: W (out x in, row-major) and b (out), FP32 host; resets Adam state */
int altar_mlp_set_layer(altar_mlp *h, int l, const float *W_host, const float *b_host);
/* read layer l's master weights, or (grads != 0) the last gradients; synchronizes the device */
int altar_mlp_get_layer(altar_mlp *h, int l, float *W_host, float *b_host, int grads);
/* z (B x ldz FP32) = network(theta (B x nin FP64)) */
int altar_mlp_forward(altar_mlp *h, const double *theta_dev, int B, float *z_dev, int ldz, void *stream);
/* gradients of sum_b sum_k w_k (net_k - target_bk)^2 / B (target B x ldt FP32, w nout FP32) */
int altar_mlp_backward(altar_mlp *h, const double *theta_dev, const float *target_dev, int ldt, const float *w_dev,
                       int B, float *loss_dev, void *stream);
/* one Adam step (Kingma and Ba 2015) with the last gradients */
int altar_mlp_adam(altar_mlp *h, float lr, float beta1, float beta2, float eps, void *stream);
/* one epoch in one call: batch b (b < nbatches) takes the rows
 * order[b batch .. (b + 1) batch) of theta_store (FP64, nin columns) and of
 * targets (FP32, row pitch ldt >= nout), runs the gradients of
 * altar_mlp_backward and one Adam step with the cosine-decayed rate
 * lr_final + (lr - lr_final)(1 + cos(pi it / (total - 1))) / 2, it = it0 + b;
 * loss_dev (optional) receives the last batch's loss */
int altar_mlp_train_epoch(altar_mlp *h, const double *theta_store_dev, const float *targets_dev, int ldt,
                          const float *w_dev, const int *order_dev, int nbatches, int batch, int it0, int total,
                          float lr, float lr_final, float beta1, float beta2, float eps, float *loss_dev, void *stream);
/* surrogate data llk from normalized PCA coefficients z (C x ldz):
 * llk = normalization - (cperp + sum_k (e_k + s_k z_k)^2) / 2; e, s device FP64 (K) */
int altar_surrogate_llk(const float *z_dev, int ldz, int C, int K, const double *e_dev, const double *s_dev,
                        double cperp, double normalization, double *llk_dev, void *stream);

/* ---- two-stage delayed-acceptance Metropolis (Christen and Fox 2005) */
/* stage 1: cpost_s = cprior + beta cdata_s; chain c passes if mask[c] == 0 and
 * log u1 <= cpost_s - post_s (u1: screening stream).  index_dev (2 C ints)
 * receives the passing chains in chain order, count_dev their number;
 * counters_dev[1] += rejected, counters_dev[3] += screened out. */
int altar_da_screen(const double *cprior_dev, const double *cdata_s_dev, const double *post_s_dev,
                    const int *mask_dev, double beta, int C, uint64_t seed, uint32_t stage, uint32_t step,
                    double *cpost_s_dev, int *index_dev, int *count_dev, unsigned long long *counters_dev,
                    void *stream);
/* packed[j] = cand[index[j]] for j < n (FP64 rows of P) */
int altar_da_pack(const double *cand_dev, const int *index_dev, int n, int P, double *packed_dev, void *stream);
/* stage 2 for the n screened-in chains with their exact data llk (packed):
 * accept if log u2 <= (cpost - post) - (cpost_s - post_s); on acceptance the
 * state, post and post_s take the candidate's values.  counters_dev[0] +=
 * accepted, counters_dev[2] += unlikely at stage 2. */
int altar_da_accept(double *theta_dev, double *prior_dev, double *data_dev, double *post_dev, double *post_s_dev,
                    const double *cand_dev, const double *cprior_dev, const double *cdata_packed_dev,
                    const double *cpost_s_dev, const int *index_dev, int n, double beta, int P, uint64_t seed,
                    uint32_t stage, uint32_t step, int *decision_dev, unsigned long long *counters_dev,
                    void *stream);

/* ---- surrogate trajectories (Liu 2001, the surrogate transition method; no
 * reference counterpart).  Each exact step runs `substeps` Metropolis steps of
 * the surrogate posterior pi* = prior + beta llk_s from the chain's state theta
 * to an end point phi, then altar_da_accept judges phi with the exact model:
 * accept if log u2 <= (post(phi) - post(theta)) - (post_s(phi) - post_s(theta)).
 * The ratio keeps the exact posterior invariant (a Metropolis kernel is
 * reversible with respect to its target, so are its powers).
 * Sub-step streams: step' = step * 1024 + sub (sub < 1024, step < 2^22). */
/* cand = phi + (L xi)^T with xi from the trajectory proposal stream */
int altar_da_trajectory_propose(const double *phi_dev, int C, int P, const double *L_dev, uint64_t seed,
                                uint32_t stage, uint32_t step, uint32_t sub, double *xi_dev, double *cand_dev,
                                void *stream);
/* one surrogate Metropolis sub-step on (phi, prior_phi, post_phi):
 * cpost_s = cprior + beta cdata_s; accept cand if mask[c] == 0 and
 * log u <= cpost_s - post_phi (u: trajectory dice stream).  The upper half of
 * index_dev (2 C ints) flags the chains that moved in this trajectory (sub 0
 * resets the flags).  counters_dev[0] += accepted, [1] += rejected by the
 * model, [2] += unlikely. */
int altar_da_trajectory_step(double *phi_dev, double *prior_phi_dev, double *post_phi_dev, const double *cand_dev,
                             const double *cprior_dev, const double *cdata_s_dev, double *cpost_s_dev,
                             const int *mask_dev, double beta, int C, int P, uint64_t seed, uint32_t stage,
                             uint32_t step, uint32_t sub, int *index_dev, unsigned long long *counters_dev,
                             void *stream);
/* The whole trajectory of every chain in one launch (a workgroup runs 64
 * chains through all sub-steps with their state and the network's
 * activations in LDS).  Sub-step for sub-step the same arithmetic as the
 * separate calls: altar_da_trajectory_propose, the prior's support check,
 * the source's check (CDM and joint sources; src may be NULL), the prior
 * llk, the network (altar_mlp_forward), altar_surrogate_llk and
 * altar_da_trajectory_step, so the end points agree bit for bit.
 * Also writes post_s_dev[c] = prior_theta[c] + beta llk_s(theta_c), the
 * surrogate posterior of the starting state for the exact stage.
 * Needs P <= 32 and a network of hidden width 256 with at most 64 inputs and
 * outputs.  Moved flags and counters as altar_da_trajectory_step. */
int altar_da_trajectory_fused(altar_mlp *mlp, const altar_prior *prior, const altar_source *src,
                              const double *theta_dev, const double *prior_theta_dev, double *post_s_dev,
                              double *phi_dev, double *prior_phi_dev, double *post_phi_dev, const double *L_dev,
                              const double *e_dev, const double *s_dev, int K, double cperp, double normalization,
                              double beta, int C, int P, uint64_t seed, uint32_t stage, uint32_t step,
                              int substeps, int *index_dev, unsigned long long *counters_dev, void *stream);
/* index_dev[0..count) = the chains whose trajectory moved, in chain order */
int altar_da_trajectory_list(int *index_dev, int C, int *count_dev, void *stream);

/* Philox normals into device memory (the same streams the kernels draw),
 * for parity checks */
int altar_random_normals(double *out_dev, int C, int P, uint64_t seed, uint32_t stage, uint32_t step,
                         uint32_t tag, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ALTAR_AMD_H */
