// Metropolis propose/verify/accept, priors and the COV-stage helpers.
//
// Reference: altar/altar/bayesian/Metropolis.py:93-281 (propose with the
// Cholesky factor of scaling^2 Sigma, verify against the prior support,
// accept if log u <= post' - post), altar/altar/distributions/Uniform.py
// :44-72 and Base.py:60-79 (prior), altar/altar/bayesian/COV.py:105-251
// (weighted covariance, resampling).
//
// All sample state is FP64 and row-major (chains x parameters).  Random
// draws come from the Philox streams in common.hpp: normals by Box-Muller,
// uniforms on (0, 1] with 53 bits.
#include "common.hpp"
#include "kernels.hpp"
#include "chain_math.hpp"

// The state updates (posterior = prior + beta data, the accept test) follow
// the reference's separate multiply and add; no fused multiply-add
// contraction in this file.
#pragma clang fp contract(off)

namespace altar {

// ---------------------------------------------------------------- draws
// xi[c, 2w], xi[c, 2w+1] from one Philox block: one thread per (chain, pair)
__global__ void normals_kernel(double *__restrict__ xi, int C, int P, uint64_t seed, uint32_t stage,
                               uint32_t step, uint32_t tag)
{
    const int pairs = (P + 1) / 2;
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * pairs) return;
    const int c = (int)(idx / pairs), w = (int)(idx % pairs);
    u32x4 r = draw(seed, stage, step, (uint32_t)c, tag, (uint32_t)w);
    double u1 = u53(r.x, r.y), u2 = u53(r.z, r.w);
    double rad = sqrt(-2.0 * log(u1));
    double s, co;
    sincos(2.0 * PI * u2, &s, &co);
    xi[(size_t)c * P + 2 * w] = rad * co;
    if (2 * w + 1 < P) xi[(size_t)c * P + 2 * w + 1] = rad * s;
}

// candidate = theta + (L xi)^T, one thread per (chain, parameter); L lower.
// For small P (the point-source and CDM models).
__global__ void propose_kernel(const double *__restrict__ theta, const double *__restrict__ L,
                               const double *__restrict__ xi, int C, int P, double *__restrict__ cand)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    const int c = (int)(idx / P), i = (int)(idx % P);
    cand[idx] = propose_elem(theta[idx], L + (size_t)i * P, xi + (size_t)c * P, i);
}

// The same product for large P (the linear model), tiled through LDS: a
// block computes 128 chains x 64 parameters, each thread 8 consecutive
// chains x 4 consecutive parameters, over k in slices of 16 (the next slice
// is fetched into registers while the current one is multiplied).  Every
// output sums k = 0, 1, ..., i in order with fused multiply-adds, so it
// differs from the small-P kernel by rounding only; slices below the tile's
// diagonal skip the k <= i test.
constexpr int PTM = 128, PTN = 64, PK = 16;
__global__ __launch_bounds__(256) void propose_tiled_kernel(const double *__restrict__ theta,
                                                            const double *__restrict__ L,
                                                            const double *__restrict__ xi, int C, int P,
                                                            double *__restrict__ cand)
{
    // pitch PTM + 1 / PTN + 1: the 16 k of one row land in distinct banks
    __shared__ double Xs[PK][PTM + 1];
    __shared__ double Ls[PK][PTN + 1];
    const int m0 = blockIdx.x * PTM, n0 = blockIdx.y * PTN;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    double acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
    const int kmax = min(P, n0 + PTN);  // L is lower triangular: k <= n < n0 + PTN
    // staging: thread t moves (row r, k) = (t / 16 + 16 q, t % 16), q < 8 for
    // xi and q < 4 for L: 16 consecutive k of a row are one 128-byte line
    const int sk = threadIdx.x & 15, sr = threadIdx.x >> 4;
    double rx[8], rl[4];
    auto fetch = [&](int k0) {
        const int k = k0 + sk;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = m0 + sr + 16 * q;
            rx[q] = (m < C && k < P) ? xi[(size_t)m * P + k] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int n = n0 + sr + 16 * q;
            rl[q] = (n < P && k < P) ? L[(size_t)n * P + k] : 0.0;
        }
    };
    fetch(0);
    for (int k0 = 0; k0 < kmax; k0 += PK) {
#pragma unroll
        for (int q = 0; q < 8; ++q) Xs[sk][sr + 16 * q] = rx[q];
#pragma unroll
        for (int q = 0; q < 4; ++q) Ls[sk][sr + 16 * q] = rl[q];
        __syncthreads();
        if (k0 + PK < kmax) fetch(k0 + PK);
        const int kn = min(PK, kmax - k0);
        if (k0 + PK <= n0) {
            // the whole slice lies at or below the diagonal of every column
            for (int kk = 0; kk < kn; ++kk) {
                double x[8], l[4];
#pragma unroll
                for (int i = 0; i < 8; ++i) x[i] = Xs[kk][ty * 8 + i];
#pragma unroll
                for (int j = 0; j < 4; ++j) l[j] = Ls[kk][tx * 4 + j];
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = fma(l[j], x[i], acc[i][j]);
            }
        } else {
            for (int kk = 0; kk < kn; ++kk) {
                const int k = k0 + kk;
                double x[8], l[4];
#pragma unroll
                for (int i = 0; i < 8; ++i) x[i] = Xs[kk][ty * 8 + i];
#pragma unroll
                for (int j = 0; j < 4; ++j) l[j] = k <= n0 + tx * 4 + j ? Ls[kk][tx * 4 + j] : 0.0;
#pragma unroll
                for (int i = 0; i < 8; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) acc[i][j] = fma(l[j], x[i], acc[i][j]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int m = m0 + ty * 8 + i;
        if (m >= C) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + tx * 4 + j;
            if (n < P) cand[(size_t)m * P + n] = theta[(size_t)m * P + n] + acc[i][j];
        }
    }
}

// ---------------------------------------------------------------- priors
// kind 0: uniform on [a, b]; kind 1: gaussian with mean a, sigma b.
// mask[c] += 1 if any uniform parameter is outside its closed support.
__global__ void prior_verify_kernel(const double *__restrict__ theta, int C, int P,
                                    const int *__restrict__ kind, const double *__restrict__ a,
                                    const double *__restrict__ b, int *__restrict__ mask)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double *t = theta + (size_t)c * P;
    int bad = 0;
    for (int j = 0; j < P; ++j)
        if (kind[j] == 0 && !(a[j] <= t[j] && t[j] <= b[j])) bad = 1;
    if (bad) mask[c] += 1;
}

// prior log-likelihood: sum of log densities (a uniform contributes
// -log(b - a) on its support)
__global__ void prior_llk_kernel(const double *__restrict__ theta, int C, int P, const int *__restrict__ kind,
                                 const double *__restrict__ a, const double *__restrict__ b,
                                 double *__restrict__ llk)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const double *t = theta + (size_t)c * P;
    double acc = 0.0;
    for (int j = 0; j < P; ++j) {
        if (kind[j] == 0) {
            acc += log(1.0 / (b[j] - a[j]));
        } else {
            acc += gauss_log_term(t[j], a[j], b[j], log(b[j] * sqrt(2.0 * PI)));
        }
    }
    llk[c] = acc;
}

// The prior for long parameter vectors: one wavefront per chain, lanes over
// the parameters, a fixed-order butterfly sum (deterministic).
__global__ __launch_bounds__(256) void prior_llk_wave_kernel(const double *__restrict__ theta, int C, int P,
                                                             const int *__restrict__ kind,
                                                             const double *__restrict__ a,
                                                             const double *__restrict__ b, double *__restrict__ llk)
{
    const int c = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (c >= C) return;  // uniform over the wavefront
    const double *t = theta + (size_t)c * P;
    double acc = 0.0;
    for (int j = lane; j < P; j += 64) {
        if (kind[j] == 0) {
            acc += log(1.0 / (b[j] - a[j]));
        } else {
            acc += gauss_log_term(t[j], a[j], b[j], log(b[j] * sqrt(2.0 * PI)));
        }
    }
    acc = wave_sum(acc);
    if (lane == 0) llk[c] = acc;
}

__global__ __launch_bounds__(256) void prior_verify_wave_kernel(const double *__restrict__ theta, int C, int P,
                                                                const int *__restrict__ kind,
                                                                const double *__restrict__ a,
                                                                const double *__restrict__ b, int *__restrict__ mask)
{
    const int c = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6), lane = threadIdx.x & 63;
    if (c >= C) return;
    const double *t = theta + (size_t)c * P;
    int bad = 0;
    for (int j = lane; j < P; j += 64)
        if (kind[j] == 0 && !(a[j] <= t[j] && t[j] <= b[j])) bad = 1;
    if (__any(bad) && lane == 0) mask[c] += 1;
}

// initial sample from the prep distributions, one thread per (chain, param)
__global__ void prior_sample_kernel(double *__restrict__ theta, int C, int P, const int *__restrict__ kind,
                                    const double *__restrict__ a, const double *__restrict__ b,
                                    uint64_t seed, uint32_t stage, uint32_t attempt,
                                    const int *__restrict__ only)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    const int c = (int)(idx / P), j = (int)(idx % P);
    if (only && !only[c]) return;
    u32x4 r = draw(seed, stage, attempt, (uint32_t)c, TAG_INIT, (uint32_t)j);
    double u1 = u53(r.x, r.y);
    double v;
    if (kind[j] == 0) {
        v = a[j] + (b[j] - a[j]) * (1.0 - u1);  // [a, b)
    } else {
        double u2 = u53(r.z, r.w);
        v = a[j] + b[j] * sqrt(-2.0 * log(u1)) * cos(2.0 * PI * u2);
    }
    theta[idx] = v;
}

// ---------------------------------------------------------------- Metropolis
// rows of candidates rejected by verify revert to the current sample
// (Metropolis.py:177-182)
__global__ void restore_kernel(double *__restrict__ cand, const double *__restrict__ theta,
                               const int *__restrict__ mask, int C, int P)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    const int c = (int)(idx / P);
    if (mask[c]) cand[idx] = theta[idx];
}

// Metropolis.py:189-223 with posteriorLikelihood (Bayesian.py:104-115)
// fused: cpost = cprior + beta cdata.  decision: 0 accepted, 1 rejected by
// the model, 2 unlikely.  A non-finite candidate posterior is unlikely.
__global__ void accept_kernel(double *__restrict__ theta, double *__restrict__ prior, double *__restrict__ data,
                              double *__restrict__ post, const double *__restrict__ ctheta,
                              const double *__restrict__ cprior, const double *__restrict__ cdata,
                              double *__restrict__ cpost_out, const int *__restrict__ mask, double beta,
                              int C, int P, uint64_t seed, uint32_t stage, uint32_t step,
                              const double *__restrict__ dice, int *__restrict__ decision,
                              unsigned long long *__restrict__ counters)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int dec = -1;
    if (c < C) {
        const double cp = tempered(cprior[c], beta, cdata[c]);
        if (cpost_out) cpost_out[c] = cp;
        if (mask[c]) {
            dec = 1;
        } else {
            double u;
            if (dice) {
                u = dice[c];
            } else {
                u32x4 r = draw(seed, stage, step, (uint32_t)c, TAG_ACCEPT, 0u);
                u = u53(r.x, r.y);
            }
            const double diff = cp - post[c];
            if (!(log(u) <= diff)) {
                dec = 2;
            } else {
                dec = 0;
                if (P <= 32)
                    for (int j = 0; j < P; ++j) theta[(size_t)c * P + j] = ctheta[(size_t)c * P + j];
                prior[c] = cprior[c];
                data[c] = cdata[c];
                post[c] = cp;
            }
        }
        if (decision) decision[c] = dec;
    }
    if (P > 32) {
        // long rows: the block copies each accepted row with all its threads
        __shared__ int sdec[256];
        sdec[threadIdx.x] = dec;
        __syncthreads();
        const size_t c0 = (size_t)blockIdx.x * blockDim.x;
        for (int k = 0; k < (int)blockDim.x; ++k) {
            if (sdec[k] != 0) continue;
            const size_t base = (c0 + k) * P;
            for (int j = threadIdx.x; j < P; j += blockDim.x) theta[base + j] = ctheta[base + j];
        }
    }
    // block-level counts, then one atomic per counter per block
    __shared__ unsigned cnt[3];
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (dec >= 0) atomicAdd(&cnt[dec], 1u);
    __syncthreads();
    if (threadIdx.x < 3 && cnt[threadIdx.x]) atomicAdd(&counters[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

// ---------------------------------------------------------------- COV stage
// post = prior + beta data (CoolingStep.computePosterior, CoolingStep.py:103-113)
__global__ void posterior_kernel(const double *__restrict__ prior, const double *__restrict__ data, double beta,
                                 int C, double *__restrict__ post)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < C) {
        post[c] = tempered(prior[c], beta, data[c]);
    }
}

// Weighted moments of the sample (COV.py:141-152): mean[j] = sum_i w_i
// theta_ij / sum_i w_i and Sigma_ab = sum_i w_i theta_ia theta_ib -
// mean_a mean_b.  Every sum runs over the chains in order (the host
// restatement's order, so results agree bit for bit); the blocks stage
// chunks of the sample through LDS with coalesced loads and each output is
// one thread's running sum.
constexpr int WCH = 64;  // chains per LDS round

// one block per 64 parameters; thread t stages (chain t / 64 + 4 q, param
// t % 64) for q < 16, the next chunk fetched while the current one is summed
__global__ __launch_bounds__(256) void wmean_kernel(const double *__restrict__ theta, const double *__restrict__ w,
                                                    int C, int P, double *__restrict__ mean)
{
    __shared__ double st[WCH][64 + 1];
    __shared__ double sw[WCH];
    const int p0 = blockIdx.x * 64, j = p0 + (int)threadIdx.x;
    const int sq = threadIdx.x & 63, sr = threadIdx.x >> 6;
    double rt[WCH / 4], rw = 0.0;
    auto fetch = [&](int i0) {
#pragma unroll
        for (int q = 0; q < WCH / 4; ++q) {
            const int i = i0 + sr + 4 * q;
            rt[q] = (i < C && p0 + sq < P) ? theta[(size_t)i * P + p0 + sq] : 0.0;
        }
        if (threadIdx.x < WCH) rw = i0 + (int)threadIdx.x < C ? w[i0 + threadIdx.x] : 0.0;
    };
    double wsum = 0.0, acc = 0.0;
    fetch(0);
    for (int i0 = 0; i0 < C; i0 += WCH) {
#pragma unroll
        for (int q = 0; q < WCH / 4; ++q) st[sr + 4 * q][sq] = rt[q];
        if (threadIdx.x < WCH) sw[threadIdx.x] = rw;
        __syncthreads();
        if (i0 + WCH < C) fetch(i0 + WCH);
        if (threadIdx.x < 64) {
            const int n = min(WCH, C - i0);
#pragma unroll 16
            for (int r = 0; r < n; ++r) {
                wsum += sw[r];
                acc += sw[r] * st[r][threadIdx.x];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 64 && j < P) mean[j] = acc / wsum;
}

// one block per 16 x 16 tile of the lower triangle (tile row ta >= column
// tb); thread t stages (chain t / 16 + 16 q, column t % 16) of both column
// blocks for q < 4, the next chunk fetched while the current one is summed
__global__ __launch_bounds__(256) void wcov_kernel(const double *__restrict__ theta, const double *__restrict__ w,
                                                   const double *__restrict__ mean, int C, int P,
                                                   double *__restrict__ sigma)
{
    __shared__ double sa[WCH][16 + 1], sb[WCH][16 + 1];
    __shared__ double sw[WCH];
    const int t = blockIdx.x;
    int ta = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
    while (ta * (ta + 1) / 2 > t) --ta;
    while ((ta + 1) * (ta + 2) / 2 <= t) ++ta;
    const int tb = t - ta * (ta + 1) / 2;
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
    const int a = ta * 16 + r, b = tb * 16 + c;
    const int ca = ta * 16 + c, cb = tb * 16 + c;  // staged columns (c doubles as the column slot)
    double ra[WCH / 16], rb[WCH / 16], rw = 0.0;
    auto fetch = [&](int i0) {
#pragma unroll
        for (int q = 0; q < WCH / 16; ++q) {
            const int i = i0 + r + 16 * q;
            ra[q] = (i < C && ca < P) ? theta[(size_t)i * P + ca] : 0.0;
            rb[q] = (i < C && cb < P) ? theta[(size_t)i * P + cb] : 0.0;
        }
        if (threadIdx.x < WCH) rw = i0 + (int)threadIdx.x < C ? w[i0 + threadIdx.x] : 0.0;
    };
    double acc = 0.0;
    fetch(0);
    for (int i0 = 0; i0 < C; i0 += WCH) {
#pragma unroll
        for (int q = 0; q < WCH / 16; ++q) {
            sa[r + 16 * q][c] = ra[q];
            sb[r + 16 * q][c] = rb[q];
        }
        if (threadIdx.x < WCH) sw[threadIdx.x] = rw;
        __syncthreads();
        if (i0 + WCH < C) fetch(i0 + WCH);
        const int n = min(WCH, C - i0);
#pragma unroll 16
        for (int ii = 0; ii < n; ++ii) acc += sw[ii] * sa[ii][r] * sb[ii][c];
        __syncthreads();
    }
    if (a < P && b < P && a >= b) {
        const double v = acc - mean[a] * mean[b];
        sigma[(size_t)a * P + b] = v;
        sigma[(size_t)b * P + a] = v;
    }
}

// resampled state: row c takes old row src[c]
__global__ void gather_kernel(const double *__restrict__ theta_old, const double *__restrict__ prior_old,
                              const double *__restrict__ data_old, const double *__restrict__ post_old,
                              const int *__restrict__ src, int C, int P, double *__restrict__ theta,
                              double *__restrict__ prior, double *__restrict__ data, double *__restrict__ post)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    const int c = (int)(idx / P), j = (int)(idx % P);
    const int s = src[c];
    theta[idx] = theta_old[(size_t)s * P + j];
    if (j == 0) {
        prior[c] = prior_old[s];
        data[c] = data_old[s];
        post[c] = post_old[s];
    }
}

// ---------------------------------------------------------------- launchers
static inline dim3 grid1(long n, int bs = 256) { return dim3((unsigned)((n + bs - 1) / bs)); }

hipError_t launch_normals(double *xi, int C, int P, uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                          hipStream_t st)
{
    hipLaunchKernelGGL(normals_kernel, grid1((long)C * ((P + 1) / 2)), dim3(256), 0, st, xi, C, P, seed, stage,
                       step, tag);
    return hipGetLastError();
}

hipError_t launch_propose(const double *theta, const double *L, const double *xi, int C, int P, double *cand,
                          hipStream_t st)
{
    if (P <= 32)
        hipLaunchKernelGGL(propose_kernel, grid1((long)C * P), dim3(256), 0, st, theta, L, xi, C, P, cand);
    else
        hipLaunchKernelGGL(propose_tiled_kernel, dim3((C + PTM - 1) / PTM, (P + PTN - 1) / PTN), dim3(256), 0,
                           st, theta, L, xi, C, P, cand);
    return hipGetLastError();
}

hipError_t launch_prior_verify(const double *theta, int C, int P, const int *kind, const double *a,
                               const double *b, int *mask, hipStream_t st)
{
    if (P <= 32)
        hipLaunchKernelGGL(prior_verify_kernel, grid1(C), dim3(256), 0, st, theta, C, P, kind, a, b, mask);
    else
        hipLaunchKernelGGL(prior_verify_wave_kernel, grid1((long)C * 64), dim3(256), 0, st, theta, C, P, kind, a,
                           b, mask);
    return hipGetLastError();
}

hipError_t launch_prior_llk(const double *theta, int C, int P, const int *kind, const double *a, const double *b,
                            double *llk, hipStream_t st)
{
    if (P <= 32)
        hipLaunchKernelGGL(prior_llk_kernel, grid1(C), dim3(256), 0, st, theta, C, P, kind, a, b, llk);
    else
        hipLaunchKernelGGL(prior_llk_wave_kernel, grid1((long)C * 64), dim3(256), 0, st, theta, C, P, kind, a, b,
                           llk);
    return hipGetLastError();
}

hipError_t launch_prior_sample(double *theta, int C, int P, const int *kind, const double *a, const double *b,
                               uint64_t seed, uint32_t stage, uint32_t attempt, const int *only, hipStream_t st)
{
    hipLaunchKernelGGL(prior_sample_kernel, grid1((long)C * P), dim3(256), 0, st, theta, C, P, kind, a, b, seed,
                       stage, attempt, only);
    return hipGetLastError();
}

hipError_t launch_restore(double *cand, const double *theta, const int *mask, int C, int P, hipStream_t st)
{
    hipLaunchKernelGGL(restore_kernel, grid1((long)C * P), dim3(256), 0, st, cand, theta, mask, C, P);
    return hipGetLastError();
}

hipError_t launch_accept(const AcceptArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(accept_kernel, grid1(a.C), dim3(256), 0, st, a.theta, a.prior, a.data, a.post, a.ctheta,
                       a.cprior, a.cdata, a.cpost, a.mask, a.beta, a.C, a.P, a.seed, a.stage, a.step, a.dice,
                       a.decision, a.counters);
    return hipGetLastError();
}

hipError_t launch_posterior(const double *prior, const double *data, double beta, int C, double *post,
                            hipStream_t st)
{
    hipLaunchKernelGGL(posterior_kernel, grid1(C), dim3(256), 0, st, prior, data, beta, C, post);
    return hipGetLastError();
}

hipError_t launch_wcov(const double *theta, const double *w, int C, int P, double *mean, double *sigma,
                       hipStream_t st)
{
    const int nt = (P + 15) / 16;
    hipLaunchKernelGGL(wmean_kernel, dim3((P + 63) / 64), dim3(256), 0, st, theta, w, C, P, mean);
    hipLaunchKernelGGL(wcov_kernel, dim3(nt * (nt + 1) / 2), dim3(256), 0, st, theta, w, mean, C, P, sigma);
    return hipGetLastError();
}

hipError_t launch_gather(const double *theta_old, const double *prior_old, const double *data_old,
                         const double *post_old, const int *src, int C, int P, double *theta, double *prior,
                         double *data, double *post, hipStream_t st)
{
    hipLaunchKernelGGL(gather_kernel, grid1((long)C * P), dim3(256), 0, st, theta_old, prior_old, data_old,
                       post_old, src, C, P, theta, prior, data, post);
    return hipGetLastError();
}

}  // namespace altar

namespace altar {

// ---------------------------------------------------------------- delayed acceptance
// Two-stage Metropolis (Christen and Fox 2005).  Stage 1 screens the
// candidate with the surrogate posterior cpost_s = cprior + beta cdata_s:
// pass if the model accepts it and log u1 <= cpost_s - post_s (u1 from the
// screening stream).  A second kernel lists the passing chains in chain
// order (a one-block ballot scan), so the list does not depend on
// scheduling.
__global__ void da_screen_kernel(const double *__restrict__ cprior, const double *__restrict__ cdata_s,
                                 const double *__restrict__ post_s, const int *__restrict__ mask, double beta,
                                 int C, uint64_t seed, uint32_t stage, uint32_t step, double *__restrict__ cpost_s,
                                 int *__restrict__ flag, unsigned long long *__restrict__ counters)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int pass = 0, rej = 0, scr = 0;
    if (c < C) {
        const double cp = tempered(cprior[c], beta, cdata_s[c]);
        cpost_s[c] = cp;
        if (mask[c]) {
            rej = 1;
        } else {
            u32x4 r = draw(seed, stage, step, (uint32_t)c, TAG_SCREEN, 0u);
            const double u = u53(r.x, r.y);
            if (log(u) <= cp - post_s[c]) pass = 1;
            else scr = 1;
        }
        flag[c] = pass;
    }
    __shared__ unsigned cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (rej) atomicAdd(&cnt[0], 1u);
    if (scr) atomicAdd(&cnt[1], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (cnt[0]) atomicAdd(&counters[1], (unsigned long long)cnt[0]);
        if (cnt[1]) atomicAdd(&counters[3], (unsigned long long)cnt[1]);
    }
}

// exclusive scan of the pass flags in one block of 1024 threads (C chains in
// chunks), writing the ordered list of passing chains and their count
__global__ void da_compact_kernel(const int *__restrict__ flag, int C, int *__restrict__ index,
                                  int *__restrict__ count)
{
    __shared__ int wsum[16];
    __shared__ int base;
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int start = 0; start < C; start += blockDim.x) {
        const int c = start + threadIdx.x;
        const int f = c < C ? flag[c] : 0;
        const unsigned long long bal = __ballot(f);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wid] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wid; ++w) off += wsum[w];
        if (f) index[off + before] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += wsum[w];
            base += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *count = base;
}

// packed[j] = cand[index[j]] for the n screened-in chains
__global__ void da_pack_kernel(const double *__restrict__ cand, const int *__restrict__ index, int n, int P,
                               double *__restrict__ packed)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)n * P) return;
    const int j = (int)(idx / P), k = (int)(idx % P);
    packed[idx] = cand[(size_t)index[j] * P + k];
}

// Stage 2 for the n screened-in chains: accept if
// log u2 <= (cpost - post) - (cpost_s - post_s) with the exact cpost.
// decision (optional, per chain): 0 accepted, 2 unlikely at stage 2.
__global__ void da_accept_kernel(double *__restrict__ theta, double *__restrict__ prior, double *__restrict__ data,
                                 double *__restrict__ post, double *__restrict__ post_s,
                                 const double *__restrict__ cand, const double *__restrict__ cprior,
                                 const double *__restrict__ cdata_packed, const double *__restrict__ cpost_s,
                                 const int *__restrict__ index, int n, double beta, int P, uint64_t seed,
                                 uint32_t stage, uint32_t step, int *__restrict__ decision,
                                 unsigned long long *__restrict__ counters)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    int acc = 0, unl = 0;
    if (j < n) {
        const int c = index[j];
        const double cp = tempered(cprior[c], beta, cdata_packed[j]);
        u32x4 r = draw(seed, stage, step, (uint32_t)c, TAG_ACCEPT, 0u);
        const double u = u53(r.x, r.y);
        const double lhs = (cp - post[c]) - (cpost_s[c] - post_s[c]);
        if (log(u) <= lhs) {
            acc = 1;
            for (int k = 0; k < P; ++k) theta[(size_t)c * P + k] = cand[(size_t)c * P + k];
            prior[c] = cprior[c];
            data[c] = cdata_packed[j];
            post[c] = cp;
            post_s[c] = cpost_s[c];
        } else {
            unl = 1;
        }
        if (decision) decision[c] = acc ? 0 : 2;
    }
    __shared__ unsigned cnt[2];
    if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (acc) atomicAdd(&cnt[0], 1u);
    if (unl) atomicAdd(&cnt[1], 1u);
    __syncthreads();
    if (threadIdx.x == 0) {
        if (cnt[0]) atomicAdd(&counters[0], (unsigned long long)cnt[0]);
        if (cnt[1]) atomicAdd(&counters[2], (unsigned long long)cnt[1]);
    }
}

// Surrogate trajectory (Liu 2001, the surrogate transition method): each
// exact step runs `substeps` Metropolis steps of the surrogate posterior
// pi* = prior + beta llk_s from the chain's state, then judges the end point
// phi with the exact posterior, accepting with
//     min(1, pi(phi) pi*(theta) / (pi(theta) pi*(phi))).
// A Metropolis kernel is reversible with respect to its own target, so a
// k-fold composition of it is too, and this ratio keeps pi invariant.
// One sub-step: cpost_s = cprior + beta cdata_s; accept phi <- cand if the
// model accepts it and log u <= cpost_s - post_phi (u from the trajectory
// dice stream at step' = step * 1024 + sub).  moved[c] records whether any
// sub-step of this trajectory moved the chain (sub 0 resets it).
// counters: [0] accepted, [1] rejected by the model, [2] unlikely.
__global__ void da_traj_step_kernel(double *__restrict__ phi, double *__restrict__ prior_phi,
                                    double *__restrict__ post_phi, const double *__restrict__ cand,
                                    const double *__restrict__ cprior, const double *__restrict__ cdata_s,
                                    double *__restrict__ cpost_s, const int *__restrict__ mask, double beta, int C,
                                    int P, uint64_t seed, uint32_t stage, uint32_t step2, int first,
                                    int *__restrict__ moved, unsigned long long *__restrict__ counters)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    int acc = 0, rej = 0, unl = 0;
    if (c < C) {
        const double cp = tempered(cprior[c], beta, cdata_s[c]);
        cpost_s[c] = cp;
        if (mask[c]) {
            rej = 1;
        } else {
            u32x4 r = draw(seed, stage, step2, (uint32_t)c, TAG_TRAJ_ACCEPT, 0u);
            const double u = u53(r.x, r.y);
            if (log(u) <= cp - post_phi[c]) {
                acc = 1;
                for (int k = 0; k < P; ++k) phi[(size_t)c * P + k] = cand[(size_t)c * P + k];
                prior_phi[c] = cprior[c];
                post_phi[c] = cp;
            } else {
                unl = 1;
            }
        }
        moved[c] = first ? acc : (moved[c] | acc);
    }
    __shared__ unsigned cnt[3];
    if (threadIdx.x < 3) cnt[threadIdx.x] = 0;
    __syncthreads();
    if (acc) atomicAdd(&cnt[0], 1u);
    if (rej) atomicAdd(&cnt[1], 1u);
    if (unl) atomicAdd(&cnt[2], 1u);
    __syncthreads();
    if (threadIdx.x < 3 && cnt[threadIdx.x]) atomicAdd(&counters[threadIdx.x], (unsigned long long)cnt[threadIdx.x]);
}

hipError_t launch_da_traj_propose(const double *phi, int C, int P, const double *L, uint64_t seed, uint32_t stage,
                                  uint32_t step2, double *xi, double *cand, hipStream_t st)
{
    hipError_t e = launch_normals(xi, C, P, seed, stage, step2, TAG_TRAJ_PROPOSAL, st);
    if (e != hipSuccess) return e;
    return launch_propose(phi, L, xi, C, P, cand, st);
}

hipError_t launch_da_traj_step(double *phi, double *prior_phi, double *post_phi, const double *cand,
                               const double *cprior, const double *cdata_s, double *cpost_s, const int *mask,
                               double beta, int C, int P, uint64_t seed, uint32_t stage, uint32_t step2, int first,
                               int *index, unsigned long long *counters, hipStream_t st)
{
    // the moved flags live in the upper half of `index` (2 C ints)
    hipLaunchKernelGGL(da_traj_step_kernel, grid1(C), dim3(256), 0, st, phi, prior_phi, post_phi, cand, cprior,
                       cdata_s, cpost_s, mask, beta, C, P, seed, stage, step2, first, index + C, counters);
    return hipGetLastError();
}

hipError_t launch_da_traj_list(int *index, int C, int *count, hipStream_t st)
{
    hipLaunchKernelGGL(da_compact_kernel, dim3(1), dim3(1024), 0, st, index + C, C, index, count);
    return hipGetLastError();
}

hipError_t launch_da_screen(const double *cprior, const double *cdata_s, const double *post_s, const int *mask,
                            double beta, int C, uint64_t seed, uint32_t stage, uint32_t step, double *cpost_s,
                            int *index, int *count, unsigned long long *counters, hipStream_t st)
{
    // the pass flags are staged in the upper half of `index` (2 C ints)
    int *flag = index + C;
    hipLaunchKernelGGL(da_screen_kernel, grid1(C), dim3(256), 0, st, cprior, cdata_s, post_s, mask, beta, C, seed,
                       stage, step, cpost_s, flag, counters);
    hipLaunchKernelGGL(da_compact_kernel, dim3(1), dim3(1024), 0, st, flag, C, index, count);
    return hipGetLastError();
}

hipError_t launch_da_pack(const double *cand, const int *index, int n, int P, double *packed, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(da_pack_kernel, grid1((long)n * P), dim3(256), 0, st, cand, index, n, P, packed);
    return hipGetLastError();
}

hipError_t launch_da_accept(double *theta, double *prior, double *data, double *post, double *post_s,
                            const double *cand, const double *cprior, const double *cdata_packed,
                            const double *cpost_s, const int *index, int n, double beta, int P, uint64_t seed,
                            uint32_t stage, uint32_t step, int *decision, unsigned long long *counters,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(da_accept_kernel, grid1(n), dim3(256), 0, st, theta, prior, data, post, post_s, cand, cprior,
                       cdata_packed, cpost_s, index, n, beta, P, seed, stage, step, decision, counters);
    return hipGetLastError();
}

}  // namespace altar
