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

// candidate = theta + (L xi)^T, one thread per (chain, parameter); L lower
__global__ void propose_kernel(const double *__restrict__ theta, const double *__restrict__ L,
                               const double *__restrict__ xi, int C, int P, double *__restrict__ cand)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)C * P) return;
    const int c = (int)(idx / P), i = (int)(idx % P);
    const double *x = xi + (size_t)c * P;
    const double *l = L + (size_t)i * P;
    double acc = 0.0;
    for (int j = 0; j <= i; ++j) acc += l[j] * x[j];
    cand[idx] = theta[idx] + acc;
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
            double z = (t[j] - a[j]) / b[j];
            acc += -0.5 * z * z - log(b[j] * sqrt(2.0 * PI));
        }
    }
    llk[c] = acc;
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
        const double bd = beta * cdata[c];
        const double cp = cprior[c] + bd;
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
                for (int j = 0; j < P; ++j) theta[(size_t)c * P + j] = ctheta[(size_t)c * P + j];
                prior[c] = cprior[c];
                data[c] = cdata[c];
                post[c] = cp;
            }
        }
        if (decision) decision[c] = dec;
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
        const double bd = beta * data[c];
        post[c] = prior[c] + bd;
    }
}

// weighted mean of every parameter (COV.py:137-140): one thread per parameter
__global__ void wmean_kernel(const double *__restrict__ theta, const double *__restrict__ w, int C, int P,
                             double *__restrict__ mean)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    double sw = 0.0, acc = 0.0;
    for (int i = 0; i < C; ++i) {
        sw += w[i];
        acc += w[i] * theta[(size_t)i * P + j];
    }
    mean[j] = acc / sw;
}

// Sigma = sum_i w_i theta_i theta_i^T - mean mean^T (COV.py:141-152), one
// thread per lower-triangle entry, fixed summation order
__global__ void wcov_kernel(const double *__restrict__ theta, const double *__restrict__ w,
                            const double *__restrict__ mean, int C, int P, double *__restrict__ sigma)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ne = (long)P * (P + 1) / 2;
    if (e >= ne) return;
    // e -> (a, b) with a >= b
    int a = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
    while ((long)a * (a + 1) / 2 > e) --a;
    while ((long)(a + 1) * (a + 2) / 2 <= e) ++a;
    const int b = (int)(e - (long)a * (a + 1) / 2);
    double acc = 0.0;
    for (int i = 0; i < C; ++i) acc += w[i] * theta[(size_t)i * P + a] * theta[(size_t)i * P + b];
    const double v = acc - mean[a] * mean[b];
    sigma[(size_t)a * P + b] = v;
    sigma[(size_t)b * P + a] = v;
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
    hipLaunchKernelGGL(propose_kernel, grid1((long)C * P), dim3(256), 0, st, theta, L, xi, C, P, cand);
    return hipGetLastError();
}

hipError_t launch_prior_verify(const double *theta, int C, int P, const int *kind, const double *a,
                               const double *b, int *mask, hipStream_t st)
{
    hipLaunchKernelGGL(prior_verify_kernel, grid1(C), dim3(256), 0, st, theta, C, P, kind, a, b, mask);
    return hipGetLastError();
}

hipError_t launch_prior_llk(const double *theta, int C, int P, const int *kind, const double *a, const double *b,
                            double *llk, hipStream_t st)
{
    hipLaunchKernelGGL(prior_llk_kernel, grid1(C), dim3(256), 0, st, theta, C, P, kind, a, b, llk);
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
    hipLaunchKernelGGL(wmean_kernel, grid1(P, 64), dim3(64), 0, st, theta, w, C, P, mean);
    hipLaunchKernelGGL(wcov_kernel, grid1((long)P * (P + 1) / 2, 64), dim3(64), 0, st, theta, w, mean, C, P,
                       sigma);
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
        const double bd = beta * cdata_s[c];
        const double cp = cprior[c] + bd;
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
        const double bd = beta * cdata_packed[j];
        const double cp = cprior[c] + bd;
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
