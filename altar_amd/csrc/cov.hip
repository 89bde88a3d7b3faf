// GPU-resident COV stage (altar/altar/bayesian/COV.py:65-251,
// altar/lib/libaltar/bayesian/COV.cc:222-368): the median of the data
// log-likelihoods, the dbeta grid search and the normalized weights, and the
// multinomial resampling, on the device.  Only the new temperature and COV
// come back to the host (two doubles).
//
// * median: exact order statistics by radix selection on order-preserving
//   64-bit keys of the FP64 values (8 passes of 8 bits, LDS histograms, one
//   workgroup); for an even count the mean of the two middle values, as
//   GSL's gsl_stats_median_from_sorted_data (altar/altar/bayesian/Grid.py:57).
// * dbeta: the grid search of COV.cc:222-326 in one workgroup of 1024
//   threads; each COV evaluation (cov::cov, COV.cc:334-368) reduces in FP64
//   in a fixed order, so results are deterministic.
// * resampling: inclusive scan of the weights (fixed order), the chain's
//   uniform (Philox, resampling stream) located by binary search on the
//   edges {0, cumsum(w)} with the last edge forced to 1, integer counts, an
//   exclusive scan of the counts, and every old sample writing its index into
//   its slots: rows in original-sample order, each repeated by its count.
#include "common.hpp"
#include "kernels.hpp"

namespace altar {

constexpr int COV_THREADS = 1024;

__device__ inline uint64_t f64_key(double v)
{
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ inline double key_f64(uint64_t k)
{
    const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

// fixed-order FP64 sum over a 1024-thread workgroup; every thread gets it
__device__ inline double block_sum(double v, double *red)
{
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < COV_THREADS / 64; ++w) s += red[w];
        red[COV_THREADS / 64] = s;
    }
    __syncthreads();
    const double s = red[COV_THREADS / 64];
    __syncthreads();
    return s;
}

// k-th smallest (0-based) of n values
__device__ double radix_select(const double *__restrict__ x, int n, int k, unsigned *hist, uint64_t *state)
{
    if (threadIdx.x == 0) state[0] = 0ull, state[1] = (uint64_t)k;
    __syncthreads();
    for (int shift = 56; shift >= 0; shift -= 8) {
        for (int b = threadIdx.x; b < 256; b += blockDim.x) hist[b] = 0u;
        __syncthreads();
        const uint64_t prefix = state[0];
        const uint64_t mask = shift == 56 ? 0ull : (~0ull << (shift + 8));
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const uint64_t key = f64_key(x[i]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t kk = state[1];
            unsigned b = 0;
            while (b < 255u && kk >= hist[b]) {
                kk -= hist[b];
                ++b;
            }
            state[0] = prefix | ((uint64_t)b << shift);
            state[1] = kk;
        }
        __syncthreads();
    }
    const double r = key_f64(state[0]);
    __syncthreads();  // every thread has read state before a next call resets it
    return r;
}

// COV of w = exp(dbeta (llk - median)) normalized to unit sum; w written
struct CovEval {
    double cov, metric;
};

__device__ CovEval cov_eval(double dbeta, const double *__restrict__ llk, int C, double median, double target,
                            double *__restrict__ w, double *red)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        const double v = exp(dbeta * (llk[i] - median));
        w[i] = v;
        s += v;
    }
    const double sum = block_sum(s, red);
    const double wsum = C * (sum / C);
    const double scale = 1.0 / wsum;
    double m = 0.0;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        const double v = w[i] * scale;
        w[i] = v;
        m += v;
    }
    const double mean = block_sum(m, red) / C;
    double q = 0.0;
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        const double dv = w[i] - mean;
        q += dv * dv;
    }
    const double ss = block_sum(q, red);
    const double cov = sqrt(ss / (C - 1)) / mean;
    CovEval r;
    if (isinf(cov) || isnan(cov)) {
        r.cov = 1e100;
        r.metric = 1e100;
    } else {
        r.cov = cov;
        r.metric = (cov - target) * (cov - target);
    }
    return r;
}

// out[0] = new beta, out[1] = cov, out[2] = median; w (C) the weights
__global__ __launch_bounds__(COV_THREADS) void cov_update_kernel(const double *__restrict__ llk, int C, double beta,
                                                                  double tolerance, double target,
                                                                  double *__restrict__ w, double *__restrict__ out)
{
    __shared__ double red[COV_THREADS / 64 + 1];
    __shared__ unsigned hist[256];
    __shared__ uint64_t state[2];
    double median;
    if (C % 2) {
        median = radix_select(llk, C, C / 2, hist, state);
    } else {
        const double lo = radix_select(llk, C, C / 2 - 1, hist, state);
        const double hi = radix_select(llk, C, C / 2, hist, state);
        median = (lo + hi) / 2.0;
    }
    double lo = 0.0, hi = 1.0 - beta;
    CovEval e = cov_eval(hi, llk, C, median, target, w, red);
    double dbeta;
    if (e.cov < target || fabs(e.cov - target) < tolerance) {
        dbeta = hi;
        beta = 1.0;
    } else {
        const int Nb = 10;
        int nb = 0, loops = 0;
        bool found = false;
        double guess = lo, step;
        do {
            ++loops;
            step = (hi - lo) / Nb;
            for (guess = lo, nb = 0; nb <= Nb; guess += step, ++nb) {
                e = cov_eval(guess, llk, C, median, target, w, red);
                if (fabs(e.cov - target) < tolerance) {
                    found = true;
                    break;
                }
                if (e.cov >= target) {
                    if (nb == 0) found = true;
                    break;
                }
            }
            if (nb > Nb) guess -= step;
            hi = guess;
            lo = guess - step;
        } while (loops <= Nb && !found);
        dbeta = guess;
        e = cov_eval(dbeta, llk, C, median, target, w, red);
        beta += dbeta;
    }
    if (threadIdx.x == 0) {
        out[0] = beta;
        out[1] = e.cov;
        out[2] = median;
    }
}

// edges[0] = 0, edges[i+1] = w[0] + ... + w[i] (fixed order), edges[C] = 1
__global__ __launch_bounds__(COV_THREADS) void cumsum_kernel(const double *__restrict__ w, int C,
                                                              double *__restrict__ edges)
{
    __shared__ double part[COV_THREADS];
    const int per = (C + COV_THREADS - 1) / COV_THREADS;
    const int b = threadIdx.x * per, e = min(C, b + per);
    double s = 0.0;
    for (int i = b; i < e; ++i) s += w[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < COV_THREADS; ++t) {
            const double v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = part[threadIdx.x];
    for (int i = b; i < e; ++i) {
        run += w[i];
        edges[i + 1] = run;
    }
    if (threadIdx.x == 0) edges[0] = 0.0;
    __syncthreads();
    if (threadIdx.x == 0) edges[C] = 1.0;
}

// count[k] = number of draws in (edges[k], edges[k+1]]
__global__ void multiplicity_kernel(const double *__restrict__ edges, int C, uint64_t seed, uint32_t stage,
                                    int *__restrict__ count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const u32x4 r = draw(seed, stage, 0u, (uint32_t)i, TAG_RESAMPLE, 0u);
    const double u = u53(r.x, r.y);
    int lo = 0, hi = C - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (u <= edges[mid + 1]) hi = mid;
        else lo = mid + 1;
    }
    atomicAdd(&count[lo], 1);
}

// exclusive scan of the counts (one workgroup), then src fill
__global__ __launch_bounds__(COV_THREADS) void count_scan_kernel(const int *__restrict__ count, int C,
                                                                  int *__restrict__ offset)
{
    __shared__ int part[COV_THREADS];
    const int per = (C + COV_THREADS - 1) / COV_THREADS;
    const int b = threadIdx.x * per, e = min(C, b + per);
    int s = 0;
    for (int i = b; i < e; ++i) s += count[i];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < COV_THREADS; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int i = b; i < e; ++i) {
        offset[i] = run;
        run += count[i];
    }
}

__global__ void zero_int_kernel(int *__restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void src_fill_kernel(const int *__restrict__ count, const int *__restrict__ offset, int C,
                                int *__restrict__ src)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const int o = offset[i], n = count[i];
    for (int j = 0; j < n; ++j) src[o + j] = i;
}

}  // namespace altar

namespace altar {

hipError_t launch_cov_update(const double *llk, int C, double beta, double tolerance, double target, double *w,
                             double *out, hipStream_t st)
{
    hipLaunchKernelGGL(cov_update_kernel, dim3(1), dim3(COV_THREADS), 0, st, llk, C, beta, tolerance, target, w, out);
    return hipGetLastError();
}

hipError_t launch_cov_resample_src(const double *w, int C, uint64_t seed, uint32_t stage, double *edges, int *count,
                                   int *offset, int *src, hipStream_t st)
{
    hipLaunchKernelGGL(cumsum_kernel, dim3(1), dim3(COV_THREADS), 0, st, w, C, edges);
    hipLaunchKernelGGL(zero_int_kernel, dim3((C + 255) / 256), dim3(256), 0, st, count, C);
    hipLaunchKernelGGL(multiplicity_kernel, dim3((C + 255) / 256), dim3(256), 0, st, edges, C, seed, stage, count);
    hipLaunchKernelGGL(count_scan_kernel, dim3(1), dim3(COV_THREADS), 0, st, count, C, offset);
    hipLaunchKernelGGL(src_fill_kernel, dim3((C + 255) / 256), dim3(256), 0, st, count, offset, C, src);
    return hipGetLastError();
}

}  // namespace altar
