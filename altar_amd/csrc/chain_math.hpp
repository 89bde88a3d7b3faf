// Per-chain arithmetic and the network's activation, shared by the separate
// sampler and network kernels and the fused trajectory kernel.  Contraction into fused multiply-adds is switched off
// in these functions: every product and sum is rounded on its own, in the
// order written, whoever inlines them (and as the FP64 oracle rounds them).
#pragma once

#include <hip/hip_runtime.h>

namespace altar {

// candidate_i = theta_i + sum_{j <= i} L_ij xi_j, the sum taken over j in order
__device__ inline double propose_elem(double t, const double *l, const double *x, int i)
{
#pragma clang fp contract(off)
    double acc = 0.0;
#pragma unroll 4
    for (int j = 0; j <= i; ++j) acc += l[j] * x[j];
    return t + acc;
}

// tempered posterior: prior + beta data
__device__ inline double tempered(double prior, double beta, double data)
{
#pragma clang fp contract(off)
    return prior + beta * data;
}

// log density of a gaussian prior at t without its constant c = log(sd sqrt(2 pi))
__device__ inline double gauss_log_term(double t, double mean, double sd, double c)
{
#pragma clang fp contract(off)
    const double z = (t - mean) / sd;
    return -0.5 * z * z - c;
}

// surrogate data llk from K network outputs z (unit stride):
// norm - (cperp + sum_k (e_k + s_k z_k)^2) / 2.  The sum is taken as eight
// interleaved partial sums (part q over k = q mod 8, in order of k), added
// to cperp in the order q = 0 .. 7, so that eight threads can share a row.
constexpr int SLLK_PARTS = 8;

__device__ inline double surrogate_llk_part(const float *z, const double *e, const double *s, int K, int q)
{
#pragma clang fp contract(off)
    double acc = 0.0;
#pragma unroll 4
    for (int k = q; k < K; k += SLLK_PARTS) {
        const double t = e[k] + s[k] * (double)z[k];
        acc += t * t;
    }
    return acc;
}

// surrogate_llk_part for K <= KMAX with a fixed trip count, the loads of
// every term issued ahead of the arithmetic: the same terms in the same
// order, each k >= K adding +0.0 to a sum of squares that is never -0.0,
// which leaves it unchanged.  z, e and s must be readable up to KMAX - 1
// (the values past K are not used).
template <int KMAX>
__device__ inline double surrogate_llk_part_upto(const float *z, const double *e, const double *s, int K, int q)
{
#pragma clang fp contract(off)
    constexpr int N = KMAX / SLLK_PARTS;
    double ev[N], sv[N];
    float zv[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        ev[i] = e[q + SLLK_PARTS * i];
        sv[i] = s[q + SLLK_PARTS * i];
        zv[i] = z[q + SLLK_PARTS * i];
    }
    double acc = 0.0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const double t = ev[i] + sv[i] * (double)zv[i];
        acc += q + SLLK_PARTS * i < K ? t * t : 0.0;
    }
    return acc;
}

__device__ inline double surrogate_llk_combine(const double (&p)[SLLK_PARTS], double cperp, double norm)
{
#pragma clang fp contract(off)
    double acc = cperp;
#pragma unroll
    for (int q = 0; q < SLLK_PARTS; ++q) acc += p[q];
    return norm - 0.5 * acc;
}

__device__ inline double surrogate_llk_sum(const float *z, const double *e, const double *s, int K, double cperp,
                                           double norm)
{
    double p[SLLK_PARTS];
#pragma unroll
    for (int q = 0; q < SLLK_PARTS; ++q) p[q] = surrogate_llk_part(z, e, s, K, q);
    return surrogate_llk_combine(p, cperp, norm);
}

// GELU in its tanh form (Hendrycks and Gimpel 2016):
//   gelu(x) = x/2 (1 + tanh(u)),  u = sqrt(2/pi) (x + 0.044715 x^3),
// evaluated as the identical function x s with s = 1 / (1 + exp(-2 u)) the
// logistic of 2u, and exp(-2u) = 2^(x (a + b x^2)) with the constants
//   a = -2 sqrt(2/pi) log2(e),  b = 0.044715 a.
// Six operations plus exp2 and a reciprocal on the transcendental unit;
// branch-free: 2^(...) overflows to +inf for very negative x, which gives
// s = 0, and underflows to 0 for large x, s = 1.
constexpr float GELU_A = -2.302208198144325f;     // -2 sqrt(2/pi) / ln 2
constexpr float GELU_B = -0.10294323958f;         // 0.044715 GELU_A

__device__ inline float gelu_logistic(float x)
{
#pragma clang fp contract(off)
    const float e = __builtin_amdgcn_exp2f(x * fmaf(GELU_B, x * x, GELU_A));
    return __builtin_amdgcn_rcpf(1.0f + e);
}

__device__ inline float gelu(float x)
{
#pragma clang fp contract(off)
    return x * gelu_logistic(x);
}

// d gelu / dx = s + x s (1 - s) 2 u'(x),  u'(x) = sqrt(2/pi) (1 + 3 * 0.044715 x^2)
__device__ inline float gelu_grad(float x)
{
#pragma clang fp contract(off)
    const float s = gelu_logistic(x);
    return s + x * s * (1.0f - s) * 1.5957691216057308f * (1.0f + 0.134145f * x * x);
}

}  // namespace altar
