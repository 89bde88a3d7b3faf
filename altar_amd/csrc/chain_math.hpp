// Per-chain arithmetic shared by the separate sampler kernels and the fused
// trajectory kernel.  Contraction into fused multiply-adds is switched off
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
// norm - (cperp + sum_k (e_k + s_k z_k)^2) / 2, summed over k in order
__device__ inline double surrogate_llk_sum(const float *z, const double *e, const double *s, int K, double cperp,
                                           double norm)
{
#pragma clang fp contract(off)
    double acc = cperp;
    for (int k = 0; k < K; ++k) {
        const double t = e[k] + s[k] * (double)z[k];
        acc += t * t;
    }
    return norm - 0.5 * acc;
}

}  // namespace altar
