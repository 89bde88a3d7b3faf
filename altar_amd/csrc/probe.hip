// VALU throughput calibration: the issue rate of the instruction kinds the
// forward-model kernels are made of, measured on the device (the published
// peaks do not say what a non-packed FP32 multiply or a reciprocal sustains).
// Each thread runs 8 independent dependency chains of one instruction kind,
// enough waves to fill every SIMD; the result is lane-operations per second.
#include "../../include/altar_amd.h"
#include "common.hpp"

#include <hip/hip_runtime.h>

namespace altar {

template <int KIND>
__global__ __launch_bounds__(256) void valu_probe_kernel(float *__restrict__ out, const float *__restrict__ in,
                                                         float sx, float sy, int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const float x = in[t & 1023], y = in[(t + 7) & 1023];
    float a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = x + (float)j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            if constexpr (KIND == 0) a[j] = fmaf(a[j], x, y);        // FMA, per-lane operands
            else if constexpr (KIND == 1) a[j] = fmaf(a[j], sx, sy); // FMA, wave-uniform operands
            else if constexpr (KIND == 2) a[j] = a[j] * x;           // multiply
            else if constexpr (KIND == 3) a[j] = a[j] + x;           // add
            else a[j] = __builtin_amdgcn_rcpf(a[j]);                 // hardware reciprocal
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += a[j];
    out[t] = s;
}

__global__ __launch_bounds__(256) void valu_probe_f64_kernel(double *__restrict__ out, const double *__restrict__ in,
                                                             int iters)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const double x = in[t & 1023], y = in[(t + 7) & 1023];
    double a[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) a[j] = x + (double)j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = fma(a[j], x, y);
    }
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < 8; ++j) s += a[j];
    out[t] = s;
}

}  // namespace altar

using namespace altar;

extern "C" int altar_valu_probe(int kind, double *lane_ops_per_s)
{
    if (!lane_ops_per_s || kind < 0 || kind > 5) return ALTAR_EINVAL;
    const int blocks = 256 * 32, threads = 256, iters = 2048;
    const size_t n = (size_t)blocks * threads;
    void *out = nullptr, *in = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = ALTAR_OK;
    auto ok = [&](hipError_t e) {
        if (e != hipSuccess && rc == ALTAR_OK) rc = ALTAR_EHIP;
        return rc == ALTAR_OK;
    };
    if (ok(hipMalloc(&out, n * sizeof(double))) && ok(hipMalloc(&in, 1024 * sizeof(double))) &&
        ok(hipEventCreate(&a)) && ok(hipEventCreate(&b))) {
        double host[1024];
        for (int i = 0; i < 1024; ++i) host[i] = 0.999 + 1e-6 * i;
        float hostf[1024];
        for (int i = 0; i < 1024; ++i) hostf[i] = (float)host[i];
        if (kind == 5) ok(hipMemcpy(in, host, sizeof(host), hipMemcpyHostToDevice));
        else ok(hipMemcpy(in, hostf, sizeof(hostf), hipMemcpyHostToDevice));
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == ALTAR_OK; ++rep) {  // the second launch is timed
            ok(hipEventRecord(a, nullptr));
            dim3 g(blocks), t(threads);
            float *o = (float *)out;
            const float *i = (const float *)in;
            switch (kind) {
            case 0: hipLaunchKernelGGL(valu_probe_kernel<0>, g, t, 0, 0, o, i, 0.999f, 1e-3f, iters); break;
            case 1: hipLaunchKernelGGL(valu_probe_kernel<1>, g, t, 0, 0, o, i, 0.999f, 1e-3f, iters); break;
            case 2: hipLaunchKernelGGL(valu_probe_kernel<2>, g, t, 0, 0, o, i, 0.999f, 1e-3f, iters); break;
            case 3: hipLaunchKernelGGL(valu_probe_kernel<3>, g, t, 0, 0, o, i, 0.999f, 1e-3f, iters); break;
            case 4: hipLaunchKernelGGL(valu_probe_kernel<4>, g, t, 0, 0, o, i, 0.999f, 1e-3f, iters); break;
            default:
                hipLaunchKernelGGL(valu_probe_f64_kernel, g, t, 0, 0, (double *)out, (const double *)in, iters);
            }
            ok(hipGetLastError());
            ok(hipEventRecord(b, nullptr));
            ok(hipEventSynchronize(b));
            ok(hipEventElapsedTime(&ms, a, b));
        }
        if (rc == ALTAR_OK) *lane_ops_per_s = (double)n * iters * 8 / (ms * 1e-3);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (out) (void)hipFree(out);
    if (in) (void)hipFree(in);
    return rc;
}
