// Shared device/host helpers for the altar_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

namespace altar {

// ----------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror and Shaw, SC'11).  The same generator
// and counter layout run on the host and on the device, so every random
// stream is reproducible bit for bit: counter = (chain, tag<<24 | word, step,
// stage), key = 64-bit seed.
struct u32x4 {
    uint32_t x, y, z, w;
};

__host__ __device__ inline u32x4 philox4x32_10(u32x4 c, uint32_t k0, uint32_t k1)
{
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)M0 * c.x;
        uint64_t p1 = (uint64_t)M1 * c.z;
        u32x4 n;
        n.x = (uint32_t)(p1 >> 32) ^ c.y ^ k0;
        n.y = (uint32_t)p1;
        n.z = (uint32_t)(p0 >> 32) ^ c.w ^ k1;
        n.w = (uint32_t)p0;
        c = n;
        k0 += W0;
        k1 += W1;
    }
    return c;
}

enum StreamTag : uint32_t {
    TAG_PROPOSAL = 0,   // Metropolis displacement normals
    TAG_ACCEPT = 1,     // Metropolis acceptance dice
    TAG_RESAMPLE = 2,   // COV multiplicity draws
    TAG_INIT = 3,       // initial sample from the prep distributions
    TAG_SCREEN = 4,     // delayed-acceptance first-stage dice
    TAG_TRAJ_PROPOSAL = 5,  // surrogate-trajectory sub-step normals
    TAG_TRAJ_ACCEPT = 6,    // surrogate-trajectory sub-step dice
};

__host__ __device__ inline u32x4 draw(uint64_t seed, uint32_t stage, uint32_t step,
                                      uint32_t chain, uint32_t tag, uint32_t word)
{
    u32x4 c = {chain, (tag << 24) | word, step, stage};
    return philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// 53-bit uniform on (0, 1]; exact in FP64
__host__ __device__ inline double u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)m + 1.0) * (1.0 / 9007199254740992.0);
}

constexpr double PI = 3.14159265358979323846;

// ----------------------------------------------------------------------------
// wavefront helpers (64 lanes on CDNA)
__device__ inline float wave_sum(float v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

inline size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

}  // namespace altar
