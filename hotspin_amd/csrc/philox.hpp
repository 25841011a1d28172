// philox.hpp -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller.
// Counter (cell, step, attempt, stream), key = 64-bit seed: any cell's noise
// at any step is computed independently, identically on every device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 24-bit uniform in (0,1), exact in FP32 and FP64
__device__ __forceinline__ float u24f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
__device__ __forceinline__ double u24d(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

// three standard normals of cell I (components 0,1 from the first
// Box-Muller pair, component 2 from the cosine of the second)
__device__ __forceinline__ void thermal_normals_f(uint32_t I, uint32_t step, uint32_t attempt, uint32_t stream,
                                                  uint64_t seed, float& n0, float& n1, float& n2) {
    uint32_t c0 = I, c1 = step, c2 = attempt, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    float r0 = sqrtf(-2.0f * logf(u24f(c0)));
    float r1 = sqrtf(-2.0f * logf(u24f(c2)));
    float s0, cs0, s1, cs1;
    sincospif(2.0f * u24f(c1), &s0, &cs0);
    sincospif(2.0f * u24f(c3), &s1, &cs1);
    n0 = r0 * cs0;
    n1 = r0 * s0;
    n2 = r1 * cs1;
}
__device__ __forceinline__ void thermal_normals_d(uint32_t I, uint32_t step, uint32_t attempt, uint32_t stream,
                                                  uint64_t seed, double& n0, double& n1, double& n2) {
    uint32_t c0 = I, c1 = step, c2 = attempt, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    double r0 = sqrt(-2.0 * log(u24d(c0)));
    double r1 = sqrt(-2.0 * log(u24d(c2)));
    double a0 = 2.0 * 3.14159265358979323846 * u24d(c1);
    double a1 = 2.0 * 3.14159265358979323846 * u24d(c3);
    n0 = r0 * cos(a0);
    n1 = r0 * sin(a0);
    n2 = r1 * cos(a1);
}
