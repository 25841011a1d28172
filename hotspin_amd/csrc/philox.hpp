// philox.hpp -- counter-based Philox4x32-10 (Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3", SC'11) and Box-Muller.
// Counter (cell, step, attempt, stream), key = 64-bit seed: any cell's noise
// at any step is computed independently, identically on every device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                              uint32_t k0, uint32_t k1) {
    // each 32x32 -> 64-bit product is one v_mad_u64_u32 (instead of a
    // v_mul_lo_u32 + v_mul_hi_u32 pair)
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// 24-bit uniform in (0,1), exact in FP32 and FP64
__device__ __forceinline__ float u24f(uint32_t x) { return ((float)(x >> 8) + 0.5f) * (1.0f / 16777216.0f); }
__device__ __forceinline__ double u24d(uint32_t x) { return ((double)(x >> 8) + 0.5) * (1.0 / 16777216.0); }

// Three standard normals of cell I (components 0, 1 from the first Box-Muller
// pair, component 2 from the cosine of the second), FP32 from the hardware
// transcendentals: v_log_f32 (log2),
// v_sqrt_f32 and v_sin_f32 / v_cos_f32, which take their argument in
// revolutions (sin(2 pi u) = v_sin_f32(u), exact range reduction for u in
// (0, 1)).  Absolute error ~1e-6 against the FP64 Box-Muller of the same
// Philox words; a quarter of the instructions of the libm forms.
__device__ __forceinline__ void thermal_normals_hw(uint32_t I, uint32_t step, uint32_t attempt, uint32_t stream,
                                                   uint64_t seed, float& n0, float& n1, float& n2) {
    uint32_t c0 = I, c1 = step, c2 = attempt, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)seed, (uint32_t)(seed >> 32));
    const float m2ln2 = -1.38629436111989061883f;  // -2 ln 2
    float r0 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24f(c0)));
    float r1 = __builtin_amdgcn_sqrtf(m2ln2 * __builtin_amdgcn_logf(u24f(c2)));
    float u1 = u24f(c1), u3 = u24f(c3);
    n0 = r0 * __builtin_amdgcn_cosf(u1);
    n1 = r0 * __builtin_amdgcn_sinf(u1);
    n2 = r1 * __builtin_amdgcn_cosf(u3);
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
