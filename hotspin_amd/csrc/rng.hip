// rng.hip -- thermal-noise entry points (replace the cuRAND host generator of
// src/hotspin-core/modules/mod_brown.go:98-122 with a counter-based stream).
#include "hs_common.hpp"
#include "philox.hpp"

__global__ void k_thermal_normals_f64(double* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                      uint32_t attempt) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    double a, b, c;
    thermal_normals_d((uint32_t)I, step, attempt, stream, seed, a, b, c);
    xi[I] = a;
    xi[(long)N + I] = b;
    xi[2L * N + I] = c;
}

__global__ void k_thermal_normals_f32(float* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                      uint32_t attempt) {
    int I = blockIdx.x * blockDim.x + threadIdx.x;
    if (I >= N) return;
    float a, b, c;
    thermal_normals_hw((uint32_t)I, step, attempt, stream, seed, a, b, c);
    xi[I] = a;
    xi[(long)N + I] = b;
    xi[2L * N + I] = c;
}

__global__ void k_philox_raw(uint32_t* out, const uint32_t* ctr, int n, uint64_t key) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = ctr[4 * i], c1 = ctr[4 * i + 1], c2 = ctr[4 * i + 2], c3 = ctr[4 * i + 3];
    philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
    out[4 * i] = c0;
    out[4 * i + 1] = c1;
    out[4 * i + 2] = c2;
    out[4 * i + 3] = c3;
}

extern "C" void hs_thermal_normals_f64(double* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                       uint32_t attempt, CUstream s) {
    if (!xi || N < 0) {
        hs::set_error(HS_EINVAL, "hs_thermal_normals_f64");
        return;
    }
    if (N == 0) return;
    hipLaunchKernelGGL(k_thermal_normals_f64, hs::grid1d(N), hs::kBlock, 0, hs::S(s), xi, N, seed, stream, step,
                       attempt);
    HS_LAUNCH_CHECK("hs_thermal_normals_f64");
}

extern "C" void hs_thermal_normals_f32(float* xi, int N, uint64_t seed, uint32_t stream, uint32_t step,
                                       uint32_t attempt, CUstream s) {
    if (!xi || N < 0) {
        hs::set_error(HS_EINVAL, "hs_thermal_normals_f32");
        return;
    }
    if (N == 0) return;
    hipLaunchKernelGGL(k_thermal_normals_f32, hs::grid1d(N), hs::kBlock, 0, hs::S(s), xi, N, seed, stream, step,
                       attempt);
    HS_LAUNCH_CHECK("hs_thermal_normals_f32");
}

extern "C" void hs_philox4x32_10(uint32_t* out, const uint32_t* ctr, int n, uint64_t key, CUstream s) {
    if (!out || !ctr || n < 0) {
        hs::set_error(HS_EINVAL, "hs_philox4x32_10");
        return;
    }
    if (n == 0) return;
    hipLaunchKernelGGL(k_philox_raw, hs::grid1d(n), hs::kBlock, 0, hs::S(s), out, ctr, n, key);
    HS_LAUNCH_CHECK("hs_philox4x32_10");
}
