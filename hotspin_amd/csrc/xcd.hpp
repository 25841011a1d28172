// xcd.hpp -- XCD-aware block numbering shared by the FFT passes and the
// stage kernels.
#pragma once

#include <hip/hip_runtime.h>

// XCD-aware block order.  Workgroups are handed to the 8 XCDs round-robin
// (block b runs on XCD b % 8), and each XCD has its own L2.  Renumbering the
// blocks of a 2-D grid so that consecutive logical tiles run on the same XCD
// lets two tiles that share a 128-byte line hit one L2.  Bijective on [0, nb).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_swizzle(int pid, int nb) {
    const int q = nb / kXcds, r = nb % kXcds;
    const int x = pid % kXcds, i = pid / kXcds;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}
