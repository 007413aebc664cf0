// Device helpers shared by the kernels of ssimu2_kernels.hip and map_mfma.hip.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "ssimu2_dev.hpp"

namespace vship {

// Global-memory view of a device pointer that the compiler cannot place (one
// read from a descriptor), so that loads through it are global loads rather
// than flat ones, which also wait on LDS traffic.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* as_global(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ float ldf(const float* p) { return *p; }
__device__ __forceinline__ float ldf(const __half* p) { return __half2float(*p); }
__device__ __forceinline__ void stf(float* p, float v) { *p = v; }
__device__ __forceinline__ void stf(__half* p, float v) { *p = __float2half_rn(v); }

// Virtual workgroup index such that the workgroups of one XCD (hardware
// workgroup b runs on XCD b % 8) hold one contiguous range [0, nwg).
__device__ __forceinline__ int xcd_block(int b, int nwg) {
    constexpr int kXcd = 8;
    const int x = b % kXcd, j = b / kXcd;
    const int q = nwg / kXcd, r = nwg % kXcd;
    return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

// Sum of the in-bounds normalised taps at position p of an axis of length n,
// from the cumulative tap table (renormalisation of src/ssimu2/gaussianblur.hpp:58-60).
template <typename AT>
__device__ __forceinline__ AT tap_total(const AT* cum, int p, int n) {
    const int lo = max(0, 8 - p);
    const int hi = max(min(16, n - 1 - p + 8), lo - 1);
    return cum[hi + 1] - cum[lo];
}

}  // namespace vship
