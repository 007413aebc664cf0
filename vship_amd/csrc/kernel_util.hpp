// Device helpers shared by the kernels of ssimu2_kernels.hip and map_mfma.hip.
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "ssimu2_dev.hpp"

namespace vship {

// threadIdx.x read anew where it is used: the index arithmetic derived from it
// is recomputed per use instead of being hoisted out of a loop and held (or
// spilled) for the whole kernel.
__device__ __forceinline__ int tid_in_loop() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

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

// Next tile of a persistent grid's dynamic queue, or -1 when none is left.
// The launch's tiles [0, n) are split into 8 contiguous ranges, one per XCD
// (workgroup b runs on XCD b % 8, each XCD with its own L2), so that
// neighbouring tiles, whose halos overlap, are taken on one XCD. A workgroup
// takes the next tile of its own XCD's range from that range's counter
// (q[0..7], one vector atomic), then tiles of the other ranges once its own
// is used up. The counters are 0 at the start of every launch (they live in
// the batch's descriptor block, which every batch writes afresh). One lane
// calls it; it only hands out indices, so relaxed ordering suffices.
__device__ __forceinline__ int queue_next(int32_t* q, int n) {
    const int x = blockIdx.x & 7;
    for (int k = 0; k < 8; k++) {
        const int r = (x + k) & 7;
        const int lo = (int)(((int64_t)r * n) >> 3), hi = (int)(((int64_t)(r + 1) * n) >> 3);
        if (lo >= hi) continue;
        const int i = __hip_atomic_fetch_add(q + r, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (i < hi - lo) return lo + i;
    }
    return -1;
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
