// Host launchers of the SSIMULACRA2 kernels (defined in ssimu2_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ssimu2_dev.hpp"

namespace vship {

hipError_t launch_prep(bool fp16, int grid, const PairDev* pairs, const int32_t* prep_start, const BatchHdr* hdr,
                       void* arena, hipStream_t st);
// cls 0: FP32 moments, tile height th (32 or 64); cls 1: FP64 moments, tile height kSmallTH
hipError_t launch_map(bool fp16, int cls, int th, int grid, const PairDev* pairs, const int32_t* order_seg,
                      const int32_t* order_start, const BatchHdr* hdr, const void* arena, double* partials,
                      const Taps& taps, hipStream_t st);
hipError_t launch_finalize(int grid, const PairDev* pairs, const int32_t* seg_start, const int32_t* seg_count,
                           const BatchHdr* hdr, const double* partials, double* scores, double* features,
                           hipStream_t st);
hipError_t launch_check(int* out, hipStream_t st);

}  // namespace vship
