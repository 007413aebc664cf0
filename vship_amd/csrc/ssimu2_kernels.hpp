// Host launchers of the SSIMULACRA2 kernels (defined in ssimu2_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "ssimu2_dev.hpp"

namespace vship {

// generic: the prep kernel of the pairs on the generic chroma path
// (prep_generic), else the one of every other pair.
hipError_t launch_prep(bool fp16, bool generic, int grid, const PairDev* pairs, const int32_t* prep_start,
                       const BatchHdr* hdr, void* arena, hipStream_t st);
// Map tiles of one group: cls 0 FP32 moments with tile height th (48, 64 or
// 112), cls 1 FP64 moments with tile height kSmallTH; half: the group's
// planes are FP16.
hipError_t launch_map(bool half, int cls, int th, int grid, const PairDev* pairs, const int32_t* order_seg,
                      const int32_t* order_start, const BatchHdr* hdr, const void* arena, double* partials,
                      const Taps& taps, int csplit, hipStream_t st);
// Map tiles of the FP32 class on the matrix cores (map_mfma.hip); exact: the
// groups of float-input pairs (three-way bf16 split), else the scaled fp16
// split. csplit: 1 (a workgroup per tile) or 3 (a workgroup per tile and
// channel; the same partials).
hipError_t launch_map_mx(bool half, bool exact, int grid, const PairDev* pairs, const int32_t* order_seg,
                         const int32_t* order_start, const BatchHdr* hdr, const void* arena, double* partials,
                         int csplit, hipStream_t st);
hipError_t launch_finalize(int grid, const PairDev* pairs, const int32_t* seg_start, const int32_t* seg_count,
                           const BatchHdr* hdr, const double* partials, double* scores, double* features,
                           hipStream_t st);
hipError_t launch_check(int* out, hipStream_t st);

}  // namespace vship
