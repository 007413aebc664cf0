/* Prints sizeof/offsetof of the public ABI structs so the ctypes mirror can be checked. */
#include <stddef.h>
#include <stdio.h>
#include "VshipAPI.h"
#include "VshipButteraugliBatch.h"
#include "VshipSSIMU2Batch.h"
int main(void) {
    printf("Vship_Version %zu\n", sizeof(Vship_Version));
    printf("Vship_DeviceInfo %zu %zu %zu\n", sizeof(Vship_DeviceInfo), offsetof(Vship_DeviceInfo, VRAMSize),
           offsetof(Vship_DeviceInfo, WarpSize));
    printf("Vship_SSIMU2Pair %zu %zu %zu\n", sizeof(Vship_SSIMU2Pair), offsetof(Vship_SSIMU2Pair, width),
           offsetof(Vship_SSIMU2Pair, matrix));
    printf("Vship_SSIMU2BatchConfig %zu %zu %zu\n", sizeof(Vship_SSIMU2BatchConfig),
           offsetof(Vship_SSIMU2BatchConfig, arena_bytes), offsetof(Vship_SSIMU2BatchConfig, profile));
    printf("Vship_SSIMU2BatchStats %zu %zu\n", sizeof(Vship_SSIMU2BatchStats),
           offsetof(Vship_SSIMU2BatchStats, map_bytes));
    printf("Vship_ButteraugliScore %zu\n", sizeof(Vship_ButteraugliScore));
    printf("Vship_ButteraugliBatchConfig %zu %zu\n", sizeof(Vship_ButteraugliBatchConfig),
           offsetof(Vship_ButteraugliBatchConfig, intensity_multiplier));
    printf("Vship_ButteraugliPair %zu %zu\n", sizeof(Vship_ButteraugliPair), offsetof(Vship_ButteraugliPair, distmap));
    printf("Vship_ButteraugliBatchStats %zu %zu\n", sizeof(Vship_ButteraugliBatchStats),
           offsetof(Vship_ButteraugliBatchStats, stage_bytes));
    return 0;
}
