/*
 * Vship C API, MI355X (gfx950) build.
 *
 * Drop-in replacement for the reference's public header src/VshipAPI.h:
 * every type, enum value and function below has the same name, signature and
 * meaning as there (src/VshipAPI.h:20-148), so existing clients recompile or
 * relink unchanged. SSIMULACRA2 and Butteraugli run on the GPU.
 *
 * Additive entry points live in VshipSSIMU2Batch.h (batched and served
 * SSIMULACRA2) and VshipButteraugliBatch.h (batched Butteraugli).
 */
#ifndef VSHIP_API_HEADER
#define VSHIP_API_HEADER

#include <stdint.h>

#if defined(_WIN32)
#if defined(EXPORTVSHIPLIB)
#define EXPORTPREPROCESS __declspec(dllexport)
#else
#define EXPORTPREPROCESS __declspec(dllimport)
#endif
#else
#define EXPORTPREPROCESS __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* src/VshipAPI.h:20-30 */
typedef enum {
    Vship_HIP = 0,
    Vship_Cuda = 1,
} Vship_Backend;

typedef struct {
    int major;
    int minor;
    int minorMinor;
    Vship_Backend backend;
} Vship_Version;

/* Returns 3.1.0 / Vship_HIP, as the reference's library (src/VshipLib.cpp:71-80). */
EXPORTPREPROCESS Vship_Version Vship_GetVersion(void);

/* Error codes; numeric values are those of src/VshipAPI.h:35-58. */
typedef enum {
    Vship_NoError = 0,
    Vship_OutOfVRAM,
    Vship_OutOfRAM,
    Vship_DifferingInputType,
    Vship_NonRGBSInput,
    Vship_DeviceCountError,
    Vship_NoDeviceDetected,
    Vship_BadDeviceArgument,
    Vship_BadDeviceCode,
    Vship_BadHandler,
    Vship_BadPointer,
    Vship_BadErrorType,
} Vship_Exception;

/* src/VshipAPI.h:61 — number of visible GPUs. */
EXPORTPREPROCESS Vship_Exception Vship_GetDeviceCount(int* number);

/* src/VshipAPI.h:64 — selects the calling thread's device. */
EXPORTPREPROCESS Vship_Exception Vship_SetDevice(int gpu_id);

/* src/VshipAPI.h:67-75 */
typedef struct {
    char name[256];
    uint64_t VRAMSize;
    int integrated;
    int MultiProcessorCount;
    int WarpSize;
} Vship_DeviceInfo;

EXPORTPREPROCESS Vship_Exception Vship_GetDeviceInfo(Vship_DeviceInfo* device_info, int gpu_id);

/* src/VshipAPI.h:78 — count, select, run a test kernel. */
EXPORTPREPROCESS Vship_Exception Vship_GPUFullCheck(int gpu_id);

/* src/VshipAPI.h:82 — len == 0 returns the size needed (NUL included);
 * otherwise copies at most len-1 characters and NUL-terminates. */
EXPORTPREPROCESS int Vship_GetErrorMessage(Vship_Exception exception, char* out_message, int len);

/* src/VshipAPI.h:87-89 — an index into the library's handle table. */
typedef struct {
    int id;
} Vship_SSIMU2Handler;

/* src/VshipAPI.h:91-93 — page-locked host memory. */
EXPORTPREPROCESS Vship_Exception Vship_PinnedMalloc(void** ptr, uint64_t size);
EXPORTPREPROCESS Vship_Exception Vship_PinnedFree(void* ptr);

/* src/VshipAPI.h:96-99 — one handle per concurrent frame; width/height fixed at init. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2Init(Vship_SSIMU2Handler* handler, int width, int height);
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2Free(Vship_SSIMU2Handler handler);

/* src/VshipAPI.h:101-111 — planar R, G, B (BT.709 transfer, gamma-encoded) host
 * planes, float [0,1] or uint16 full range, strides in bytes. Synchronous; the
 * score is written only on success. A handle used by two threads at once
 * returns Vship_BadHandler for the second call. */
EXPORTPREPROCESS Vship_Exception Vship_ComputeSSIMU2Float(Vship_SSIMU2Handler handler, double* score,
                                                          const uint8_t* srcp1[3], const uint8_t* srcp2[3],
                                                          int64_t stride);
EXPORTPREPROCESS Vship_Exception Vship_ComputeSSIMU2Floatv2(Vship_SSIMU2Handler handler, double* score,
                                                            const uint8_t* srcp1[3], const uint8_t* srcp2[3],
                                                            int64_t stride, int64_t stride2);
EXPORTPREPROCESS Vship_Exception Vship_ComputeSSIMU2Uint16(Vship_SSIMU2Handler handler, double* score,
                                                           const uint8_t* srcp1[3], const uint8_t* srcp2[3],
                                                           int64_t stride);
EXPORTPREPROCESS Vship_Exception Vship_ComputeSSIMU2Uint16v2(Vship_SSIMU2Handler handler, double* score,
                                                             const uint8_t* srcp1[3], const uint8_t* srcp2[3],
                                                             int64_t stride, int64_t stride2);

/* src/VshipAPI.h:113-148 (src/VshipLib.cpp:252-365): Butteraugli of one
 * frame pair of fixed size. The handle owns a GPU engine
 * (vship_amd/csrc/butter.hip); Compute copies the host planes in, scores
 * them, and, when dstp is not NULL, writes the distortion map as float rows
 * of dststride bytes. The score's fields are ints: the norms are stored by
 * plain conversion, i.e. truncated, as the reference stores its float norms
 * in them. VshipButteraugliBatch.h returns them in double precision. */
typedef struct {
    int id;
} Vship_ButteraugliHandler;

typedef struct {
    int norm2;
    int norm3;
    int norminf;
} Vship_ButteraugliScore;

EXPORTPREPROCESS Vship_Exception Vship_ButteraugliInit(Vship_ButteraugliHandler* handler, int width, int height,
                                                       float intensity_multiplier);
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliInitv2(Vship_ButteraugliHandler* handler, int width, int height,
                                                         int Qnorm, float intensity_multiplier);
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliFree(Vship_ButteraugliHandler handler);
EXPORTPREPROCESS Vship_Exception Vship_ComputeButteraugliFloat(Vship_ButteraugliHandler handler,
                                                               Vship_ButteraugliScore* score, const uint8_t* dstp,
                                                               int64_t dststride, const uint8_t* srcp1[3],
                                                               const uint8_t* srcp2[3], int64_t stride);
EXPORTPREPROCESS Vship_Exception Vship_ComputeButteraugliFloatv2(Vship_ButteraugliHandler handler,
                                                                 Vship_ButteraugliScore* score, const uint8_t* dstp,
                                                                 int64_t dststride, const uint8_t* srcp1[3],
                                                                 const uint8_t* srcp2[3], int64_t stride,
                                                                 int64_t stride2);
EXPORTPREPROCESS Vship_Exception Vship_ComputeButteraugliUint16(Vship_ButteraugliHandler handler,
                                                                Vship_ButteraugliScore* score, const uint8_t* dstp,
                                                                int64_t dststride, const uint8_t* srcp1[3],
                                                                const uint8_t* srcp2[3], int64_t stride);
EXPORTPREPROCESS Vship_Exception Vship_ComputeButteraugliUint16v2(Vship_ButteraugliHandler handler,
                                                                  Vship_ButteraugliScore* score, const uint8_t* dstp,
                                                                  int64_t dststride, const uint8_t* srcp1[3],
                                                                  const uint8_t* srcp2[3], int64_t stride,
                                                                  int64_t stride2);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* VSHIP_API_HEADER */
