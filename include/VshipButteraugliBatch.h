/*
 * Additive Butteraugli entry points of the MI355X build: batches of equally
 * sized RGB frame pairs already resident in device memory, with the norms
 * returned in double precision. These symbols have no counterpart in the
 * reference's src/VshipAPI.h, whose Butteraugli entry points
 * (Vship_ButteraugliInit ... Vship_ComputeButteraugliUint16v2, VshipAPI.h)
 * take host planes one pair at a time and return the norms truncated to int.
 *
 * The metric is the reference's (src/butter/main.hpp:22-195): the same
 * stages, constants and FP32 arithmetic, a map at full and half scale blended
 * as full * 0.85 + 0.5 * half, and the norms
 *   Q-norm   (mean of map^q)^(1/q)      (q = qnorm, 2 by default)
 *   3-norm   (mean of map^3)^(1/3)
 *   inf-norm max of map
 * accumulated in FP64 here (the reference sums in FP32).
 */
#ifndef VSHIP_BUTTERAUGLI_BATCH_HEADER
#define VSHIP_BUTTERAUGLI_BATCH_HEADER

#include "VshipAPI.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int id;
} Vship_ButteraugliBatchHandler;

typedef struct {
    int width;
    int height;
    int max_pairs;              /* pairs per batch */
    int qnorm;                  /* exponent of the first norm (the reference's Qnorm) */
    float intensity_multiplier; /* display luminance in nits (the reference's default is 203) */
    int profile;                /* 1: time each kernel stage with HIP events */
} Vship_ButteraugliBatchConfig;

/* One pair. Planes are device pointers, strides in bytes, in the batch's
 * Vship_PixelFormat (VshipSSIMU2Batch.h): planar R, G, B, gamma-encoded
 * (BT.709 transfer), or limited-range YUV of any of its layouts, converted to
 * R'G'B' with the pair's matrix and chroma location as the SSIMULACRA2 entry
 * points do. distmap, if not NULL, is a device buffer of height rows of
 * distmap_stride bytes that receives the blended map. */
typedef struct {
    const uint8_t* ref[3];
    const uint8_t* dist[3];
    int64_t ref_stride[3];
    int64_t dist_stride[3];
    float* distmap;
    int64_t distmap_stride;
    int bits;            /* sample depth of the 16-bit-container YUV formats (9..16); ignored otherwise */
    int matrix;          /* Vship_Matrix for the YUV formats */
    int chroma_location; /* Vship_ChromaLocation for the subsampled YUV formats; 0 for the others */
} Vship_ButteraugliPair;

/* Device bytes one in-flight pair needs (work planes of both scales). */
EXPORTPREPROCESS uint64_t Vship_ButteraugliPairFootprint(int width, int height);

EXPORTPREPROCESS Vship_Exception Vship_ButteraugliBatchInit(Vship_ButteraugliBatchHandler* handler,
                                                            const Vship_ButteraugliBatchConfig* config);
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliBatchFree(Vship_ButteraugliBatchHandler handler);

/* Scores npairs <= max_pairs pairs of one Vship_PixelFormat (the
 * reference's Butteraugli entry points take Vship_PF_RGB_U16 and
 * Vship_PF_RGB_F32; u8, f16 and the YUV layouts are additive).
 * norms[3 * i ...] receives the Q-norm, 3-norm and inf-norm of pair i.
 * Synchronous. */
EXPORTPREPROCESS Vship_Exception Vship_ButteraugliBatchCompute(Vship_ButteraugliBatchHandler handler, int format,
                                                               const Vship_ButteraugliPair* pairs, int npairs,
                                                               double* norms);

typedef struct {
    float input_ms; /* sampling, linearisation, half-scale mean */
    float opsin_ms; /* sigma 1.2 blur + opsin dynamics, both scales */
    float low_ms;   /* sigma 7.156 blur: low band */
    float mid_ms;   /* sigma 3.225 blur: middle band */
    float high_ms;  /* sigma 1.564 blur: high and ultra-high bands */
    float diff_ms;  /* Malta lines, L2 terms, mask inputs */
    float mask_ms;  /* sigma 2.7 blur of the masks, erosion, map, blend, norm partials */
    float norm_ms;  /* per-pair norm reduction */
    int npairs;
    int timed;      /* 1 when the last batch ran with profile on */
    uint64_t stage_bytes[8]; /* algorithmic HBM bytes of each stage, whole batch */
} Vship_ButteraugliBatchStats;

EXPORTPREPROCESS Vship_Exception Vship_ButteraugliBatchGetStats(Vship_ButteraugliBatchHandler handler,
                                                                Vship_ButteraugliBatchStats* stats);

#ifdef __cplusplus
}  // extern "C"
#endif
#endif
