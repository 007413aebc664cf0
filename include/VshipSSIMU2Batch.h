/*
 * Additive SSIMULACRA2 entry points of the MI355X build: batched frame pairs
 * of uneven resolution already resident in device memory, and a continuous-
 * batching request server. These symbols have no counterpart in the
 * reference's src/VshipAPI.h; the reference entry points are unchanged
 * (VshipAPI.h).
 */
#ifndef VSHIP_SSIMU2_BATCH_HEADER
#define VSHIP_SSIMU2_BATCH_HEADER

#include "VshipAPI.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sample formats. RGB formats are the reference's InputMemType
 * (src/util/preprocessor.hpp:94): planar R, G, B, gamma-encoded (BT.709
 * transfer), u8/u16 full range or f16/f32 in [0,1].
 *
 * The YUV formats are planar, limited range, in the chroma layouts the
 * reference's FFVship front end accepts
 * (src/ffvship_utility/gpuColorToLinear/vshipColor.hpp:12-128): chroma planes
 * of ceil(width / 2^sw) x ceil(height / 2^sh) samples with (sw, sh) = 4:2:0
 * (1, 1), 4:2:2 (1, 0), 4:4:4 (0, 0), 4:4:0 (0, 1), 4:1:1 (2, 0), 4:1:0
 * (2, 1). Even codes hold 8-bit samples, odd codes 9..16-bit samples (the
 * pair's bits) in little-endian 16-bit words; samples above 2^bits - 1 are
 * read as 2^bits - 1. Chroma is upsampled with Catmull-Rom cubic weights at
 * the positions the pair's chroma_location gives, with edge samples clamped
 * (src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp:5-205), and
 * converted with the pair's matrix before the same pipeline runs. */
typedef enum {
    Vship_PF_RGB_U8 = 0,
    Vship_PF_RGB_U16 = 1,
    Vship_PF_RGB_F16 = 2,
    Vship_PF_RGB_F32 = 3,
    Vship_PF_YUV420_U8 = 4,
    Vship_PF_YUV420_U16 = 5,
    Vship_PF_YUV422_U8 = 6,
    Vship_PF_YUV422_U16 = 7,
    Vship_PF_YUV444_U8 = 8,
    Vship_PF_YUV444_U16 = 9,
    Vship_PF_YUV440_U8 = 10,
    Vship_PF_YUV440_U16 = 11,
    Vship_PF_YUV411_U8 = 12,
    Vship_PF_YUV411_U16 = 13,
    Vship_PF_YUV410_U8 = 14,
    Vship_PF_YUV410_U16 = 15,
} Vship_PixelFormat;

/* Position of the chroma samples relative to the luma grid, as
 * AVChromaLocation (chromaUpsample.hpp:131-205): co-sited ("left", "top")
 * puts chroma sample i at luma position 2^s * i, centred at
 * 2^s * i + (2^s - 1) / 2. */
typedef enum {
    Vship_Chroma_Left = 0,    /* horizontally co-sited, vertically centred (MPEG-2, H.264 4:2:0) */
    Vship_Chroma_Center = 1,  /* centred both ways (JPEG, MPEG-1) */
    Vship_Chroma_TopLeft = 2, /* co-sited both ways */
    Vship_Chroma_Top = 3,     /* horizontally centred, vertically co-sited */
} Vship_ChromaLocation;

typedef enum {
    Vship_Matrix_BT709 = 0,
    Vship_Matrix_BT601 = 1,
    Vship_Matrix_BT2020NCL = 2,
} Vship_Matrix;

/* One frame pair. Plane pointers are device pointers; strides in bytes. */
typedef struct {
    const uint8_t* ref[3];
    const uint8_t* dist[3];
    int64_t ref_stride[3];
    int64_t dist_stride[3];
    int width;
    int height;
    int format; /* Vship_PixelFormat */
    int bits;   /* sample depth of the 16-bit-container YUV formats (9..16); ignored otherwise */
    int matrix; /* Vship_Matrix for the YUV formats */
    int chroma_location; /* Vship_ChromaLocation for the subsampled YUV formats; 0 for the others */
} Vship_SSIMU2Pair;

typedef struct {
    int id;
} Vship_SSIMU2BatchHandler;

#define VSHIP_MAP_MFMA 1

typedef struct {
    int max_pairs;         /* pairs per batch */
    int max_width;         /* largest frame: sizes the default arena */
    int max_height;
    int fp16_planes;       /* 1: store XYB scales 0 and 1 in FP16, scales 2-5 in FP32 (FP32 math, FP64 sums) */
    uint64_t arena_bytes;  /* 0: max_pairs * Vship_SSIMU2PairFootprint(max_width, max_height, fp16_planes) */
    int map_tile_height;   /* 0: tuned shape if recorded, else VSHIP_MAP_MFMA; VSHIP_MAP_MFMA: tiles of 64
                              columns x 128 rows with the blur on the matrix cores; 48, 64, 112: 32-column
                              tiles of that many rows on the vector ALUs */
    int persistent_grid;   /* 0: tuned shape if recorded, else one workgroup per tile; or a fixed grid */
    int use_graph;         /* 1: capture the batch launch into a hipGraph once and replay it */
    int profile;           /* 1: time each kernel with HIP events (eager launches only) */
} Vship_SSIMU2BatchConfig;

/* Device bytes one in-flight pair needs in the XYB arena (both images, six scales). */
EXPORTPREPROCESS uint64_t Vship_SSIMU2PairFootprint(int width, int height, int fp16_planes);

EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchInit(Vship_SSIMU2BatchHandler* handler,
                                                       const Vship_SSIMU2BatchConfig* config);
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchFree(Vship_SSIMU2BatchHandler handler);

/* Enqueue a batch on the handler's stream (asynchronous). Two batches may be
 * outstanding (submitted and not yet waited for) at once; a third Submit
 * returns Vship_BadHandler. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchSubmit(Vship_SSIMU2BatchHandler handler,
                                                         const Vship_SSIMU2Pair* pairs, int npairs);
/* Submit for pairs whose plane pointers are HOST memory (the reference's
 * per-frame copies, src/ssimu2/main.hpp:116-135, for a whole batch). Each
 * plane is uploaded as one block of (rows - 1) strides plus one row on the
 * handler's copy stream into one of two device input buffers: pageable memory
 * through a ring of page-locked chunks filled by CPU threads, page-locked
 * memory (Vship_PinnedMalloc) by one DMA per plane. The kernels wait for the
 * upload, so batch k+1's upload overlaps batch k's kernels when batch k+1 is
 * submitted before batch k is waited for. Returns once the caller's planes
 * may be reused. Same limits as Submit. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchSubmitHost(Vship_SSIMU2BatchHandler handler,
                                                             const Vship_SSIMU2Pair* pairs, int npairs);
/* Shape of SubmitHost's staging ring: chunk bytes (4 KiB .. 1 GiB), slots
 * (2 .. 64) and CPU copy threads (1 .. 64). Defaults 8 MiB, 4, 4. Only
 * before the handler's first SubmitHost (Vship_BadHandler after it). */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchHostStaging(Vship_SSIMU2BatchHandler handler,
                                                              uint64_t chunk_bytes, int slots, int copy_threads);
/* Wait for the oldest outstanding batch; scores[npairs] and, if not NULL,
 * features[npairs*108] (reference feature order) receive the results. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchWait(Vship_SSIMU2BatchHandler handler, double* scores,
                                                       double* features);
/* Submit + Wait; Vship_BadHandler while a batch is outstanding. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchCompute(Vship_SSIMU2BatchHandler handler,
                                                          const Vship_SSIMU2Pair* pairs, int npairs, double* scores,
                                                          double* features);

typedef struct {
    float prep_ms;      /* convert + pyramid + XYB kernel */
    float map_ms;       /* blur + SSIM/edge maps, both precision classes */
    float map_fp32_ms;  /* map kernel of the segments with FP32 moment accumulation */
    float map_fp64_ms;  /* map kernel of the small segments (FP64 moment accumulation) */
    float finalize_ms;  /* FP64 norms + score kernel */
    float total_ms;     /* whole batch, first copy to last copy */
    int npairs;
    int prep_tiles;
    int map_tiles;
    int map_tiles_fp32;
    int map_tiles_fp64;
    int map_tile_height;
    uint64_t map_bytes;      /* algorithmic bytes the map kernels read (both images, all scales) */
    uint64_t map_bytes_fp32; /* the part read by the FP32 map kernel */
    uint64_t prep_bytes;     /* algorithmic bytes the prep kernel reads and writes */
} Vship_SSIMU2BatchStats;

/* ---- autotuner -------------------------------------------------------------
 * Times the batch on the sample for every map kernel shape (VSHIP_MAP_MFMA; 48, 64, 112),
 * persistent grid (one workgroup per tile, 2/4/8 x CU count) and graph
 * capture, returns the table and the fastest shape, and records it for the
 * sample's frame-size class: Vship_SSIMU2BatchInit then uses it when
 * map_tile_height and persistent_grid are 0. */
#define VSHIP_TUNE_MAX_CANDIDATES 32
typedef struct {
    int map_tile_height;
    int persistent_grid;
    int use_graph;
    float ms;
} Vship_SSIMU2TuneCandidate;

typedef struct {
    int map_tile_height; /* best shape */
    int persistent_grid;
    int use_graph;
    float best_ms;       /* wall time of one batch with the best shape */
    int ncandidates;
    Vship_SSIMU2TuneCandidate candidates[VSHIP_TUNE_MAX_CANDIDATES];
} Vship_SSIMU2TuneResult;

EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchTune(const Vship_SSIMU2Pair* pairs, int npairs, int fp16_planes,
                                                       int repeats, Vship_SSIMU2TuneResult* result);
/* Text table of the recorded shapes (one line per size class and precision). */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2TuneSave(const char* path);
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2TuneLoad(const char* path);

/* Per-kernel HIP-event timings of the last waited batch (profile = 1). */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2BatchGetStats(Vship_SSIMU2BatchHandler handler,
                                                           Vship_SSIMU2BatchStats* stats);

/* ---- continuous-batching server ------------------------------------------
 * Clips of device-resident frame pairs are submitted at any time and scored
 * by a worker thread. Before each GPU step, waiting clips are admitted in
 * arrival order while their per-pair device footprint
 * (Vship_SSIMU2PairFootprint) fits the memory budget and a slot is free;
 * the step then packs the next frames of every admitted clip. Two steps are
 * in flight at a time on two streams, so planning overlaps execution. A
 * clip's scores are the same whichever clips it shares steps with. */
typedef struct {
    int id;
} Vship_SSIMU2ServerHandler;

typedef struct {
    uint64_t memory_budget;  /* device bytes for the XYB arenas of the two in-flight steps */
    int max_pairs_per_step;  /* pairs packed into one GPU step */
    int max_inflight_clips;  /* admitted clips at a time (0: max_pairs_per_step) */
    int max_width;           /* largest frame a clip may have */
    int max_height;
    int fp16_planes;
    int map_tile_height;     /* 0: default */
    int persistent_grid;     /* 0: one workgroup per tile */
    int use_graph;
} Vship_SSIMU2ServerConfig;

typedef struct {
    int64_t pairs_done;
    int64_t steps;
    int64_t clips_done;
    double host_ms;          /* host time spent planning and enqueueing steps */
    int waiting_clips;
    int active_clips;
} Vship_SSIMU2ServerStats;

EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerInit(Vship_SSIMU2ServerHandler* handler,
                                                        const Vship_SSIMU2ServerConfig* config);
/* Drains the admitted and waiting work, then stops the worker. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerFree(Vship_SSIMU2ServerHandler handler);
/* frames: nframes pair descriptors (device planes, which must stay valid until
 * the clip is finished). arrival_ms: the request's arrival time on
 * Vship_SSIMU2ServerClockMs's clock, or 0 for "now"; latency is measured
 * from it. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerSubmitClip(Vship_SSIMU2ServerHandler handler,
                                                              const Vship_SSIMU2Pair* frames, int nframes,
                                                              double arrival_ms, int* clip_id);
/* *done = 1 and scores[nframes] / latency written once the clip is finished. */
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerPollClip(Vship_SSIMU2ServerHandler handler, int clip_id,
                                                            int* done, double* scores, double* latency_ms);
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerWaitClip(Vship_SSIMU2ServerHandler handler, int clip_id,
                                                            double* scores, double* latency_ms);
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ServerGetStats(Vship_SSIMU2ServerHandler handler,
                                                            Vship_SSIMU2ServerStats* stats);
/* Monotonic clock (ms) the server measures latency with. */
EXPORTPREPROCESS double Vship_SSIMU2ServerClockMs(void);

/* Device memory helpers so callers without a GPU framework can stage inputs. */
EXPORTPREPROCESS Vship_Exception Vship_DeviceMalloc(void** ptr, uint64_t size);
EXPORTPREPROCESS Vship_Exception Vship_DeviceFree(void* ptr);
EXPORTPREPROCESS Vship_Exception Vship_MemcpyHtoD(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception Vship_MemcpyDtoH(void* dst, const void* src, uint64_t size);
EXPORTPREPROCESS Vship_Exception Vship_MemcpyDtoD(void* dst, const void* src, uint64_t size);

/* ---- allocation-free entry point -------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc. Scores one device-resident pair with the kernels and
 * launch shape of Vship_SSIMU2BatchCompute at its defaults (matrix-core map
 * kernel, one workgroup per tile, FP32 planes), so the results are that entry
 * point's. It allocates nothing: it works in a fixed scratch area of
 * VSHIP_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends.
 *
 * Asynchronous: enqueues on `stream` (a hipStream_t; NULL for the null
 * stream) and returns; *score (and features[108], reference feature order,
 * if not NULL) are device memory written when the stream reaches that work.
 * Frames whose XYB planes exceed the scratch (above about 1920 x 1080)
 * return Vship_OutOfVRAM; bad pointers or sizes return Vship_BadPointer,
 * both before anything is enqueued. Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU race, so order them (one stream,
 * or events). */
#define VSHIP_NOALLOC_SCRATCH_BYTES (2048ull + 4096ull * 18ull * 8ull + (64ull << 20))
EXPORTPREPROCESS Vship_Exception Vship_SSIMU2ComputeNoAlloc(const Vship_SSIMU2Pair* pair, double* score,
                                                            double* features, void* stream);

#ifdef __cplusplus
} /* extern "C" */
#endif
#endif /* VSHIP_SSIMU2_BATCH_HEADER */
