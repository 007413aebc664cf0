// Batch engine: plans a batch of frame pairs into device descriptors, owns the
// XYB arena, the partial-sum buffer and one HIP stream, and launches the three
// kernels (optionally through one captured hipGraph).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ssimu2_dev.hpp"

namespace vship {

// Error codes with the numeric values of the public header
// (src/VshipAPI.h:35-58 of the reference).
enum ErrCode : int {
    E_NoError = 0,
    E_OutOfVRAM = 1,
    E_OutOfRAM = 2,
    E_DifferingInputType = 3,
    E_NonRGBSInput = 4,
    E_DeviceCountError = 5,
    E_NoDeviceDetected = 6,
    E_BadDeviceArgument = 7,
    E_BadDeviceCode = 8,
    E_BadHandler = 9,
    E_BadPointer = 10,
    E_BadErrorType = 11,
};

struct VshipErr : std::exception {
    ErrCode code;
    std::string file;
    int line;
    std::string where;
    VshipErr(ErrCode c, const char* f, int l)
        : code(c), file(f), line(l), where(std::string(f) + ":" + std::to_string(l)) {}
    const char* what() const noexcept override { return where.c_str(); }
};
#define VSHIP_THROW(code) throw ::vship::VshipErr((code), __FILE__, __LINE__)

// Any HIP failure inside the engine surfaces as this code (the reference maps
// failed device allocations to OutOfVRAM, src/ssimu2/main.hpp:116-119).
#define HIP_OR_THROW(expr, code)                 \
    do {                                         \
        hipError_t e_ = (expr);                  \
        if (e_ != hipSuccess) VSHIP_THROW(code); \
    } while (0)

// Luma coefficients (kr, kb) of a YUV matrix: 0 BT.709, 1 BT.601,
// 2 BT.2020 non-constant luminance.
void matrix_coefs(int matrix, float& kr, float& kb);

// Throws E_BadPointer unless `rows` rows of `row_bytes` bytes at a pitch of
// `stride` bytes from `ptr` lie inside one device allocation and span less
// than 4 GiB (the kernels address planes with 32-bit byte offsets), so that a
// host pointer or a short buffer is an error on the host and never a GPU fault.
void require_device_rows(const void* ptr, int64_t stride, int rows, int64_t row_bytes);

struct PairInput {
    const uint8_t* src[2][3];  // device pointers [image][plane]
    int64_t stride[2][3];
    int w, h;
    int fmt;     // InFmt
    int bits;
    int matrix;  // 0 BT.709, 1 BT.601, 2 BT.2020 NCL
    int chroma_loc = 0;  // ChromaLoc of the subsampled YUV formats
};

struct EngineOptions {
    int device = 0;
    int max_pairs = 1;
    int64_t arena_bytes = 0;   // XYB arena capacity
    int64_t max_map_tiles = 0;  // 0: derived from the arena for the smallest tile count per element
    bool fp16 = false;
    int map_th = 48;       // FP32-class map tiles: 32 x (48, 64 or 112) on the vector ALUs, or kMapMfma
    int persistent_grid = 0;   // 0: grid = tile count
    bool use_graph = false;
    bool profile = false;
    // submit_host: page-locked staging ring (chunk size, slots) and the CPU
    // threads that copy pageable planes into it
    int64_t host_chunk_bytes = 8ll << 20;
    int host_slots = 4;
    int host_copy_threads = 1;
};

struct Timings {
    float prep_ms = 0, map_ms = 0, map32_ms = 0, map64_ms = 0, finalize_ms = 0, total_ms = 0;
};

// Geometry of one pair in the arena (offsets in bytes, pitches in elements).
struct PairLayout {
    int sw[kScales], sh[kScales], pitch[kScales];
    int64_t scale_off[kScales], plane_elems[kScales];
    int64_t img_bytes;
};

constexpr int kMapGroups = 6;  // (FP32 | FP64 moments) x (FP32 | FP16 planes), + exact-split FP32 moments
inline int scale_elem_bytes(int s, bool fp16) { return fp16 && s < kHalfScales ? 2 : 4; }
PairLayout pair_layout(int w, int h, bool fp16);
int64_t pair_bytes(int w, int h, bool fp16);  // arena bytes of both images
int prep_tiles(int w, int h);
int map_tiles(int w, int h, int th);
uint64_t map_algorithmic_bytes(int w, int h, bool fp16);
uint64_t prep_algorithmic_bytes(const PairInput& p, bool fp16);

// Descriptor block of a batch, one allocation (host copy, then device copy):
//   [BatchHdr][PairDev x max_pairs][prep_start x (max_pairs+1)]
//   [seg_start x 6 max_pairs][seg_count x 6 max_pairs][order_seg x 6 max_pairs][order_start x (6 max_pairs+1)]
struct DescLayout {
    size_t off_pairs = 0, off_prep = 0, off_segstart = 0, off_segcount = 0, off_order = 0, off_ostart = 0;
    size_t bytes = 0;
};
DescLayout desc_layout(int max_pairs);

// Pointers into a descriptor block at `base`.
struct DescView {
    const BatchHdr* hdr;
    const PairDev* pairs;
    const int32_t *prep_start, *seg_start, *seg_count, *order_seg, *order_start;
};
DescView desc_view(const uint8_t* base, const DescLayout& L);

// What planning a batch found: tile counts per kernel and group, bytes.
struct PlanInfo {
    int npairs = 0, ntiles_prep = 0, ntiles_map = 0, nprep_generic = 0;
    int ntiles_cls[2] = {0, 0}, ntiles_grp[kMapGroups] = {};
    uint64_t map_bytes = 0, prep_bytes = 0, cls_bytes[2] = {0, 0};
};

// Plans pairs [0, n) into the host descriptor block `block` (layout L for at
// most max_pairs pairs): checks every pair (bad sizes, formats and planes
// throw E_BadPointer; an arena or tile table too small throws E_OutOfVRAM)
// and lays the pairs out in an arena of arena_bytes.
PlanInfo plan_batch(const PairInput* in, int n, int max_pairs, bool fp16, int map_th, int64_t arena_bytes,
                    int64_t max_map_tiles, uint8_t* block, const DescLayout& L);

// Launch shape of a batch's kernels.
struct LaunchShape {
    bool fp16 = false;
    int map_th = kMapMfma;
    int persistent_grid = 0;  // 0: one workgroup per tile
    bool graph = false;       // fixed grids (capture): every group launched, sized by max_pairs
    int max_pairs = 1;
    int num_cus = 256;
};

// Enqueues the prep, map and finalize kernels of a planned batch on st (no
// copies). ev, if not null, holds 4 events recorded after the prep kernels,
// after each map class and after finalize.
void launch_batch(const DescView& d, const PlanInfo& P, const LaunchShape& S, void* arena, double* partials,
                  double* scores, double* features, const Taps& taps, hipStream_t st, hipEvent_t* ev);

Taps make_taps();

class HostStager;

class Engine {
public:
    explicit Engine(const EngineOptions& opt);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Batches in flight at once: submitted and not yet waited for.
    static constexpr int kDepth = 2;

    // Plan and enqueue one batch of device-resident pairs (asynchronous).
    // Up to kDepth batches may be outstanding; a further submit is an error.
    void submit(const PairInput* pairs, int n);
    // The same for pairs whose plane pointers are host memory. The planes
    // are uploaded on the engine's copy stream into one of kDepth device
    // input buffers (each plane as one block of (rows - 1) strides plus a
    // row, keeping the caller's stride): through the page-locked staging ring
    // for pageable memory, by one DMA each for page-locked memory. The
    // kernels wait for the upload, so the upload of batch k+1 runs while
    // batch k's kernels do. Returns once the caller's memory may be reused.
    void submit_host(const PairInput* pairs, int n);
    // Staging-ring shape of submit_host; only before its first call.
    void set_host_staging(int64_t chunk_bytes, int slots, int copy_threads);
    // Wait for the oldest outstanding batch and copy its results out.
    void wait(double* scores, double* features);
    int outstanding() const { return (int)(head_ - tail_); }

    const EngineOptions& options() const { return opt_; }
    hipStream_t stream() const { return stream_; }
    const Timings& timings() const { return timings_; }
    int last_npairs() const { return plan_.npairs; }
    int last_prep_tiles() const { return plan_.ntiles_prep; }
    int last_map_tiles() const { return plan_.ntiles_map; }
    int last_map_tiles_class(int c) const { return plan_.ntiles_cls[c]; }
    uint64_t last_map_bytes() const { return plan_.map_bytes; }
    uint64_t last_map_bytes_class(int c) const { return plan_.cls_bytes[c]; }
    uint64_t last_prep_bytes() const { return plan_.prep_bytes; }
    uint64_t host_pageable_bytes() const;
    uint64_t host_direct_bytes() const;
    int num_cus() const { return num_cus_; }

private:
    void enqueue_body(int slot);
    void release();

    EngineOptions opt_;
    hipStream_t stream_ = nullptr;
    int num_cus_ = 0;
    Taps taps_{};

    DescLayout desc_;
    uint8_t* desc_h_[kDepth] = {};  // pinned, one per outstanding batch
    uint8_t* desc_d_ = nullptr;
    void* arena_d_ = nullptr;
    int64_t arena_bytes_ = 0;
    double* partials_d_ = nullptr;
    int64_t max_map_tiles_ = 0;
    double* out_d_ = nullptr;   // scores [max_pairs] then features [max_pairs*108]
    double* out_h_[kDepth] = {};  // pinned, same layout
    hipEvent_t ev_[kDepth][7] = {};
    hipEvent_t done_[kDepth] = {};
    hipGraph_t graph_[kDepth] = {};
    hipGraphExec_t graph_exec_[kDepth] = {};
    int npairs_[kDepth] = {};

    // host input (created on the first submit_host)
    hipStream_t copy_ = nullptr;
    std::unique_ptr<HostStager> stager_;
    uint8_t* in_d_[kDepth] = {};
    size_t in_cap_[kDepth] = {};
    hipEvent_t in_ready_[kDepth] = {}, in_free_[kDepth] = {};

    PlanInfo plan_;  // the last planned batch
    uint64_t head_ = 0, tail_ = 0;  // batches submitted, waited for
    Timings timings_;
};

}  // namespace vship
