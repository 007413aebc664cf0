// Butteraugli engine for MI355X: the two-scale distortion map and its norms
// for batches of equally sized RGB frame pairs resident in device memory
// (the metric of the reference's src/butter/main.hpp:22-195).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ssimu2_dev.hpp"  // InFmt, chroma layouts

namespace vship {
namespace bt {

// Taps of one of the five Gaussians, normalised to sum 1, and their running
// sums (cum[i] = sum of g[0..i-1]) for the renormalisation at image edges.
struct Taps {
    float g[43];
    float cum[44];
};

// One pair as the kernels read it: device planes, strides in bytes; map, if
// not null, receives the blended distortion map (row stride map_stride bytes).
// YUV formats also read bits (sample depth of the 16-bit containers), cloc
// (ChromaLoc) and the matrix's kr, kb.
struct PairDesc {
    const uint8_t* src[2][3];
    int64_t stride[2][3];
    float* map;
    int64_t map_stride;
    int32_t bits;
    int32_t cloc;
    float kr, kb;
};

// Kernel groups, for per-kernel timing.
enum Stage : int {
    ST_INPUT = 0,  // sampling, linearisation, half-scale mean
    ST_OPSIN,      // sigma 1.2 blur + opsin dynamics
    ST_LOW,        // sigma 7.156 blur: low band, residual
    ST_MID,        // sigma 3.225 blur: middle band, X suppression
    ST_HIGH,       // sigma 1.564 blur: high / ultra-high bands
    ST_DIFF,       // Malta lines, L2 terms, mask inputs
    ST_MASK,       // sigma 2.7 blur of the masks, erosion, map, blend, norms
    ST_NORM,       // per-pair norm reduction
    ST_COUNT
};

class Engine {
public:
    // Throws VshipErr (E_OutOfVRAM, E_BadDeviceArgument) on failure.
    Engine(int width, int height, int capacity, int qnorm, float intensity, int device);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Scores n <= capacity pairs of input format fmt (InFmt: planar RGB
    // u8/u16/f16/f32, gamma-encoded, or a limited-range YUV layout);
    // norms[3 * i .. 3 * i + 2] = Q-norm, 3-norm, max of pair i.
    // Synchronous; the work runs on stream().
    void run(int fmt, const PairDesc* pairs, int n, double* norms);

    // Enqueue only; the norms land in norms_host() after a stream sync.
    void submit(int fmt, const PairDesc* pairs, int n);
    const double* norms_host() const { return norms_h_; }

    hipStream_t stream() const { return stream_; }
    int width() const { return w_; }
    int height() const { return h_; }
    int capacity() const { return cap_; }
    int device() const { return device_; }
    int last_npairs() const { return last_n_; }

    // Per-stage times of the last submit (ms), when profiling is on.
    void set_profile(bool on) { profile_ = on; }
    bool stage_ms(float* out) const;
    // Algorithmic HBM bytes of one pair for a stage (both scales).
    uint64_t stage_bytes(int stage) const;
    // Device bytes per pair of work planes.
    static uint64_t pair_bytes(int width, int height);

private:
    void launch(int fmt, int n);
    void release();

    int w_, h_, w2_, h2_, cap_, q_;
    float intensity_;
    int device_;
    bool profile_ = false;
    bool timed_ = false;
    int last_n_ = 0;
    int last_fmt_ = IN_RGB_F32;
    bool staged420_ = false;
    hipStream_t stream_ = nullptr;
    float* arena_ = nullptr;
    int64_t plane0_ = 0, plane1_ = 0;
    Taps* taps_d_ = nullptr;
    PairDesc* desc_d_ = nullptr;
    PairDesc* desc_h_ = nullptr;
    double* partials_d_ = nullptr;
    double* norms_d_ = nullptr;
    double* norms_h_ = nullptr;
    int mask_tiles_ = 0;
    hipEvent_t ev_[ST_COUNT + 1] = {};
};

}  // namespace bt
}  // namespace vship
