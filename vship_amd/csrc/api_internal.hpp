// Internals shared by the C ABI (vship_api.hip) and the VapourSynth filter.
#pragma once
#include <atomic>
#include <cstdint>
#include <memory>

#include "butter.hpp"
#include "engine.hpp"
#include "host_stage.hpp"

namespace vship {

const char* error_message(int code);
int device_count_or_throw();
bool gpu_kernel_check();
void gpu_full_check(int gpu_id);
int default_map_th(int w, int h);

// Autotuned launch shape of one frame-size class (tune.hip).
struct TuneEntry {
    int size_class = 0;
    int fp16 = 0;
    int th = 48;
    int grid = 0;
    int graph = 0;
    double ms_per_pair = 0;
};
int size_class(int w, int h);
bool tuned_shape(int w, int h, bool fp16, int* th, int* grid, int* graph);
void record_shape(const TuneEntry& e);

// One per-frame engine of fixed size, as a Vship_SSIMU2Handler refers to
// (src/ssimu2/main.hpp:153-191 of the reference): host RGB planes are
// uploaded through the engine's page-locked staging ring (Engine::submit_host)
// and scored as a batch of one.
struct FrameHandle {
    int width, height, device;
    std::unique_ptr<Engine> engine;
    std::atomic<bool> busy{false};

    FrameHandle(int w, int h, int device);
    ~FrameHandle();
    // fmt: IN_RGB_U8 / IN_RGB_U16 / IN_RGB_F16 / IN_RGB_F32; host planes
    double compute(int fmt, const uint8_t* const p1[3], const uint8_t* const p2[3], int64_t s1, int64_t s2);
};

// One Butteraugli engine of fixed size, as a Vship_ButteraugliHandler refers
// to (ButterComputingImplementation, src/butter/main.hpp:197-240 of the
// reference): host RGB planes are uploaded into device staging through a
// page-locked staging ring and scored as a batch of one.
struct ButterHandle {
    int width, height, device;
    std::unique_ptr<bt::Engine> engine;
    std::unique_ptr<HostStager> stager;
    uint8_t* staging = nullptr;
    size_t staging_bytes = 0;
    float* map = nullptr;  // device map, rows of width floats
    std::atomic<bool> busy{false};

    ButterHandle(int w, int h, int qnorm, float intensity, int device);
    ~ButterHandle();
    // fmt: IN_RGB_F32 / IN_RGB_U16; host planes. dst (host, may be null)
    // receives the map in rows of dst_stride bytes. norms: Q, 3, inf.
    void compute(int fmt, const uint8_t* const p1[3], const uint8_t* const p2[3], int64_t s1, int64_t s2, float* dst,
                 int64_t dst_stride, double norms[3]);
};

}  // namespace vship
