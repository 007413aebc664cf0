// Device-visible batch descriptors shared by the host engine and the kernels.
//
// A batch is a list of frame pairs of possibly different sizes and input
// formats. For each pair the engine reserves one region of the XYB arena:
//   image (0 = reference, 1 = distorted) -> scale 0..5 -> plane X, Y, B
// every plane row-major with a pitch that is a multiple of 64 elements, every
// plane offset a multiple of 64 elements (256-byte aligned rows in FP32).
#pragma once
#include <stdint.h>

namespace vship {

constexpr int kScales = 6;
constexpr int kPrepTile = 64;   // scale-0 tile edge of the convert/pyramid kernel
constexpr int kMapTW = 32;      // map tile width (columns)
constexpr int kFeat = 18;       // partial sums per map tile: 3 channels x 3 maps x {sum d, sum d^4}

// Input sample formats. RGB formats are the reference's InputMemType
// (src/util/preprocessor.hpp:94); YUV 4:2:0 is the fused front end.
enum InFmt : int32_t {
    IN_RGB_U8 = 0,
    IN_RGB_U16 = 1,
    IN_RGB_F16 = 2,
    IN_RGB_F32 = 3,
    IN_YUV420_U8 = 4,
    IN_YUV420_U16 = 5,  // 9..16-bit samples in little-endian 16-bit containers
};

struct PairDev {
    const uint8_t* src[2][3];   // [image][plane] device pointers
    int64_t stride[2][3];       // bytes per row
    int32_t fmt;
    int32_t bits;               // YUV sample depth
    float kr, kb;               // YUV matrix coefficients
    int32_t w, h;               // scale-0 size
    int64_t arena_off;          // element offset of the pair's region
    int64_t img_elems;          // elements of one image region
    int64_t scale_off[kScales]; // element offset of scale s (X plane) in an image region
    int64_t plane_elems[kScales];
    int32_t pitch[kScales];
    int32_t sw[kScales], sh[kScales];
};

struct BatchHdr {
    int32_t npairs;
    int32_t prep_tiles;  // total convert/pyramid tiles
    int32_t map_tiles;   // total map tiles of both classes
    int32_t map_th;      // map tile height of class 0
    // Map segments (one per pair and scale) fall in two classes: class 0
    // accumulates the moment blur in FP32, class 1 (segments of fewer than
    // kSmallSegment pixels, where FP32 rounding noise would not average out)
    // in FP64 with tiles of kSmallTH rows. Class-0 tiles are numbered first.
    int32_t nseg[2];
    int32_t tiles[2];
};

constexpr int64_t kSmallSegment = 32768;
constexpr int kSmallTH = 48;

// The 17-tap Gaussian (sigma 1.5, radius 8) of src/ssimu2/gaussianblur.hpp:7-18
// and src/util/preprocessor.hpp:89-92, normalised to sum 1 and rounded to
// float: kTap[i] = float(exp(-i^2 / 4.5) / sum_{k=-8..8} exp(-k^2 / 4.5)).
// Normalising does not change the renormalised blur; it lets interior
// positions skip the division. The host recomputes these at start-up and
// refuses to run if they differ.
constexpr float kTap[9] = {
    0x1.105838p-2f, 0x1.b4272cp-3f, 0x1.bfdb5ap-4f, 0x1.26dcd8p-5f, 0x1.f1e5f2p-8f,
    0x1.0d8866p-10f, 0x1.763768p-14f, 0x1.4d2166p-18f, 0x1.7c4a7ep-23f,
};
__host__ __device__ constexpr float tap(int k) { return kTap[k < 8 ? 8 - k : k - 8]; }  // k = 0..16

struct Taps {
    float g[17];
};

}  // namespace vship
