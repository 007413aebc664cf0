// Device-visible batch descriptors shared by the host engine and the kernels.
//
// A batch is a list of frame pairs of possibly different sizes and input
// formats. For each pair the engine reserves one region of the XYB arena:
//   image (0 = reference, 1 = distorted) -> scale 0..5 -> plane X, Y, B
// every plane row-major with a pitch that is a multiple of 64 elements, every
// plane offset a multiple of 128 bytes. Planes are FP32, except that with
// FP16 storage the two finest scales (s < kHalfScales, 15/16 of the plane
// bytes) are FP16: rounding the coarse scales to FP16 moves the score by up
// to 2e-2 (the coarse-scale features carry large weights), rounding scales
// 0 and 1 by about 2e-4.
#pragma once
#include <stdint.h>

namespace vship {

constexpr int kScales = 6;
constexpr int kPrepTile = 64;   // scale-0 tile edge of the convert/pyramid kernel
constexpr int kMapTW = 32;      // map tile width (columns)
constexpr int kFeat = 18;       // partial sums per map tile: 3 channels x 3 maps x {sum d, sum d^4}
constexpr int kHalfScales = 2;  // scales stored in FP16 when FP16 storage is on

// Input sample formats. RGB formats are the reference's InputMemType
// (src/util/preprocessor.hpp:94); the planar YUV formats are the fused front
// end, in the chroma layouts of src/ffvship_utility/gpuColorToLinear/
// vshipColor.hpp:12-128. YUV codes alternate 8-bit and 16-bit containers.
enum InFmt : int32_t {
    IN_RGB_U8 = 0,
    IN_RGB_U16 = 1,
    IN_RGB_F16 = 2,
    IN_RGB_F32 = 3,
    IN_YUV420_U8 = 4,
    IN_YUV420_U16 = 5,  // 9..16-bit samples in little-endian 16-bit containers
    IN_YUV422_U8 = 6,
    IN_YUV422_U16 = 7,
    IN_YUV444_U8 = 8,
    IN_YUV444_U16 = 9,
    IN_YUV440_U8 = 10,
    IN_YUV440_U16 = 11,
    IN_YUV411_U8 = 12,
    IN_YUV411_U16 = 13,
    IN_YUV410_U8 = 14,
    IN_YUV410_U16 = 15,
    IN_NUM_FORMATS = 16,
};

__host__ __device__ constexpr bool fmt_yuv(int f) { return f >= IN_YUV420_U8; }
// 16-bit container (YUV codes are odd for 16-bit, even for 8-bit)
__host__ __device__ constexpr bool fmt_wide(int f) { return fmt_yuv(f) && (f & 1); }
// log2 of the horizontal / vertical chroma subsampling of a YUV format, by
// layout: 4:2:0, 4:2:2, 4:4:4, 4:4:0, 4:1:1, 4:1:0
__host__ __device__ constexpr int fmt_csw(int f) {
    constexpr int8_t kW[6] = {1, 1, 0, 0, 2, 2};
    return fmt_yuv(f) ? kW[(f - IN_YUV420_U8) >> 1] : 0;
}
__host__ __device__ constexpr int fmt_csh(int f) {
    constexpr int8_t kH[6] = {1, 0, 0, 1, 0, 1};
    return fmt_yuv(f) ? kH[(f - IN_YUV420_U8) >> 1] : 0;
}

// Chroma sample positions (AVChromaLocation,
// src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp:131-205)
enum ChromaLoc : int32_t { CLOC_LEFT = 0, CLOC_CENTER = 1, CLOC_TOPLEFT = 2, CLOC_TOP = 3 };
__host__ __device__ constexpr bool cloc_hcentred(int c) { return c == CLOC_CENTER || c == CLOC_TOP; }
__host__ __device__ constexpr bool cloc_vcentred(int c) { return c == CLOC_LEFT || c == CLOC_CENTER; }

// A YUV pair whose chroma is not 4:2:0 with left-located samples goes through
// the generic chroma path, in its own prep kernel instantiation.
__host__ __device__ constexpr bool prep_generic(int fmt, int csw, int csh, int cloc) {
    return fmt_yuv(fmt) && !(csw == 1 && csh == 1 && cloc == CLOC_LEFT);
}

struct PairDev {
    const uint8_t* src[2][3];   // [image][plane] device pointers
    int64_t stride[2][3];       // bytes per row
    int32_t fmt;
    int32_t bits;               // YUV sample depth
    int32_t csw, csh;           // log2 chroma subsampling (YUV)
    int32_t cloc;               // ChromaLoc (YUV)
    float kr, kb;               // YUV matrix coefficients
    float ikg;                  // 1 / (1 - kr - kb)
    int32_t w, h;               // scale-0 size
    int64_t arena_off;          // byte offset of the pair's region
    int64_t img_bytes;          // bytes of one image region
    int64_t scale_off[kScales]; // byte offset of scale s (X plane) in an image region
    int64_t plane_elems[kScales];  // elements of one plane (pitch x height)
    int32_t pitch[kScales];
    int32_t sw[kScales], sh[kScales];
};

struct BatchHdr {
    int32_t npairs;
    int32_t prep_tiles;  // total convert/pyramid tiles
    int32_t prep_generic_pairs;  // pairs on the generic chroma path (prep_generic)
    int32_t map_tiles;   // total map tiles of all groups
    int32_t map_th;      // map tile height of the FP32 groups
    // Map segments (one per pair and scale) fall in two classes: class 0
    // accumulates the moment blur in FP32, class 1 (segments of fewer than
    // kSmallSegment pixels, where FP32 rounding noise would not average out)
    // in FP64 with tiles of kSmallTH rows. Each class splits by plane storage
    // into group 2*class (FP32 planes) and 2*class+1 (FP16 planes). On the
    // matrix cores, class-0 segments of pairs with float RGB input go to
    // groups 4 (FP32 planes) and 5 (FP16 planes) instead: they take the exact
    // bf16 split, the others the scaled fp16 one. Segments and tiles are
    // numbered group by group.
    int32_t nseg[6];
    int32_t tiles[6];
    // per map group, the 8 per-XCD tile counters of a persistent grid's
    // dynamic queue (queue_next); written as 0 with the rest of the header
    int32_t queue[6][8];
};

__host__ __device__ constexpr int map_group(int cls, bool half, bool exact = false) {
    return exact ? 4 + (half ? 1 : 0) : 2 * cls + (half ? 1 : 0);
}

constexpr int64_t kSmallSegment = 32768;
// The matrix-core map kernel centres each image on its own tile-centre value
// and keeps Var(x1 - x2) free of large cancelling terms, so its FP32 class
// reaches down to segments of kSmallSegmentMx pixels.
constexpr int64_t kSmallSegmentMx = 4096;
constexpr int kSmallTH = 48;
// Tile height value that selects the matrix-core map kernel for the FP32
// class; its tiles are kMxTile columns by kMxTileH rows (a wave streams down
// its 16-column strip of the tile, so a taller tile re-reads a smaller share
// of halo rows: 144 input rows for 128 output rows).
constexpr int kMapMfma = 1;
constexpr int kMxTile = 64;
constexpr int kMxTileH = 128;
__host__ __device__ constexpr int map_tile_w(int th) { return th == kMapMfma ? kMxTile : kMapTW; }
__host__ __device__ constexpr int map_tile_h(int th) { return th == kMapMfma ? kMxTileH : th; }

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
