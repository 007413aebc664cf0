// Device-side YUV sampling shared by the SSIMULACRA2 prep kernel and the
// Butteraugli input kernel: sample reads with the declared depth's clamp, and
// Catmull-Rom chroma upsampling at the sample positions of every layout and
// chroma location (src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "kernel_util.hpp"
#include "ssimu2_dev.hpp"

namespace vship {

// One YUV sample; 16-bit-container samples above the declared depth's
// largest code are clamped to it. The byte offset is 32-bit (the host
// refuses planes of 4 GiB or more), so a plane base in scalar registers and a
// vector offset address it.
__device__ __forceinline__ float sample_u(const uint8_t* plane, int64_t stride, int x, int y, bool wide,
                                          int maxcode) {
    if (wide) {
        const uint32_t o = (uint32_t)y * (uint32_t)stride + 2u * (uint32_t)x;
        return (float)min((int)*as_global(reinterpret_cast<const uint16_t*>(plane + o)), maxcode);
    }
    return (float)*as_global(plane + ((uint32_t)y * (uint32_t)stride + (uint32_t)x));
}

// Catmull-Rom weights of samples c-1 .. c+2 at fraction t of [c, c+1]: the
// cubic Hermite interpolant with central-difference tangents
// (src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp:5-24). Exact in
// FP32 for the fractions k/8 that occur here.
__device__ __forceinline__ void cr_weights(float t, float w[4]) {
    w[0] = 0.5f * t * ((2.0f - t) * t - 1.0f);
    w[1] = 0.5f * ((3.0f * t - 5.0f) * t * t + 2.0f);
    w[2] = 0.5f * t * ((4.0f - 3.0f * t) * t + 1.0f);
    w[3] = 0.5f * t * t * (t - 1.0f);
}

// Chroma interval c and fraction t of luma position x under subsampling 2^s:
// chroma sample i sits at luma position 2^s i (co-sited) or
// 2^s i + (2^s - 1) / 2 (centred), so x maps to u = x / 2^s or
// (2x + 1 - 2^s) / 2^(s+1), c = floor(u), t = u - c
// (chromaUpsample.hpp:44-128: left/top t = 0, 1/2 or k/4; centre 1/4, 3/4 or
// 1/8, 3/8, 5/8, 7/8).
__device__ __forceinline__ void chroma_pos(int x, int s, bool centred, int& c, float& t) {
    if (!centred) {
        c = x >> s;
        t = (float)(x & ((1 << s) - 1)) / (float)(1 << s);
    } else {
        const int n = 2 * x + 1 - (1 << s);
        c = n >> (s + 1);  // floor, n may be negative
        t = (float)(n & ((2 << s) - 1)) / (float)(2 << s);
    }
}

// One row of chroma at interval cx: the sample itself without horizontal
// subsampling, else the four-tap horizontal interpolation (indices clamped).
template <bool WIDE>
__device__ __forceinline__ float chroma_row(const uint8_t* plane, int64_t stride, int cw, int y, int cx,
                                            const float wx[4], bool direct, int maxcode) {
    if (direct) return sample_u(plane, stride, cx, y, WIDE, maxcode);
    float a = 0.f;
#pragma unroll
    for (int b = 0; b < 4; b++) a = fmaf(wx[b], sample_u(plane, stride, min(max(cx - 1 + b, 0), cw - 1), y, WIDE, maxcode), a);
    return a;
}

// Chroma sample value at luma pixel (x, y) for any layout and chroma location:
// horizontal interpolation of up to four rows, then vertical
// (chromaUpsample.hpp:131-205 order).
template <bool WIDE>
__device__ __forceinline__ float chroma_at(const uint8_t* plane, int64_t stride, int cw, int chh, int x, int y,
                                           int csw, int csh, bool hc, bool vc, int maxcode) {
    int cx, cy;
    float tx, ty, wx[4], wy[4];
    chroma_pos(x, csw, hc, cx, tx);
    chroma_pos(y, csh, vc, cy, ty);
    cr_weights(tx, wx);
    cr_weights(ty, wy);
    if (csh == 0) return chroma_row<WIDE>(plane, stride, cw, cy, cx, wx, csw == 0, maxcode);
    float a = 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++)
        a = fmaf(wy[r], chroma_row<WIDE>(plane, stride, cw, min(max(cy - 1 + r, 0), chh - 1), cx, wx, csw == 0, maxcode),
                 a);
    return a;
}

// 4:2:0 with left-located chroma (the common case), staged: the 36 x 36
// chroma neighbourhood of a 64 x 64 luma tile at (x0, y0), origin
// (x0 / 2 - 1, y0 / 2 - 2), indices clamped to the plane. Thread t < 252 owns
// column t % 36 and rows t / 36 + 7 k of both planes, so a wave reads whole
// row runs and the clamped column is computed once. PITCH: floats per staged
// row (36, or 37 with a pad word).
constexpr int kCh420 = 36;

template <int PITCH>
__device__ __forceinline__ void stage_chroma420(float (*ch)[kCh420][PITCH], const uint8_t* const* src,
                                                const int64_t* stride, int W, int H, int x0, int y0, bool wide,
                                                int maxcode) {
    const int tid = tid_in_loop();
    const int cw = (W + 1) >> 1, chh = (H + 1) >> 1;
    const int cx0 = x0 / 2 - 1, cy0 = y0 / 2 - 2;
    if (tid < 252) {
        const int c = tid % 36;
        const int gx = min(max(cx0 + c, 0), cw - 1);
        for (int r = tid / 36; r < 36; r += 7) {
            const int gy = min(max(cy0 + r, 0), chh - 1);
            ch[0][r][c] = sample_u(src[1], stride[1], gx, gy, wide, maxcode);
            ch[1][r][c] = sample_u(src[2], stride[2], gx, gy, wide, maxcode);
        }
    }
}

// Chroma (scaled by cmul, offset by cadd) of the quad of luma pixels
// (2 qx + {0, 1}, 2 qy + {0, 1}) of the staged tile, in the order (xa, ya),
// (xa, yb), (xb, ya), (xb, yb); hx / hy: the quad has a second column / row
// (else it repeats the first). Horizontally the even column takes the
// co-sited sample and the odd one Catmull-Rom at 1/2, (-1, 9, 9, -1) / 16
// (chroma420_hrow, once per staged row: vertically adjacent quads share four
// of their five rows); vertically (centre-sited) the even row is Catmull-Rom
// at 3/4 from rows 0..3 and the odd row at 1/4 from rows 1..4, in integer
// weights of total 128 that cmul takes out.
// Horizontal step of one staged row for the quad column qx: the even column's
// co-sited sample and the odd column's Catmull-Rom value at 1/2.
__device__ __forceinline__ void chroma420_hrow(const float* row, int qx, float& he, float& ho) {
    he = row[qx + 1];
    ho = (9.0f * (row[qx + 1] + row[qx + 2]) - (row[qx] + row[qx + 3])) * (1.0f / 16.0f);
}

// Vertical step from the horizontal results of staged rows qy .. qy + 4.
__device__ __forceinline__ void chroma420_vquad(const float* he, const float* ho, bool hx, bool hy, float cmul,
                                                float cadd, float out[4]) {
    const float ve_e = -3.0f * he[0] + 29.0f * he[1] + 111.0f * he[2] - 9.0f * he[3];
    const float vo_e = -9.0f * he[1] + 111.0f * he[2] + 29.0f * he[3] - 3.0f * he[4];
    const float ve_o = -3.0f * ho[0] + 29.0f * ho[1] + 111.0f * ho[2] - 9.0f * ho[3];
    const float vo_o = -9.0f * ho[1] + 111.0f * ho[2] + 29.0f * ho[3] - 3.0f * ho[4];
    out[0] = fmaf(ve_e, cmul, cadd);
    out[1] = fmaf(hy ? vo_e : ve_e, cmul, cadd);
    out[2] = fmaf(hx ? ve_o : ve_e, cmul, cadd);
    out[3] = fmaf(hx ? (hy ? vo_o : ve_o) : (hy ? vo_e : ve_e), cmul, cadd);
}

// Both steps for one quad, the five rows filtered here.
__device__ __forceinline__ void chroma420_quad(const float (*ch)[kCh420 + 1], int qx, int qy, bool hx, bool hy,
                                               float cmul, float cadd, float out[4]) {
    float he[5], ho[5];
#pragma unroll
    for (int r = 0; r < 5; r++) chroma420_hrow(ch[qy + r], qx, he[r], ho[r]);
    chroma420_vquad(he, ho, hx, hy, cmul, cadd, out);
}

}  // namespace vship
