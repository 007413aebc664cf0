// Map kernel on the gfx950 matrix cores, for the FP32-moment segments.
//
// Same result as map_kernel<TH, ST, float> (renormalised 17-tap Gaussian of
// the four moment planes a, b, (a - b)^2, ab of the samples centred per image
// on the tile-centre values (a = x1 - c1, b = x2 - c2), then the SSIM, artifact and
// detail-loss maps and their 1-/4-norm sums; src/ssimu2/gaussianblur.hpp:24-92,
// src/ssimu2/score.hpp:95-196), with both blur passes done as matrix products
// on the 16x16x32 MFMA instructions:
//
//   row pass    Hr(16 rows x 16 cols) = In(16 rows x 32 input cols) . G(32 x 16)
//   column pass Out(16 rows x 16 cols) = G'(16 x 32 input rows) . Hr(32 rows x 16 cols)
//
// where G[k][j] = g[k - j] (zero outside 0..16) is the band of taps; both
// passes use the same per-lane tap operand. The operands are split into
// 16-bit parts and the partial products accumulate in FP32, one of two ways:
//
//   SplitBf16x3 (planar float RGB input, whose values are unbounded): every
//     operand x is split exactly into three bf16 parts x = h + m + l
//     (truncation: h keeps the top 8 significant bits, m the next 8, l the
//     last 8); the six products above 2^-24 are accumulated (hh + hm + mh +
//     hl + mm + lh). FP32 accuracy, FP32 range.
//   SplitF16x2 (8- to 16-bit RGB and YUV input): operands are scaled by a
//     power of two (a, b and a - b: 2^2, so (a - b)^2 and ab: 2^4; taps:
//     2^14) and split into two fp16 parts h + l; the products hh + hl + lh
//     are accumulated, each exact (the lo x lo product, below the split's own
//     error, is dropped). A value whose low part is a normal fp16 number keeps
//     22 significant bits; a smaller one keeps an absolute error of at most
//     2^-25 at its scale (under 1e-8 in XYB units, far below the SSIM
//     constants 0.0009 these moments meet). The scales are the smallest
//     powers of two that need no multiply of their own (a and b come from
//     one FMA each, (a - b)^2 and ab are products of the scaled factors) and
//     they keep every value inside the fp16 range for any input of these
//     formats: their R'G'B' lies in [-0.97, 1.98] (limited-range YUV with
//     samples clamped to the declared depth), so XYB lies in [-11.7, 12.5] x
//     [-0.15, 1.58] x [-1.18, 2.28], |a|, |b| <= 24.2, |a - b| <= 48.4 and
//     |ab| <= 586, i.e. at most 37,500 at scale 2^4. The taps are the values
//     the fp16 pair represents exactly (within 2^-22 of the FP32 taps), and
//     the renormalisation sums use the same values. On the GPU the scores of
//     real frames stay within 3e-5 of the FP64 oracle, at about half the
//     splitting work of the exact split and two thirds of its matrix products.
//
// Tile: 64 columns x 128 rows of output pixels of one scale; each of the four
// waves owns a 16-column strip (144 input rows x 32 input columns) and needs
// nothing from the other waves until the tile sums. A wave streams nine 16-row
// blocks through the row pass; the row-pass results go, split, to a two-block LDS
// ring of the wave, from which the column pass of output block ob reads
// blocks ob and ob + 1. Each output pixel's maps are formed in the registers
// the column pass leaves them in.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernel_util.hpp"
#include "ssimu2_dev.hpp"
#include "ssimu2_kernels.hpp"

namespace vship {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// exact three-way bf16 split
// ---------------------------------------------------------------------------
__device__ __forceinline__ void split3(float x, float& h, float& m, float& l) {
    h = __uint_as_float(__float_as_uint(x) & 0xffff0000u);
    const float r = x - h;
    m = __uint_as_float(__float_as_uint(r) & 0xffff0000u);
    l = r - m;
}

// Two bf16 values (the upper halves of two floats) in one word, the first in
// the low half.
__device__ __forceinline__ uint32_t pack2(float first, float second) {
    return __builtin_amdgcn_perm(__float_as_uint(second), __float_as_uint(first), 0x07060302u);
}

__device__ __forceinline__ f32x4 mfma_bf16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

struct SplitBf16x3 {
    static constexpr int kParts = 3;
    static constexpr float kScaleAB = 1.f, kScaleDD = 1.f, kScaleQ = 1.f, kTapScale = 1.f;
    struct Op {
        u32x4 p[3];  // h, m, l
    };
    // N consecutive values (8 for an operand, 4 for a row-pass result)
    template <int N>
    static __device__ __forceinline__ void split(const float* v, uint32_t (&w)[3][N / 2]) {
#pragma unroll
        for (int k = 0; k < N / 2; k++) {
            float h0, m0, l0, h1, m1, l1;
            split3(v[2 * k], h0, m0, l0);
            split3(v[2 * k + 1], h1, m1, l1);
            w[0][k] = pack2(h0, h1);
            w[1][k] = pack2(m0, m1);
            w[2][k] = pack2(l0, l1);
        }
    }
    // split of the eight products x[e] y[e] (y = nullptr: of x[e])
    static __device__ __forceinline__ void split_prod8(const float* x, const float* y, uint32_t (&w)[3][4]) {
        float p[8];
#pragma unroll
        for (int k = 0; k < 8; k++) p[k] = y ? x[k] * y[k] : x[k];
        split<8>(p, w);
    }
    // split of x * s (row-pass results, re-split at the data scale)
    template <int N>
    static __device__ __forceinline__ void split_scaled(const float* v, float s, uint32_t (&w)[3][N / 2]) {
        float x[N];
#pragma unroll
        for (int k = 0; k < N; k++) x[k] = v[k] * s;
        split<N>(x, w);
    }
    // c + a.b from the six partial products above 2^-24, smallest first
    static __device__ __forceinline__ f32x4 mma(const Op& a, const Op& b, f32x4 c) {
        c = mfma_bf16(a.p[2], b.p[0], c);
        c = mfma_bf16(a.p[1], b.p[1], c);
        c = mfma_bf16(a.p[0], b.p[2], c);
        c = mfma_bf16(a.p[1], b.p[0], c);
        c = mfma_bf16(a.p[0], b.p[1], c);
        c = mfma_bf16(a.p[0], b.p[0], c);
        return c;
    }
    static __device__ __forceinline__ float tap_value(int k) { return tap(k); }
};

// ---------------------------------------------------------------------------
// scaled two-way fp16 split
// ---------------------------------------------------------------------------
__device__ __forceinline__ f32x4 mfma_f16(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0,
                                                  0, 0);
}

struct SplitF16x2 {
    static constexpr int kParts = 2;
    static constexpr float kScaleAB = 4.f;  // a, b: 2^2
    static constexpr float kScaleDD = 4.f;  // (a - b)^2 as (4 (a - b))^2: 2^4
    static constexpr float kScaleQ = 1.f;   // ab from the scaled a, b: 2^4
    static constexpr float kTapScale = 16384.f;      // taps: 2^14
    struct Op {
        u32x4 p[2];  // h, l
    };
    template <int N>
    static __device__ __forceinline__ void split(const float* v, uint32_t (&w)[2][N / 2]) {
#pragma unroll
        for (int k = 0; k < N / 2; k++) {
            const f32x2 x = {v[2 * k], v[2 * k + 1]};
            const f16x2 h = __builtin_convertvector(x, f16x2);
            const f32x2 r = x - __builtin_convertvector(h, f32x2);
            const f16x2 l = __builtin_convertvector(r, f16x2);
            w[0][k] = __builtin_bit_cast(uint32_t, h);
            w[1][k] = __builtin_bit_cast(uint32_t, l);
        }
    }
    // split of the eight products x[e] y[e] (y = nullptr: of x[e]) for a
    // matrix-core operand: h = fp16(x y) by the compiler; the low parts
    // l = fp16(x y - h), each one v_fma_mixlo/mixhi_f16 that forms x y - h
    // exactly and rounds once, written by inline assembly (the compiler
    // splits them into a multiply-subtract and a conversion). The block ends
    // with s_nop 1: the compiler counts no VALU wait states for inline
    // assembly, and a matrix-core instruction reading a register a VALU
    // instruction wrote needs two.
    static __device__ __forceinline__ void split_prod8(const float* x, const float* y, uint32_t (&w)[2][4]) {
        float p[8];
#pragma unroll
        for (int k = 0; k < 8; k++) p[k] = y ? x[k] * y[k] : x[k];
#pragma unroll
        for (int k = 0; k < 4; k++)
            w[0][k] = __builtin_bit_cast(uint32_t, __builtin_convertvector((f32x2){p[2 * k], p[2 * k + 1]}, f16x2));
        const float one = 1.0f;
        float q[8];
#pragma unroll
        for (int k = 0; k < 8; k++) q[k] = y ? y[k] : one;
        uint32_t l0, l1, l2, l3;
        asm("v_fma_mixlo_f16 %0, %4, %12, -%20 op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixhi_f16 %0, %5, %13, -%20 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixlo_f16 %1, %6, %14, -%21 op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixhi_f16 %1, %7, %15, -%21 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixlo_f16 %2, %8, %16, -%22 op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixhi_f16 %2, %9, %17, -%22 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixlo_f16 %3, %10, %18, -%23 op_sel_hi:[0,0,1]\n\t"
            "v_fma_mixhi_f16 %3, %11, %19, -%23 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
            "s_nop 1"
            : "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3)
            : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]),
              "v"(q[0]), "v"(q[1]), "v"(q[2]), "v"(q[3]), "v"(q[4]), "v"(q[5]), "v"(q[6]), "v"(q[7]),
              "v"(w[0][0]), "v"(w[0][1]), "v"(w[0][2]), "v"(w[0][3]));
        w[1][0] = l0;
        w[1][1] = l1;
        w[1][2] = l2;
        w[1][3] = l3;
    }
    // split of x * s: h = fp16(x s), l = fp16(fma(x, s, -h)) (the residual
    // x s - h is exact in FP32 when s is a power of two times a float that
    // keeps x s in range, and fma computes it in one rounding otherwise)
    template <int N>
    static __device__ __forceinline__ void split_scaled(const float* v, float s, uint32_t (&w)[2][N / 2]) {
#pragma unroll
        for (int k = 0; k < N / 2; k++) {
            const f32x2 y = {v[2 * k] * s, v[2 * k + 1] * s};
            const f16x2 h = __builtin_convertvector(y, f16x2);
            const f32x2 r = {fmaf(v[2 * k], s, -(float)h[0]), fmaf(v[2 * k + 1], s, -(float)h[1])};
            const f16x2 l = __builtin_convertvector(r, f16x2);
            w[0][k] = __builtin_bit_cast(uint32_t, h);
            w[1][k] = __builtin_bit_cast(uint32_t, l);
        }
    }
    // c + a.b from the three products above the split's own error: with
    // a = ah + al + ea, |al| <= 2^-12 |a|, |ea| <= 2^-23 |a| (and the same for
    // b), the dropped al.bl is at most 2^-24 |a b|, below ea.b and a.eb
    static __device__ __forceinline__ f32x4 mma(const Op& a, const Op& b, f32x4 c) {
        c = mfma_f16(a.p[1], b.p[0], c);
        c = mfma_f16(a.p[0], b.p[1], c);
        c = mfma_f16(a.p[0], b.p[0], c);
        return c;
    }
    // The tap the fp16 pair represents exactly (the operand split of
    // tap(k) * 2^14 reproduces it), divided back by 2^14.
    static __device__ __forceinline__ float tap_value(int k) {
        const float x = tap(k) * kTapScale;
        const _Float16 h = (_Float16)x;
        const _Float16 l = (_Float16)(x - (float)h);
        return ((float)h + (float)l) * (1.f / kTapScale);
    }
};

// Six wave sums at once, in FP32, by a fixed butterfly of v_add_f32_dpp
// (x plus the value a DPP control brings from another lane): quad_perm
// [1,0,3,2] and [2,3,0,1] sum the quads, row_half_mirror and row_mirror the
// rows of 16 lanes; row_bcast:15 adds row 0 into row 1 and row 2 into row 3,
// row_bcast:31 row 1 into row 3 (rows outside a step's row mask keep their
// value). Lane 63 ends with ((r3 + r2) + (r1 + r0)), the same order every
// time. The six chains are interleaved, so that every DPP read comes six
// instructions after the write of its source (two wait states are needed;
// s_nop 1 covers the writes before the block).
__device__ __forceinline__ void wave_sums6(float (&x)[6]) {
    asm volatile(
        "s_nop 1\n\t"
        "v_add_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %0, %0, %0 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 quad_perm:[2,3,0,1] row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %0, %0, %0 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 row_half_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %0, %0, %0 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 row_mirror row_mask:0xf bank_mask:0xf\n\t"
        "v_add_f32_dpp %0, %0, %0 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 row_bcast:15 row_mask:0xa bank_mask:0xf\n\t"
        "v_add_f32_dpp %0, %0, %0 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_add_f32_dpp %1, %1, %1 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_add_f32_dpp %2, %2, %2 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_add_f32_dpp %3, %3, %3 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_add_f32_dpp %4, %4, %4 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        "v_add_f32_dpp %5, %5, %5 row_bcast:31 row_mask:0xc bank_mask:0xf\n\t"
        : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]));
}

// Orders this wave's LDS accesses before and after (the wave's lanes exchange
// data through LDS; no other wave touches the region).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename ST>
struct Load8;
template <>
struct Load8<float> {
    static __device__ __forceinline__ void run(const float* p, float v[8]) {
        const float4 a = *reinterpret_cast<const float4*>(p);
        const float4 b = *reinterpret_cast<const float4*>(p + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    }
};
template <>
struct Load8<__half> {
    static __device__ __forceinline__ void run(const __half* p, float v[8]) {
        const uint4 u = *reinterpret_cast<const uint4*>(p);
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float2 f = __half22float2(*reinterpret_cast<const __half2*>(&w[k]));
            v[2 * k] = f.x;
            v[2 * k + 1] = f.y;
        }
    }
};

// Byte offset of element x of row y of a plane (planes are far below 4 GiB),
// and the element at a byte offset from a plane base: the base stays in
// scalar registers and each load takes one 32-bit vector offset.
template <typename ST>
__device__ __forceinline__ uint32_t plane_off(int y, int64_t pitch, int x) {
    return ((uint32_t)y * (uint32_t)pitch + (uint32_t)x) * (uint32_t)sizeof(ST);
}
template <typename ST>
__device__ __forceinline__ const ST* at(const ST* base, uint32_t byte_off) {
    return reinterpret_cast<const ST*>(reinterpret_cast<const uint8_t*>(base) + byte_off);
}

// Shared memory of one workgroup, for a split into NP parts.
template <int NP>
struct MxLds {
    // ring[wave][slot][moment][part][row group g][column j][2 words = rows
    // 4g .. 4g+3 of 16-bit parts]: the 16 lanes of one row group write 32
    // consecutive words, and a column-pass operand is two 8-byte reads 16
    // slots apart, so that no access of 16 lanes meets a bank twice
    uint32_t ring[4][2][4][NP][4][16][2];
    double red[4][kFeat];
    float rowf[4][kMxTileH];  // per wave: renormalisation factor of the tile's output rows
    float cum[18];      // cumulative taps, as the split represents them
};

// Geometry of one tile, the same for every channel.
struct TileGeo {
    int W, H, x0, y0, xo, g, j, wave;
    int xl;     // column of this lane's 8-element input load
    int cmask;  // bit e: input column xi + e lies in the image
    int64_t pitch, pe;
};

// Tap operand of a split (row pass B, column pass A): element e of lane
// (g, j) is tap[8g + e - j] at the split's tap scale.
template <typename SP>
__device__ __forceinline__ typename SP::Op tap_operand(int g, int j) {
    float tv[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = 8 * g + e - j;
        tv[e] = (k >= 0 && k <= 16) ? SP::tap_value(k) * SP::kTapScale : 0.f;
    }
    uint32_t w[SP::kParts][4];
    SP::template split<8>(tv, w);
    typename SP::Op G;
#pragma unroll
    for (int p = 0; p < SP::kParts; p++) G.p[p] = u32x4{w[p][0], w[p][1], w[p][2], w[p][3]};
    return G;
}

// Channels c_lo .. c_hi - 1 of one tile; leaves their wave sums in red[wave]
// (6 per channel, at 6 c).
//
// EDGE is false for tiles whose 80 x 144 input region lies inside the image:
// no element is masked, every output pixel is in the image and every
// renormalisation total is the full sum of the taps, so the factors fold into
// constants. Edge tiles clamp their loads into the image, give out-of-image
// elements the centre value (all four moments zero) and renormalise by the
// in-bounds taps of each output row and column.
//
// Scales: the row pass multiplies operands at the data scale (kScaleAB for a
// and b, kScaleDD^2 for (x1 - x2)^2, kScaleAB^2 kScaleQ for ab) by taps at
// kTapScale; its results are re-split at the data scale times the column
// factor (kmid); the column pass multiplies by the taps at kTapScale again.
// An output o of moment m is therefore the renormalised blur times
// scale_m kTapScale / (row factor), and k1 (per output row) undoes that for a
// and b.
template <typename SP, typename ST, bool EDGE>
__device__ __forceinline__ void mx_tile(MxLds<SP::kParts>& L, const TileGeo& T, const typename SP::Op& G,
                                        const ST* base1, const ST* base2, const float* cum, int lane, int c_lo,
                                        int c_hi) {
    constexpr int NP = SP::kParts;
    const int W = T.W, H = T.H, g = T.g, j = T.j, wave = T.wave;
    const int64_t pitch = T.pitch;
    // moments relative to a and b: (a - b)^2 and ab carry kScaleDD^2 and
    // kScaleAB^2 kScaleQ where a and b carry kScaleAB
    constexpr float kE = SP::kScaleAB / (SP::kScaleDD * SP::kScaleDD);
    constexpr float kQ = 1.f / (SP::kScaleAB * SP::kScaleQ);
    constexpr float kOut = 1.f / (SP::kScaleAB * SP::kTapScale);
    // the full tap sum as the split represents it (interior renormalisation)
    const float inv_total = __builtin_amdgcn_rcpf(cum[17]);
    // re-split scale of this lane's row-pass results (output column T.xo)
    const float kmid =
        (EDGE ? __builtin_amdgcn_rcpf(tap_total(cum, T.xo, W)) : inv_total) * (1.f / SP::kTapScale);
    if (EDGE) {
        // per-wave table of k1 for the tile's output rows (lane l: rows y0 + l + 64 q)
#pragma unroll
        for (int q = 0; q < kMxTileH / 64; q++)
            L.rowf[wave][lane + 64 * q] =
                __builtin_amdgcn_rcpf(tap_total(cum, min(T.y0 + lane + 64 * q, H - 1), H)) * kOut;
        wave_lds_sync();
    }
    const float k1i = inv_total * kOut;
    const int xc = EDGE ? min(T.xo, W - 1) : T.xo;
    // Work that can only add zeros is skipped on edge tiles: a wave whose 16
    // output columns all lie right of the image, and the row blocks below the
    // last 16-row output block that holds an image row (the last tile row of
    // a scale often holds far fewer than 128 rows). Skipping an addition of
    // exact zeros leaves every sum bit for bit as it was.
    if (EDGE && T.x0 + 16 * wave >= W) {
        if (lane == 63)
            for (int k = 0; k < kFeat; k++) L.red[wave][k] = 0.0;
        return;
    }
    constexpr int kMb = kMxTileH / 16 + 1;
    const int nmb = EDGE ? min(kMb, (min(H - T.y0, kMxTileH) + 15) / 16 + 1) : kMb;
    // The masks and selects of an edge tile are skipped, by a wave-uniform
    // branch, in the blocks that do not need them: when this wave's 32 input
    // columns (xin) or 16 output columns (xout) lie in the image, and so do
    // the block's rows, every mask bit is set and every select keeps its value.
    const int xs0 = T.x0 + 16 * wave;
    const bool xin = xs0 >= 8 && xs0 + 24 <= W, xout = xs0 + 16 <= W;

#pragma unroll 1
    for (int c = c_lo; c < c_hi; c++) {
        const ST* p1 = base1 + c * T.pe;
        const ST* p2 = base2 + c * T.pe;
        const int cxp = min(T.x0 + kMxTile / 2, W - 1), cyp = min(T.y0 + kMxTileH / 2, H - 1);
        // each image centred on its own value at the tile centre: a = x1 - cc,
        // b = x2 - cc2; the difference moment is then (a - b)^2, which stays
        // near 0 where the images differ by a locally constant offset, so
        // Vd = E[(a - b)^2] - (mu_a - mu_b)^2 does not cancel two large terms
        const float cc = ldf(p1 + (int64_t)cyp * pitch + cxp);
        const float cc2 = ldf(p2 + (int64_t)cyp * pitch + cxp);
        const float ccs = cc * SP::kScaleAB, ccs2 = cc2 * SP::kScaleAB;
        const float dc = cc - cc2;
        float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
        for (int mb = 0; mb < nmb; mb++) {
            // ---- row pass of input rows y0 - 8 + 16 mb .. +15; this lane: row j, columns xi .. xi + 7
            const int yr = T.y0 - 8 + 16 * mb + j;
            typename SP::Op A[4];
            {
                float u1[8], u2[8];
                const uint32_t ro = plane_off<ST>(EDGE ? min(max(yr, 0), H - 1) : yr, pitch, T.xl);
                Load8<ST>::run(at(p1, ro), u1);
                Load8<ST>::run(at(p2, ro), u2);
                if (EDGE && !(xin && T.y0 - 8 + 16 * mb >= 0 && T.y0 + 8 + 16 * mb <= H)) {
                    // one vector load per image for every lane: the row is
                    // clamped into the image and an element outside it takes
                    // the centre values cc and cc2, so that all four of its
                    // moments are exactly zero (ccs = cc kScaleAB is exact:
                    // the scale is a power of two)
                    const bool row_ok = yr >= 0 && yr < H;
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const bool in = row_ok && ((T.cmask >> e) & 1);
                        u1[e] = in ? u1[e] : cc;
                        u2[e] = in ? u2[e] : cc2;
                    }
                }
                // the moments as products of two factors: a, b, dd dd, (a kScaleQ) b
                float a[8], b[8], dd[8], aq[8];
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    a[e] = fmaf(u1[e], SP::kScaleAB, -ccs);
                    b[e] = fmaf(u2[e], SP::kScaleAB, -ccs2);
                    dd[e] = (a[e] - b[e]) * (SP::kScaleDD / SP::kScaleAB);
                    aq[e] = a[e] * SP::kScaleQ;
                }
                const float* fx[4] = {a, b, dd, aq};
                const float* fy[4] = {nullptr, nullptr, dd, b};
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    uint32_t w[NP][4];
                    SP::split_prod8(fx[m], fy[m], w);
#pragma unroll
                    for (int p = 0; p < NP; p++) A[m].p[p] = u32x4{w[p][0], w[p][1], w[p][2], w[p][3]};
                }
            }
            const f32x4 z = {0.f, 0.f, 0.f, 0.f};
            const int slot = mb & 1;
#pragma unroll
            for (int m = 0; m < 4; m++) {
                const f32x4 hr = SP::mma(A[m], G, z);
                // lane (g, j) holds rows 4g .. 4g+3 of the block, column j: store the split parts
                const float rv[4] = {hr[0], hr[1], hr[2], hr[3]};
                uint32_t w[NP][2];
                SP::template split_scaled<4>(rv, kmid, w);
#pragma unroll
                for (int p = 0; p < NP; p++)
                    *reinterpret_cast<uint2*>(&L.ring[wave][slot][m][p][g][j][0]) = make_uint2(w[p][0], w[p][1]);
            }
            wave_lds_sync();
            if (mb == 0) continue;

            // ---- column pass of output rows y0 + 16 ob .. +15 from ring blocks ob, ob + 1
            const int ob = mb - 1;
            const int rs = (ob + (g >> 1)) & 1, gg = 2 * (g & 1);
            f32x4 o[4];
#pragma unroll
            for (int m = 0; m < 4; m++) {
                typename SP::Op B;
#pragma unroll
                for (int p = 0; p < NP; p++) {
                    const uint2 lo = *reinterpret_cast<const uint2*>(&L.ring[wave][rs][m][p][gg][j][0]);
                    const uint2 hi = *reinterpret_cast<const uint2*>(&L.ring[wave][rs][m][p][gg + 1][j][0]);
                    B.p[p] = u32x4{lo.x, lo.y, hi.x, hi.y};
                }
                o[m] = SP::mma(G, B, z);
            }
            wave_lds_sync();
            // ---- maps of this lane's output pixels: rows y0 + 16 ob + 4g + r, column xo
            // The eight samples are loaded together (at the position clamped
            // into the image on edge tiles, where a pixel outside the image
            // adds zero through the selects on ok, which also discard the NaN
            // that its zero renormalisation total gives).
            float k1[4];
            if (EDGE) {
                const float4 rf = *reinterpret_cast<const float4*>(&L.rowf[wave][16 * ob + 4 * g]);
                k1[0] = rf.x; k1[1] = rf.y; k1[2] = rf.z; k1[3] = rf.w;
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++) k1[r] = k1i;
            }
            const int yo = T.y0 + 16 * ob + 4 * g;
            float s1[4], s2[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t off = plane_off<ST>(EDGE ? min(yo + r, H - 1) : yo + r, pitch, xc);
                s1[r] = ldf(at(p1, off));
                s2[r] = ldf(at(p2, off));
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const float mu1 = o[0][r] * k1[r], mu2 = o[1][r] * k1[r];
                const float E2 = o[2][r] * (k1[r] * kE), Q = o[3][r] * (k1[r] * kQ);
                // SSIM: d = (Vd + md^2 num) / (Vd + num) with md the difference
                // of the two means, Vd = var(x1 - x2), num = 2 cov + C2
                // (src/ssimu2/score.hpp:141-148); mu1, mu2 are the centred
                // means, so md = mu1 - mu2 + (cc - cc2)
                const float mdc = mu1 - mu2;
                const float vd = fmaf(-mdc, mdc, E2);
                const float md = mdc + dc;
                const float md2 = md * md;
                const float num_s = fmaf(2.f, fmaf(-mu1, mu2, Q), 0.0009f);
                const float n_s = fmaf(md2, num_s, vd);
                const float den = vd + num_s;
                // edge ratio (src/ssimu2/score.hpp:150-154)
                const float fa = fabsf((s1[r] - cc) - mu1) + 1.f;
                const float fb = fabsf((s2[r] - cc2) - mu2) + 1.f;
                // one reciprocal for both quotients: d = n / den and
                // (fb - fa) / fa = (fb - fa) den / (den fa); fb - fa is
                // exactly 0 for identical inputs, so both maps stay 0
                const float rr = __builtin_amdgcn_rcpf(den * fa);
                float d = (n_s * fa) * rr;
                const float ev = ((fb - fa) * den) * rr;
                float art = fmaxf(ev, 0.f), det = fmaxf(-ev, 0.f);
                d = fmaxf(d, 0.f);
                if (EDGE && !(xout && T.y0 + 16 * ob + 16 <= H)) {
                    const bool ok = T.xo < W && yo + r < H;
                    d = ok ? d : 0.f;
                    art = ok ? art : 0.f;
                    det = ok ? det : 0.f;
                }
                const float d2 = d * d, a2 = art * art, t2 = det * det;
                acc[0] += d;
                acc[1] = fmaf(d2, d2, acc[1]);
                acc[2] += art;
                acc[3] = fmaf(a2, a2, acc[3]);
                acc[4] += det;
                acc[5] = fmaf(t2, t2, acc[5]);
            }
        }
        // wave sums of this channel by the DPP butterfly (no LDS, no lane
        // permutes through it); the per-lane sums cover 16 pixels, the wave
        // sum 1,024, and the four waves and the tiles are added in FP64
        wave_sums6(acc);
        if (lane == 63) {
#pragma unroll
            for (int k = 0; k < 6; k++) L.red[wave][c * 6 + k] = (double)acc[k];
        }
    }
}

}  // namespace

// Occupancy: the two-part split fits 4 workgroups per CU (35 KB of LDS, at
// most 128 VGPRs), the three-part split 3 (52 KB of LDS).
//
// csplit: 1, a workgroup takes the three channels of a tile; 3, one channel
// (work item w: tile w / 3, channel w % 3), for batches with too few tiles to
// fill the GPU. A channel's sums do not depend on the others, so both give
// the same partials bit for bit.
template <typename ST, typename SP>
__global__ __launch_bounds__(256, SP::kParts == 2 ? 4 : 2) void map_mx_kernel(const PairDev* __restrict__ pairs,
                                                        const int32_t* __restrict__ order_seg,
                                                        const int32_t* __restrict__ order_start,
                                                        const BatchHdr* __restrict__ hdr,
                                                        const uint8_t* __restrict__ arena,
                                                        double* __restrict__ partials, int csplit) {
    constexpr int TS = kMxTile;
    constexpr int grp = map_group(0, sizeof(ST) == 2, SP::kParts == 3);
    __shared__ MxLds<SP::kParts> L;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15;
    if (tid == 0) {
        float a = 0.f;
        L.cum[0] = 0.f;
        for (int k = 0; k < 17; k++) {
            a += SP::tap_value(k);
            L.cum[k + 1] = a;
        }
    }
    __syncthreads();

    const typename SP::Op G = tap_operand<SP>(g, j);

    int seg_lo = 0, t_begin = 0;
#pragma unroll
    for (int q = 0; q < grp; q++) {
        seg_lo += hdr->nseg[q];
        t_begin += hdr->tiles[q];
    }
    const int seg_hi = seg_lo + hdr->nseg[grp] - 1;
    const int t_end = t_begin + hdr->tiles[grp];
    // A grid of one workgroup per work item (or more) numbers its workgroups
    // with xcd_block, so that the workgroups of one XCD hold one contiguous
    // range of tiles. A smaller, persistent grid takes its work items from the
    // per-XCD dynamic queue.
    const int nwg = gridDim.x;
    const int ntl = t_end - t_begin;
    const int nwork = ntl * csplit;
    const bool persistent = nwg < nwork;
    int32_t* queue = const_cast<int32_t*>(hdr->queue[grp]);
    __shared__ int s_tile;
    // (with more workgroups than work items, item floor(v n / nwg) of virtual
    // workgroup v, so that every XCD still gets work)
    const int vb = xcd_block(blockIdx.x, nwg);
    int next = (int)(((int64_t)vb * nwork) / nwg);
    const int w_hi = (int)(((int64_t)(vb + 1) * nwork) / nwg);
    for (;;) {
        int w;
        if (persistent) {
            if (tid == 0) s_tile = queue_next(queue, nwork);
            __syncthreads();
            w = s_tile;  // read by every thread before the tile's closing barrier
            if (w < 0) break;
        } else {
            if (next >= w_hi) break;
            w = next++;
        }
        const int tq = w / csplit, part = w - tq * csplit;
        const int t = t_begin + tq;
        const int c_lo = csplit == 1 ? 0 : part, c_hi = csplit == 1 ? 3 : part + 1;
        int lo = seg_lo, hi = seg_hi;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (order_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int seg = order_seg[lo];
        const int pair = seg / kScales, s = seg - pair * kScales;
        const PairDev& P = pairs[pair];
        TileGeo T;
        T.W = P.sw[s];
        T.H = P.sh[s];
        T.pitch = P.pitch[s];
        T.pe = P.plane_elems[s];
        const int ntx = (T.W + TS - 1) / TS;
        const int local = t - order_start[lo];
        const int ty = local / ntx, tx = local - ty * ntx;
        T.x0 = tx * TS;
        T.y0 = ty * kMxTileH;
        T.g = g;
        T.j = j;
        T.wave = wave;
        const int xs0 = T.x0 + 16 * wave;  // first output column of this wave's strip
        T.xo = xs0 + j;                    // this lane's output column
        // this lane's eight input columns xi .. xi + 7 (xi a multiple of 8): a
        // group that starts in the image is read in place, past W up to the
        // row's pitch (a multiple of 64 elements); a group wholly outside it
        // is read at column 0 and masked
        const int xi = xs0 - 8 + 8 * g;
        const bool gin = xi >= 0 && xi < T.W;
        T.xl = gin ? xi : 0;
        T.cmask = gin ? (int)((1u << min(T.W - xi, 8)) - 1u) : 0;
        const ST* base1 = reinterpret_cast<const ST*>(arena + P.arena_off + P.scale_off[s]);
        const ST* base2 = reinterpret_cast<const ST*>(arena + P.arena_off + P.img_bytes + P.scale_off[s]);
        // the 80 x 144 input region in the image: no masks, constant renormalisation
        const bool interior = T.x0 >= 8 && T.x0 + kMxTile + 8 <= T.W && T.y0 >= 8 && T.y0 + kMxTileH + 8 <= T.H;
        if (interior) mx_tile<SP, ST, false>(L, T, G, base1, base2, L.cum, lane, c_lo, c_hi);
        else mx_tile<SP, ST, true>(L, T, G, base1, base2, L.cum, lane, c_lo, c_hi);

        __syncthreads();
        if (tid >= 6 * c_lo && tid < 6 * c_hi) {
            const double sum = ((L.red[0][tid] + L.red[1][tid]) + L.red[2][tid]) + L.red[3][tid];
            partials[(int64_t)t * kFeat + tid] = sum;
        }
        __syncthreads();
    }
}

hipError_t launch_map_mx(bool half, bool exact, int grid, const PairDev* pairs, const int32_t* order_seg,
                         const int32_t* order_start, const BatchHdr* hdr, const void* arena, double* partials,
                         int csplit, hipStream_t st) {
    if (csplit != 1 && csplit != 3) return hipErrorInvalidValue;
    const uint8_t* a = reinterpret_cast<const uint8_t*>(arena);
    // float RGB input may hold any value: the exact split; every other format
    // gives bounded XYB values: the scaled fp16 split
    if (exact) {
        if (half) map_mx_kernel<__half, SplitBf16x3><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials, csplit);
        else map_mx_kernel<float, SplitBf16x3><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials, csplit);
    } else {
        if (half) map_mx_kernel<__half, SplitF16x2><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials, csplit);
        else map_mx_kernel<float, SplitF16x2><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials, csplit);
    }
    return hipGetLastError();
}

}  // namespace vship
