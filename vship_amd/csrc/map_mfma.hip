// Map kernel on the gfx950 matrix cores, for the FP32-moment segments.
//
// Same result as map_kernel<TH, ST, float> (renormalised 17-tap Gaussian of
// the four moment planes a, b, (x1 - x2)^2, ab, then the SSIM, artifact and
// detail-loss maps and their 1-/4-norm sums; src/ssimu2/gaussianblur.hpp:24-92,
// src/ssimu2/score.hpp:95-196), with both blur passes done as matrix products
// on v_mfma_f32_16x16x32_bf16:
//
//   row pass    Hr(16 rows x 16 cols) = In(16 rows x 32 input cols) . G(32 x 16)
//   column pass Out(16 rows x 16 cols) = G'(16 x 32 input rows) . Hr(32 rows x 16 cols)
//
// where G[k][j] = g[k - j] (zero outside 0..16) is the band of taps; both
// passes use the same per-lane tap operand. Every operand x is split exactly
// into three bf16 parts x = h + m + l (truncation: h keeps the top 8
// significant bits of x, m the next 8, l the last 8), and the six partial
// products whose weight exceeds 2^-24 are accumulated in FP32:
// hh + hm + mh + hl + mm + lh. That is FP32 accuracy with FP32 range, so the
// maps are those of the FP32 vector kernel to rounding.
//
// Tile: 64 x 64 output pixels of one scale; each of the four waves owns a
// 16-column strip (80 input rows x 32 input columns) and needs nothing from
// the other waves until the tile sums. A wave streams five 16-row blocks
// through the row pass; the row-pass results go, split into bf16 parts, to a
// two-block LDS ring of the wave, from which the column pass of output block
// ob reads blocks ob and ob + 1. Each output pixel's maps are formed in the
// registers the column pass leaves them in.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernel_util.hpp"
#include "ssimu2_dev.hpp"
#include "ssimu2_kernels.hpp"

namespace vship {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Exact split of x into three bf16 values h + m + l.
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

// Eight values as three packed bf16 operands.
struct Split8 {
    u32x4 h, m, l;
};

__device__ __forceinline__ Split8 split8(const float v[8]) {
    Split8 s;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        float h0, m0, l0, h1, m1, l1;
        split3(v[2 * k], h0, m0, l0);
        split3(v[2 * k + 1], h1, m1, l1);
        s.h[k] = pack2(h0, h1);
        s.m[k] = pack2(m0, m1);
        s.l[k] = pack2(l0, l1);
    }
    return s;
}

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                   0, 0, 0);
}

// c + a.b from the six partial products above 2^-24, smallest first.
__device__ __forceinline__ f32x4 mfma6(const Split8& a, const Split8& b, f32x4 c) {
    c = mfma(a.l, b.h, c);
    c = mfma(a.m, b.m, c);
    c = mfma(a.h, b.l, c);
    c = mfma(a.m, b.h, c);
    c = mfma(a.h, b.m, c);
    c = mfma(a.h, b.h, c);
    return c;
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

}  // namespace

template <typename ST>
__global__ __launch_bounds__(256, 2) void map_mx_kernel(const PairDev* __restrict__ pairs,
                                                        const int32_t* __restrict__ order_seg,
                                                        const int32_t* __restrict__ order_start,
                                                        const BatchHdr* __restrict__ hdr,
                                                        const uint8_t* __restrict__ arena,
                                                        double* __restrict__ partials) {
    constexpr int T = kMxTile;
    constexpr int grp = map_group(0, sizeof(ST) == 2);
    // ring[wave][slot][moment][part][column][8 words = 16 bf16 rows]; the two
    // 4-word halves of a column are swapped for columns with swz = 4, so that
    // the 8- and 16-lane groups of one LDS access (columns j .. j+7 or j+15 at
    // a stride of 8 words) fall on distinct banks
    __shared__ uint32_t ring[4][2][4][3][16][8];
    __shared__ double red[4][kFeat];
    __shared__ float cum[18];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int g = lane >> 4, j = lane & 15;
    const int swz = (((j >> 2) ^ (j >> 3)) & 1) << 2;
    if (tid == 0) {
        float a = 0.f;
        cum[0] = 0.f;
        for (int k = 0; k < 17; k++) {
            a += tap(k);
            cum[k + 1] = a;
        }
    }
    // tap operand (row pass B, column pass A): element e of lane (g, j) is g[8g + e - j]
    Split8 G;
    {
        float tv[8];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int k = 8 * g + e - j;
            tv[e] = (k >= 0 && k <= 16) ? tap(k) : 0.f;
        }
        G = split8(tv);
    }
    __syncthreads();

    int seg_lo = 0, t_begin = 0;
#pragma unroll
    for (int q = 0; q < grp; q++) {
        seg_lo += hdr->nseg[q];
        t_begin += hdr->tiles[q];
    }
    const int seg_hi = seg_lo + hdr->nseg[grp] - 1;
    const int t_end = t_begin + hdr->tiles[grp];
    const int nwg = gridDim.x;
    const int vb = xcd_block(blockIdx.x, nwg);
    const int ntl = t_end - t_begin;
    const int t_lo = t_begin + (int)(((int64_t)vb * ntl) / nwg);
    const int t_hi = t_begin + (int)(((int64_t)(vb + 1) * ntl) / nwg);

    for (int t = t_lo; t < t_hi; t++) {
        int lo = seg_lo, hi = seg_hi;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (order_start[mid] <= t) lo = mid;
            else hi = mid - 1;
        }
        const int seg = order_seg[lo];
        const int pair = seg / kScales, s = seg - pair * kScales;
        const PairDev& P = pairs[pair];
        const int W = P.sw[s], H = P.sh[s];
        const int64_t pitch = P.pitch[s];
        const int ntx = (W + T - 1) / T;
        const int local = t - order_start[lo];
        const int ty = local / ntx, tx = local - ty * ntx;
        const int x0 = tx * T, y0 = ty * T;
        const ST* base1 = reinterpret_cast<const ST*>(arena + P.arena_off + P.scale_off[s]);
        const ST* base2 = reinterpret_cast<const ST*>(arena + P.arena_off + P.img_bytes + P.scale_off[s]);
        const int64_t pe = P.plane_elems[s];

        const int xs0 = x0 + 16 * wave;       // first output column of this wave's strip
        const int xo = xs0 + j;               // this lane's output column
        const int xi = xs0 - 8 + 8 * g;       // first of this lane's eight input columns
        const bool xin = (xs0 - 8 >= 0) && (xs0 + 24 <= W);  // every input column of the strip in the image
        const float invh = __builtin_amdgcn_rcpf(tap_total(cum, xo, W));
        // renormalisation of this lane's 16 output pixels (rows y0 + 16 ob + 4g + r), channel independent
        float inv[4][4];
#pragma unroll
        for (int ob = 0; ob < 4; ob++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                inv[ob][r] = invh * __builtin_amdgcn_rcpf(tap_total(cum, min(y0 + 16 * ob + 4 * g + r, H - 1), H));

#pragma unroll 1
        for (int c = 0; c < 3; c++) {
            const ST* p1 = base1 + c * pe;
            const ST* p2 = base2 + c * pe;
            const int cxp = min(x0 + T / 2, W - 1), cyp = min(y0 + T / 2, H - 1);
            const float cc = ldf(p1 + (int64_t)cyp * pitch + cxp);
            float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};

#pragma unroll 1
            for (int mb = 0; mb < 5; mb++) {
                // ---- row pass of input rows y0 - 8 + 16 mb .. +15; this lane: row j, columns xi .. xi + 7
                const int yr = y0 - 8 + 16 * mb + j;
                float va[8], vb2[8], ve[8], vq[8];
                {
                    float u1[8], u2[8];
                    bool in[8];
                    const bool row_ok = yr >= 0 && yr < H;
                    if (row_ok && xin) {
                        Load8<ST>::run(p1 + (int64_t)yr * pitch + xi, u1);
                        Load8<ST>::run(p2 + (int64_t)yr * pitch + xi, u2);
#pragma unroll
                        for (int e = 0; e < 8; e++) in[e] = true;
                    } else {
                        const int yc = min(max(yr, 0), H - 1);
#pragma unroll
                        for (int e = 0; e < 8; e++) {
                            const int x = xi + e;
                            in[e] = row_ok && x >= 0 && x < W;
                            const int xc = min(max(x, 0), W - 1);
                            u1[e] = ldf(p1 + (int64_t)yc * pitch + xc);
                            u2[e] = ldf(p2 + (int64_t)yc * pitch + xc);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 8; e++) {
                        const float a = in[e] ? u1[e] - cc : 0.f;
                        const float b = in[e] ? u2[e] - cc : 0.f;
                        const float dd = in[e] ? u1[e] - u2[e] : 0.f;
                        va[e] = a;
                        vb2[e] = b;
                        ve[e] = dd * dd;
                        vq[e] = a * b;
                    }
                }
                const f32x4 z = {0.f, 0.f, 0.f, 0.f};
                f32x4 hr[4];
                hr[0] = mfma6(split8(va), G, z);
                hr[1] = mfma6(split8(vb2), G, z);
                hr[2] = mfma6(split8(ve), G, z);
                hr[3] = mfma6(split8(vq), G, z);
                // lane (g, j) holds rows 4g .. 4g+3 of the block, column j: store the bf16 parts
                const int slot = mb & 1;
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    float h[4], md[4], l[4];
#pragma unroll
                    for (int r = 0; r < 4; r++) split3(hr[m][r], h[r], md[r], l[r]);
                    *reinterpret_cast<uint2*>(&ring[wave][slot][m][0][j][(2 * g) ^ swz]) =
                        make_uint2(pack2(h[0], h[1]), pack2(h[2], h[3]));
                    *reinterpret_cast<uint2*>(&ring[wave][slot][m][1][j][(2 * g) ^ swz]) =
                        make_uint2(pack2(md[0], md[1]), pack2(md[2], md[3]));
                    *reinterpret_cast<uint2*>(&ring[wave][slot][m][2][j][(2 * g) ^ swz]) =
                        make_uint2(pack2(l[0], l[1]), pack2(l[2], l[3]));
                }
                wave_lds_sync();
                if (mb == 0) continue;

                // ---- column pass of output rows y0 + 16 ob .. +15 from ring blocks ob, ob + 1
                const int ob = mb - 1;
                const int rs = (ob + (g >> 1)) & 1, rw = (4 * (g & 1)) ^ swz;
                f32x4 o[4];
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    Split8 B;
                    B.h = *reinterpret_cast<const u32x4*>(&ring[wave][rs][m][0][j][rw]);
                    B.m = *reinterpret_cast<const u32x4*>(&ring[wave][rs][m][1][j][rw]);
                    B.l = *reinterpret_cast<const u32x4*>(&ring[wave][rs][m][2][j][rw]);
                    o[m] = mfma6(G, B, z);
                }
                wave_lds_sync();
                // ---- maps of this lane's output pixels: rows y0 + 16 ob + 4g + r, column xo
                float ivr[4];
#pragma unroll
                for (int r = 0; r < 4; r++) ivr[r] = ob == 0 ? inv[0][r] : ob == 1 ? inv[1][r] : ob == 2 ? inv[2][r] : inv[3][r];
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int y = y0 + 16 * ob + 4 * g + r;
                    if (xo < W && y < H) {
                        const float iv = ivr[r];
                        const float mu1 = o[0][r] * iv, mu2 = o[1][r] * iv;
                        const float E2 = o[2][r] * iv, Q = o[3][r] * iv;
                        const float md = mu1 - mu2;
                        const float md2 = md * md;
                        const float vd = E2 - md2;  // var(x1 - x2)
                        const float num_s = fmaf(2.f, Q - mu1 * mu2, 0.0009f);
                        float d = fmaf(md2, num_s, vd) * __builtin_amdgcn_rcpf(vd + num_s);
                        d = d > 0.f ? d : 0.f;
                        const float a = ldf(p1 + (int64_t)y * pitch + xo) - cc;
                        const float b = ldf(p2 + (int64_t)y * pitch + xo) - cc;
                        const float fa = fabsf(a - mu1) + 1.f, fb = fabsf(b - mu2) + 1.f;
                        const float ev = (fb - fa) * __builtin_amdgcn_rcpf(fa);  // src/ssimu2/score.hpp:150-154
                        const float art = ev > 0.f ? ev : 0.f, det = ev < 0.f ? -ev : 0.f;
                        const float d2 = d * d, a2 = art * art, t2 = det * det;
                        acc[0] += d;
                        acc[1] += d2 * d2;
                        acc[2] += art;
                        acc[3] += a2 * a2;
                        acc[4] += det;
                        acc[5] += t2 * t2;
                    }
                }
            }
            // wave sums of this channel in FP64, fixed order
#pragma unroll
            for (int k = 0; k < 6; k++) {
                double x = (double)acc[k];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                if (lane == 0) red[wave][c * 6 + k] = x;
            }
        }
        __syncthreads();
        if (tid < kFeat) {
            const double sum = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
            partials[(int64_t)t * kFeat + tid] = sum;
        }
        __syncthreads();
    }
}

hipError_t launch_map_mx(bool half, int grid, const PairDev* pairs, const int32_t* order_seg,
                         const int32_t* order_start, const BatchHdr* hdr, const void* arena, double* partials,
                         hipStream_t st) {
    const uint8_t* a = reinterpret_cast<const uint8_t*>(arena);
    if (half) map_mx_kernel<__half><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials);
    else map_mx_kernel<float><<<grid, 256, 0, st>>>(pairs, order_seg, order_start, hdr, a, partials);
    return hipGetLastError();
}

}  // namespace vship
