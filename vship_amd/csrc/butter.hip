// Butteraugli on MI355X (gfx950).
//
// The metric is the reference's src/butter/ (main.hpp:22-195): linear RGB at
// full and half scale, opsin dynamics, four frequency bands, Malta line
// differences, L2 terms, a psycho-visual mask, a blend of the two scales and
// three norms. The arithmetic is FP32, as there.
//
// Layout and kernels. Every intermediate is an FP32 plane in one arena, per
// pair and scale: 47 planes of w x h (rows packed, planes 64-float aligned).
// The reference runs about 60 launches per scale, each a full pass over one
// or two planes; here one scale is six launches, each a 2-D tile kernel that
// stages its input with halo in LDS and fuses everything pointwise around
// its stencil:
//
//   opsin   sigma 1.2 blur of linear RGB + opsin dynamics       -> XYB
//   low     sigma 7.156 blur of XYB                             -> low band
//                                                                  (as values), XYB - low
//   mid     sigma 3.225 blur of XYB - low, range operations,
//           X high band suppressed by Y                          -> middle band, high residual
//   high    sigma 1.564 blur of the residual, range operations   -> high, ultra-high bands
//   diff    the six Malta line terms from LDS tiles with a 4-pixel
//           halo, L2 and asymmetric L2 terms, mask inputs        -> AC sum, DC sum, 2 masks
//   mask    sigma 2.7 blur of both masks, fuzzy erosion, final
//           map; at full scale also the blend with the half-scale
//           map and the norm partial sums                        -> map, partials
//
// Both images of a pair run in one launch (grid z = pair x image), and a
// batch of pairs in the same launches. The half scale runs first, so that
// the full scale's mask kernel blends and reduces in its own pass.
// Separable blurs divide each pass by the sum of the taps that fall inside
// the image, as the reference does (gaussianblur.hpp:55-85): the per-column
// and per-row factors come from the running tap sums.
#include "butter.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.hpp"  // VshipErr, HIP_OR_THROW
#include "yuv_dev.hpp"

namespace vship {
namespace bt {

namespace {

constexpr int kThreads = 256;
constexpr int kSigmaCount = 5;
// src/butter/gaussianblur.hpp:27-42: window max(floor(3 sigma), 8)
constexpr double kSigma[kSigmaCount] = {1.2, 1.56416327805, 2.7, 3.22489901262, 7.15593339443};
constexpr int kRad[kSigmaCount] = {8, 8, 8, 9, 21};
enum : int { SG_OPSIN = 0, SG_HIGH = 1, SG_MASK = 2, SG_MID = 3, SG_LOW = 4 };

// Work planes of one pair at one scale. Groups of three are [image][X, Y, B],
// groups of two [image][X, Y].
enum : int {
    P_LIN = 0,   // linear RGB
    P_XYB = 6,   // opsin XYB
    P_LFV = 12,  // low band, as values
    P_T = 18,    // XYB - low band
    P_MF = 24,   // middle band
    P_HFR = 30,  // high residual (before its blur)
    P_HF = 34,   // high band
    P_UHF = 38,  // ultra-high band
    P_ACS = 42,  // AC sum (all channels) without the mask term
    P_DCS = 43,  // DC sum
    P_M0 = 44,   // mask input of the reference image
    P_M1 = 45,   // mask input of the distorted image
    P_MAP = 46,  // distortion map (half scale)
    P_COUNT = 47
};

struct Scale {
    float* base;
    int w, h;
    int pitch;      // floats between rows (a multiple of 64)
    int64_t plane;  // floats between planes
    int64_t pair;   // floats between pairs
};

__device__ __forceinline__ float* plane_of(const Scale& S, int pair, int p) {
    return S.base + (int64_t)pair * S.pair + (int64_t)p * S.plane;
}

// plane_of for a pair and plane that are the same for the whole workgroup,
// held in scalar registers: without the read-first-lane the compiler folds
// the plane offset into every per-lane address as 64-bit vector arithmetic
// (the offset, not the pointer, goes through it, so that the pointer stays
// derived from the kernel argument and its loads stay global, not flat)
__device__ __forceinline__ float* plane_u(const Scale& S, int pair, int p) {
    const uint64_t v = (uint64_t)((int64_t)pair * S.pair + (int64_t)p * S.plane);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return S.base + (int64_t)(((uint64_t)hi << 32) | lo);
}

// Element off of a plane whose base is the same for the whole workgroup: a
// 32-bit byte offset from that base, which the compiler emits as the
// scalar-base + 32-bit vector-offset address form, with no 64-bit vector
// address arithmetic. Planes hold fewer than 2^30 floats (the engine checks).
__device__ __forceinline__ float& at(float* base, int off) {
    return *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + (uint64_t)((uint32_t)off * 4u));
}
__device__ __forceinline__ float at(const float* base, int off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (uint64_t)((uint32_t)off * 4u));
}

// Tile of this workgroup, XCD-aware: hardware workgroup b (in launch order
// x fastest, then y, then z) runs on XCD b % 8, so tiles are numbered such
// that each XCD takes one contiguous range of tile rows, and the halos that
// neighbouring tiles share are read through that XCD's L2 once.
struct TileIdx {
    int x, y, z;
};
__device__ __forceinline__ TileIdx tile_of_block() {
    const int nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
    const int b = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const int v = xcd_block(b, nx * ny * nz);
    // the divisions run on the vector unit; the tile is the workgroup's, so
    // it is moved to scalar registers, and the plane addresses built from it
    // stay scalar
    TileIdx t;
    t.x = __builtin_amdgcn_readfirstlane(v % nx);
    t.y = __builtin_amdgcn_readfirstlane((v / nx) % ny);
    t.z = __builtin_amdgcn_readfirstlane(v / (nx * ny));
    return t;
}

// 1 / (sum of the taps inside [0, n)) at position p in [0, n) for a window of
// radius R (gaussianblur.hpp:63-65, 83-84).
template <int R>
__device__ __forceinline__ float inv_total(const Taps* __restrict__ T, int p, int n) {
    const int lo = max(0, R - p), hi = min(2 * R, n - 1 - p + R);
    return 1.0f / (T->cum[hi + 1] - T->cum[lo]);
}

// Column and row factors of an output region [x0, x0 + OW) x [y0, y0 + OH):
// 0 outside the image.
template <int R, int OW, int OH>
__device__ __forceinline__ void edge_factors(float* colf, float* rowf, const Taps* __restrict__ T, int w, int h,
                                             int x0, int y0) {
    static_assert(OW <= 128 && OH <= 128, "factor threads");
    const int t = threadIdx.x;
    if (t < OW) {
        const int gx = x0 + t;
        colf[t] = (gx >= 0 && gx < w) ? inv_total<R>(T, gx, w) : 0.0f;
    } else if (t >= 128 && t < 128 + OH) {
        const int gy = y0 + t - 128;
        rowf[t - 128] = (gy >= 0 && gy < h) ? inv_total<R>(T, gy, h) : 0.0f;
    }
}

// ---------------------------------------------------------------------------
// point operations (src/butter/colors.hpp, separatefrequencies.hpp:3-50,
// maskPsycho.hpp, combineMasks.hpp)
// ---------------------------------------------------------------------------
// sign(a) |a|^2.4 (colors.hpp:7-13) as 2^(2.4 log2|a|) on the hardware
// exp2/log2 units: within a few ulp of powf, which costs an order of
// magnitude more instructions. |a| is clamped to 1e-30, whose power
// underflows to exactly 0, so 0 needs no branch.
__device__ __forceinline__ float linearize(float a) {
    const float m = fmaxf(fabsf(a), 1e-30f);
    return copysignf(__builtin_amdgcn_exp2f(2.4f * __builtin_amdgcn_logf(m)), a);
}

// 19.245 ln(v + 9.971) - 23.160 with the logarithm on the hardware log2 unit
// (1 ulp; its argument is above 11) and ln 2 folded into the factor
__device__ __forceinline__ float opsin_gamma(float v) {
    constexpr float kMul = (float)(19.245013259874995 * 0.69314718055994531);
    return fmaf(kMul, __builtin_amdgcn_logf(v + 9.9710635769299145f), -23.16046239805755f);
}

__device__ __forceinline__ void opsin_absorbance(float a0, float a1, float a2, float& o0, float& o1, float& o2) {
    o0 = fmaf(0.29956550340058319f, a0, fmaf(0.63373087833825936f, a1, fmaf(0.077705617820981968f, a2, 1.7557483643287353f)));
    o1 = fmaf(0.22158691104574774f, a0, fmaf(0.69391388044116142f, a1, fmaf(0.0987313588422f, a2, 1.7557483643287353f)));
    o2 = fmaf(0.02f, a0, fmaf(0.02f, a1, fmaf(0.20480129041026129f, a2, 12.226454707163354f)));
}

__device__ __forceinline__ float remove_range(float x, float w) { return x > w ? x - w : (x < -w ? x + w : 0.0f); }
__device__ __forceinline__ float amplify_range(float x, float w) { return x > w ? x + w : (x < -w ? x - w : 2.0f * x); }
__device__ __forceinline__ float maximum_clamp(float v, float maxval) {
    const float mul = 0.688059627878f;
    if (v >= maxval) return (v - maxval) * mul + maxval;
    if (v < -maxval) return (v + maxval) * mul - maxval;
    return v;
}

// mask input: amplitude of the high bands, compressed (maskinit + diffPrecompute)
__device__ __forceinline__ float mask_input(float hfx, float uhfx, float hfy, float uhfy) {
    const float xd = (uhfx + hfx) * 2.5f;
    const float yd = uhfy * 0.4f + hfy * 0.4f;
    // square roots of non-negative values on the hardware unit (1 ulp)
    const float m = __builtin_amdgcn_sqrtf(xd * xd + yd * yd);
    const float mul = 6.19424080439f, bias = mul * 12.61050594197f;
    return __builtin_amdgcn_sqrtf(mul * fabsf(m) + bias) - __builtin_amdgcn_sqrtf(bias);
}

constexpr double kMaskNorm = 1.0 / (double)(0.79079917404f * 17.83f);

__device__ __forceinline__ float mask_y(float d) {
    // hardware reciprocal (1 ulp) of a denominator above 0.8
    const float c = 2.5485944793f * __builtin_amdgcn_rcpf((0.451936922203f * d) + 0.829591754942f);
    const float r = (float)((1.0 + (double)c) * kMaskNorm);
    return r * r;
}
__device__ __forceinline__ float mask_dc_y(float d) {
    const float c = 0.505054525019f * __builtin_amdgcn_rcpf((3.87449418804f * d) + 0.20025578522f);
    const float r = (float)((1.0 + (double)c) * kMaskNorm);
    return r * r;
}

// a * b rounded to FP32 on its own: never contracted into a following add
// (where it would be kept exact), so that two values computed alike compare
// equal however they are used
__device__ __forceinline__ float mul_rounded(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}

// Insert v into the three smallest values m0 <= m1 <= m2 (StoreMin3,
// maskPsycho.hpp:42-56), as a min/max network: the same three values for any
// non-NaN v as the reference's branches, which the compiler would otherwise
// lower to a private array written at a computed slot.
__device__ __forceinline__ void store_min3(float v, float& m0, float& m1, float& m2) {
    const float n2 = fminf(m2, fmaxf(m1, v));
    const float n1 = fminf(m1, fmaxf(m0, v));
    m0 = fminf(m0, v);
    m1 = n1;
    m2 = n2;
}

// ---------------------------------------------------------------------------
// input: sample, linearise, 2x2 mean (Planed.hpp, colors.hpp:7-21,
// downupsample.hpp:3-31). One thread per half-scale pixel writes the up to
// four full-scale pixels it covers. RGB samples are read as the reference's
// convertPointer does (src/util/float3operations.hpp:197-218). YUV is
// converted to R'G'B' first, with the front end of the SSIMULACRA2 prep kernel:
// limited range, the pair's matrix, Catmull-Rom chroma at the pair's chroma
// location.
// ---------------------------------------------------------------------------
template <int FMT>
__device__ __forceinline__ float rgb_in(const uint8_t* __restrict__ base, int64_t stride, int x, int y) {
    const uint8_t* row = base + (int64_t)y * stride;
    if constexpr (FMT == IN_RGB_U8) return (float)as_global(row)[x] / 255.0f;
    else if constexpr (FMT == IN_RGB_U16) return (float)as_global(reinterpret_cast<const uint16_t*>(row))[x] / 65535.0f;
    else if constexpr (FMT == IN_RGB_F16)
        return __half2float(__ushort_as_half(as_global(reinterpret_cast<const unsigned short*>(row))[x]));
    else return as_global(reinterpret_cast<const float*>(row))[x];
}

// R'G'B' of luma pixel (x, y) of a YUV image (container WIDE), as prep_quads'
// generic chroma path computes it
template <bool WIDE>
__device__ __forceinline__ void yuv_rgb(const uint8_t* const* src, const int64_t* stride, int w, int h, int csw,
                                        int csh, const PairDesc& P, int x, int y, float rgb[3]) {
    const int maxcode = WIDE ? (1 << P.bits) - 1 : 255;
    const float sc = (float)(1 << (P.bits - 8));
    const float ymul = 1.0f / (219.0f * sc), yadd = -16.0f / 219.0f;
    const float cmul1 = 1.0f / (224.0f * sc), cadd = -128.0f / 224.0f;
    const float kr = P.kr, kb = P.kb, ikg = 1.0f / (1.0f - P.kr - P.kb);
    const bool hc = cloc_hcentred(P.cloc), vc = cloc_vcentred(P.cloc);
    const int cw = (w + (1 << csw) - 1) >> csw, chh = (h + (1 << csh) - 1) >> csh;
    const float yn = fmaf(sample_u(src[0], stride[0], x, y, WIDE, maxcode), ymul, yadd);
    const float cb = fmaf(chroma_at<WIDE>(src[1], stride[1], cw, chh, x, y, csw, csh, hc, vc, maxcode), cmul1, cadd);
    const float cr = fmaf(chroma_at<WIDE>(src[2], stride[2], cw, chh, x, y, csw, csh, hc, vc, maxcode), cmul1, cadd);
    const float R = fmaf(2.0f * (1.0f - kr), cr, yn);
    const float B = fmaf(2.0f * (1.0f - kb), cb, yn);
    rgb[0] = R;
    rgb[1] = (yn - kr * R - kb * B) * ikg;
    rgb[2] = B;
}

template <int FMT>
__global__ __launch_bounds__(kThreads) void butter_input_kernel(const PairDesc* __restrict__ pairs, Scale S0,
                                                                Scale S1) {
    const int pair = blockIdx.z;
    const int hx = blockIdx.x * 16 + (threadIdx.x & 15), hy = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (hx >= S1.w || hy >= S1.h) return;
    const int w = S0.w, h = S0.h;
    const bool xo = 2 * hx + 1 < w, yo = 2 * hy + 1 < h;
    const int x0 = 2 * hx, y0 = 2 * hy, x1 = xo ? x0 + 1 : x0, y1 = yo ? y0 + 1 : y0;
    const PairDesc& P = pairs[pair];
    for (int im = 0; im < 2; im++) {
        float v[4][3];  // (x0,y0), (x1,y0), (x0,y1), (x1,y1)
        if constexpr (!fmt_yuv(FMT)) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const uint8_t* base = P.src[im][c];
                const int64_t st = P.stride[im][c];
                v[0][c] = rgb_in<FMT>(base, st, x0, y0);
                v[1][c] = rgb_in<FMT>(base, st, x1, y0);
                v[2][c] = rgb_in<FMT>(base, st, x0, y1);
                v[3][c] = rgb_in<FMT>(base, st, x1, y1);
            }
        } else {
            constexpr bool wide = fmt_wide(FMT);
            constexpr int csw = fmt_csw(FMT), csh = fmt_csh(FMT);
            yuv_rgb<wide>(P.src[im], P.stride[im], w, h, csw, csh, P, x0, y0, v[0]);
            yuv_rgb<wide>(P.src[im], P.stride[im], w, h, csw, csh, P, x1, y0, v[1]);
            yuv_rgb<wide>(P.src[im], P.stride[im], w, h, csw, csh, P, x0, y1, v[2]);
            yuv_rgb<wide>(P.src[im], P.stride[im], w, h, csw, csh, P, x1, y1, v[3]);
        }
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float a = linearize(v[0][c]), b = linearize(v[1][c]);
            const float cc = linearize(v[2][c]), d = linearize(v[3][c]);
            float* L = plane_of(S0, pair, P_LIN + im * 3 + c);
            L[(int64_t)y0 * S0.pitch + x0] = a;
            if (xo) L[(int64_t)y0 * S0.pitch + x1] = b;
            if (yo) {
                L[(int64_t)y1 * S0.pitch + x0] = cc;
                if (xo) L[(int64_t)y1 * S0.pitch + x1] = d;
            }
            float m = 0.0f;
            m += a;
            m += cc;
            m += b;
            m += d;
            plane_of(S1, pair, P_LIN + im * 3 + c)[(int64_t)hy * S1.pitch + hx] = m * 0.25f;
        }
    }
}

// 4:2:0 with left-located chroma: 64 x 64 luma tiles, the chroma
// neighbourhood staged in LDS (as the SSIMULACRA2 prep kernel does); a
// thread converts four 2 x 2 quads per image.
template <bool WIDE>
__global__ __launch_bounds__(kThreads) void butter_input420_kernel(const PairDesc* __restrict__ pairs, Scale S0,
                                                                   Scale S1) {
    __shared__ float ch[2][kCh420][kCh420 + 1];
    const TileIdx ti = tile_of_block();
    const int pair = ti.z;
    const PairDesc& P = pairs[pair];
    const int W = S0.w, H = S0.h;
    const int x0 = ti.x * 64, y0 = ti.y * 64;
    const int maxcode = WIDE ? (1 << P.bits) - 1 : 255;
    const float sc = (float)(1 << (P.bits - 8));
    const float ymul = 1.0f / (219.0f * sc), yadd = -16.0f / 219.0f;
    const float cmul = 1.0f / (128.0f * 224.0f * sc), cadd = -128.0f / 224.0f;
    const float kr = P.kr, kb = P.kb, ikg = 1.0f / (1.0f - P.kr - P.kb);
    const float rcr = 2.0f * (1.0f - kr), bcb = 2.0f * (1.0f - kb);
    for (int im = 0; im < 2; im++) {
        const uint8_t* const* src = P.src[im];
        const int64_t* stride = P.stride[im];
        __syncthreads();
        stage_chroma420(ch, src, stride, W, H, x0, y0, WIDE, maxcode);
        __syncthreads();
#pragma unroll 1
        for (int i = 0; i < 4; i++) {
            const int q = threadIdx.x + kThreads * i;
            const int qy = q >> 5, qx = q & 31;
            const int xa = x0 + 2 * qx, ya = y0 + 2 * qy;
            if (xa >= W || ya >= H) continue;
            const bool hx = xa + 1 < W, hy = ya + 1 < H;
            const int xb = hx ? xa + 1 : xa, yb = hy ? ya + 1 : ya;
            const float yv[4] = {sample_u(src[0], stride[0], xa, ya, WIDE, maxcode),
                                 sample_u(src[0], stride[0], xa, yb, WIDE, maxcode),
                                 sample_u(src[0], stride[0], xb, ya, WIDE, maxcode),
                                 sample_u(src[0], stride[0], xb, yb, WIDE, maxcode)};
            float cb[4], cr[4];
            chroma420_quad(ch[0], qx, qy, hx, hy, cmul, cadd, cb);
            chroma420_quad(ch[1], qx, qy, hx, hy, cmul, cadd, cr);
            float lin[4][3];  // (xa,ya), (xa,yb), (xb,ya), (xb,yb)
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float yn = fmaf(yv[k], ymul, yadd);
                const float R = fmaf(rcr, cr[k], yn);
                const float B = fmaf(bcb, cb[k], yn);
                lin[k][0] = linearize(R);
                lin[k][1] = linearize((yn - kr * R - kb * B) * ikg);
                lin[k][2] = linearize(B);
            }
            // each row's two pixels as one 8-byte store (xa is even and rows are
            // padded to 64 floats: with no second column, at an odd width, the
            // second value lands in the row padding, which no kernel reads)
            const int o0 = ya * S0.pitch + xa, o1 = yb * S0.pitch + xa, oh = (ya / 2) * S1.pitch + xa / 2;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                float* L = plane_u(S0, pair, P_LIN + im * 3 + c);
                *reinterpret_cast<float2*>(&at(L, o0)) = make_float2(lin[0][c], lin[2][c]);
                if (hy) *reinterpret_cast<float2*>(&at(L, o1)) = make_float2(lin[1][c], lin[3][c]);
                // half scale: 2x2 mean, the reference's summation order
                float m = 0.0f;
                m += lin[0][c];
                m += lin[1][c];
                m += lin[2][c];
                m += lin[3][c];
                at(plane_u(S1, pair, P_LIN + im * 3 + c), oh) = m * 0.25f;
            }
        }
    }
}

template <int FMT>
void launch_input(dim3 g, hipStream_t st, const PairDesc* pairs, const Scale& S0, const Scale& S1) {
    butter_input_kernel<FMT><<<g, dim3(kThreads), 0, st>>>(pairs, S0, S1);
}

template <int... F>
void launch_input_fmt(int fmt, dim3 g, hipStream_t st, const PairDesc* pairs, const Scale& S0, const Scale& S1,
                      std::integer_sequence<int, F...>) {
    ((fmt == F ? launch_input<F>(g, st, pairs, S0, S1) : void()), ...);
}

// ---------------------------------------------------------------------------
// Blur tile kernels: 64 x 32 outputs, 256 threads; thread t owns the kRun
// consecutive outputs of tile row t / 8, columns 8 (t % 8) .. 8 (t % 8) + 7.
//
// The column pass runs first, over the 64 + 2R staged columns of the 32
// output rows; the row pass then forms the outputs. For this wider-than-tall
// tile that order costs fewer FMAs than the row pass over 32 + 2R staged
// rows: (64 + 2R) * 32 + 64 * 32 sums of 2R + 1 taps instead of
// (32 + 2R) * 64 + 64 * 32. Both passes are register-blocked so that LDS
// reads do not bound them: a column job sums 8 outputs of one column from
// 8 + 2R reads, and a row job 8 outputs of one row from 16-byte vector reads
// (rows padded to a multiple of 4 floats). Every output is the same sequence
// of FMAs over the taps as a plain convolution.
// ---------------------------------------------------------------------------
constexpr int kTW = 64, kTH = 32;
constexpr int kRun = 8;              // outputs per job of either pass
constexpr int kRuns = kTW / kRun;    // row jobs per tile row
static_assert(kTH * kRuns == kThreads, "one row job per thread");

// Row strides of the staged input and of the column-pass output are 4
// floats past a multiple of 8, so that the 16-byte accesses of the 16 lanes
// that the LDS serves together (two rows of eight runs) fall in distinct
// banks.
template <int R>
struct alignas(16) BlurLds {
    static constexpr int SW = kTW + 2 * R, SH = kTH + 2 * R;
    static constexpr int SWP = (SW + 3) / 8 * 8 + 4;  // padded row, 4 mod 8, >= SW
    float in[SH * SWP];
    float mid[kTH * SWP];  // column pass: the output rows over all staged columns
    float colf[kTW];
    float rowf[kTH];
};

// Staging of a region [gx0, gx0 + SW) x [gy0, gy0 + SH) of a plane into
// LDS rows of SWP floats, 0 outside the image. The row padding (columns SW
// .. SWP - 1) holds finite image values, or 0 in a region that crosses an
// edge; no pass uses it (the row jobs' 16-byte reads may load it). Thread
// (tx, ty) = (t % 64, t / 64) takes columns tx, tx + 64, ... of rows ty,
// ty + 4, ...; ty is the wave's, so the row part of an address is scalar.
//
// Every load is issued unconditionally, at the address clamped into the
// image (a 32-bit offset from the plane base, which the caller holds in
// scalar registers, plane_u), and stored as loaded; in a region that crosses
// an image edge each thread then overwrites its own elements that lie
// outside with 0. (Masking each value as it is loaded makes the compiler
// sink the load into a branch with its own wait, which serialises the loads.)
template <int SW, int SWP, int SH>
struct StageRegs {
    static constexpr int NC = (SWP + 63) / 64, NR = (SH + 3) / 4;  // column chunks, rows per thread
    float v[NC][NR];
};

__device__ __forceinline__ int stage_ty() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

__device__ __forceinline__ float stage_at(const float* __restrict__ src, int pitch, int h, int gy, int gxc) {
    return at(src, min(max(gy, 0), h - 1) * pitch + gxc);
}

// zero this thread's staged elements that lie outside the image or the
// region (only called for a region that crosses an image edge)
template <int SW, int SWP, int SH>
__device__ __forceinline__ void stage_clear_outside(float* __restrict__ in, int w, int h, int gx0, int gy0) {
    constexpr int NC = (SWP + 63) / 64, NR = (SH + 3) / 4;
    const int tx = threadIdx.x & 63, ty = stage_ty();
#pragma unroll
    for (int cc = 0; cc < NC; cc++) {
        const int c = cc * 64 + tx, gx = gx0 + c;
        const bool col_out = c >= SW || gx < 0 || gx >= w;
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const int r = ty + 4 * k, gy = gy0 + r;
            if (r < SH && c < SWP && (col_out || gy < 0 || gy >= h)) in[r * SWP + c] = 0.0f;
        }
    }
}

__device__ __forceinline__ bool crosses_edge(int gx0, int gy0, int sw, int sh, int w, int h) {
    return gx0 < 0 || gy0 < 0 || gx0 + sw > w || gy0 + sh > h;
}

// the loads of stage_padded, into registers (issued together, so that their
// latency overlaps whatever the caller does before stage_store)
template <int SW, int SWP, int SH>
__device__ __forceinline__ void stage_load(StageRegs<SW, SWP, SH>& S, const float* __restrict__ src, int w, int h,
                                           int pitch, int gx0, int gy0) {
    using SR = StageRegs<SW, SWP, SH>;
    const int tx = threadIdx.x & 63, ty = stage_ty();
#pragma unroll
    for (int cc = 0; cc < SR::NC; cc++) {
        const int gxc = min(max(gx0 + cc * 64 + tx, 0), w - 1);
#pragma unroll
        for (int k = 0; k < SR::NR; k++) S.v[cc][k] = stage_at(src, pitch, h, gy0 + ty + 4 * k, gxc);
    }
}

// the stores of stage_padded, for the same region as stage_load
template <int SW, int SWP, int SH>
__device__ __forceinline__ void stage_store(const StageRegs<SW, SWP, SH>& S, float* __restrict__ in, int w, int h,
                                            int gx0, int gy0) {
    using SR = StageRegs<SW, SWP, SH>;
    const int tx = threadIdx.x & 63, ty = stage_ty();
#pragma unroll
    for (int cc = 0; cc < SR::NC; cc++) {
        const int c = cc * 64 + tx;
        if (c >= SWP) continue;
#pragma unroll
        for (int k = 0; k < SR::NR; k++)
            if (ty + 4 * k < SH) in[(ty + 4 * k) * SWP + c] = S.v[cc][k];
    }
    if (crosses_edge(gx0, gy0, SW, SH, w, h)) stage_clear_outside<SW, SWP, SH>(in, w, h, gx0, gy0);
}

// in[r][c] = src(gx0 + c, gy0 + r) for c < SW, 0 outside the image (the
// row padding as above)
template <int SW, int SWP, int SH>
__device__ __forceinline__ void stage_padded(float* __restrict__ in, const float* __restrict__ src, int w, int h,
                                             int pitch, int gx0, int gy0) {
    // column chunk by column chunk, all loads of a chunk before its stores
    constexpr int NR = (SH + 3) / 4;
    const int tx = threadIdx.x & 63, ty = stage_ty();
#pragma unroll
    for (int c0 = 0; c0 < SWP; c0 += 64) {
        const int c = c0 + tx;
        const int gxc = min(max(gx0 + c, 0), w - 1);
        float v[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) v[k] = stage_at(src, pitch, h, gy0 + ty + 4 * k, gxc);
        if (c < SWP) {
#pragma unroll
            for (int k = 0; k < NR; k++)
                if (ty + 4 * k < SH) in[(ty + 4 * k) * SWP + c] = v[k];
        }
    }
    if (crosses_edge(gx0, gy0, SW, SH, w, h)) stage_clear_outside<SW, SWP, SH>(in, w, h, gx0, gy0);
}

// float4 loads of one row job: its kRun outputs read kRun + 2R inputs
template <int R>
constexpr int run_vectors() {
    return (kRun + 2 * R + 3) / 4;
}

// row job: acc[i] = sum_k g[k] p[i + k], i < kRun, from 16-byte reads of p
// (16-byte aligned; 4 * run_vectors<R>() floats are read)
template <int R>
__device__ __forceinline__ void hrun(const float* __restrict__ p, const Taps* __restrict__ T, float (&acc)[kRun]) {
    constexpr int NV = run_vectors<R>();
    const float4* p4 = reinterpret_cast<const float4*>(p);
    float v[4 * NV];
#pragma unroll
    for (int q = 0; q < NV; q++) {
        const float4 t = p4[q];
        v[4 * q] = t.x;
        v[4 * q + 1] = t.y;
        v[4 * q + 2] = t.z;
        v[4 * q + 3] = t.w;
    }
#pragma unroll
    for (int i = 0; i < kRun; i++) acc[i] = 0.0f;
#pragma unroll
    for (int k = 0; k <= 2 * R; k++) {
        const float g = T->g[k];
#pragma unroll
        for (int i = 0; i < kRun; i++) acc[i] = fmaf(g, v[i + k], acc[i]);
    }
}

// row pass over SH staged rows (stride SWP, 16-byte aligned) into OW
// columns (stride MW): mid[r][c] = colf[c] * sum_k g[k] in[r][c + k], in
// jobs of 8 consecutive outputs
template <int R, int OW, int SWP, int MW, int SH>
__device__ __forceinline__ void hpass_blocked(const float* __restrict__ in, float* __restrict__ mid,
                                              const float* __restrict__ colf, const Taps* __restrict__ T) {
    constexpr int RUNS = OW / kRun;
    static_assert(OW % kRun == 0 && SWP % 4 == 0 && MW % 4 == 0, "vector rows");
    static_assert((RUNS - 1) * kRun + 4 * run_vectors<R>() <= SWP, "run loads stay inside the padded row");
    for (int job = threadIdx.x; job < SH * RUNS; job += kThreads) {
        const int r = job / RUNS, c0 = (job - r * RUNS) * kRun;
        float acc[kRun];
        hrun<R>(in + r * SWP + c0, T, acc);
        float4* o = reinterpret_cast<float4*>(mid + r * MW + c0);
        o[0] = make_float4(acc[0] * colf[c0], acc[1] * colf[c0 + 1], acc[2] * colf[c0 + 2], acc[3] * colf[c0 + 3]);
        o[1] = make_float4(acc[4] * colf[c0 + 4], acc[5] * colf[c0 + 5], acc[6] * colf[c0 + 6],
                           acc[7] * colf[c0 + 7]);
    }
}

// column job: N outputs of column c, rows r0 .. r0 + N - 1 of the blur,
// from rows r0 .. r0 + N - 1 + 2R of src (stride MW: the staged input, or
// the row pass's output), not yet row-normalised
template <int R, int MW, int N>
__device__ __forceinline__ void vpass_block(const float* __restrict__ src, const Taps* __restrict__ T, int r0, int c,
                                            float (&out)[N]) {
    float v[N + 2 * R];
#pragma unroll
    for (int i = 0; i < N + 2 * R; i++) v[i] = src[(r0 + i) * MW + c];
#pragma unroll
    for (int j = 0; j < N; j++) out[j] = 0.0f;
#pragma unroll
    for (int k = 0; k <= 2 * R; k++) {
        const float g = T->g[k];
#pragma unroll
        for (int j = 0; j < N; j++) out[j] = fmaf(g, v[j + k], out[j]);
    }
}

// Output pixels of a blur-tile thread: row ty of the tile, columns
// tx .. tx + kRun - 1.
struct RunPos {
    int tx, ty;
};
__device__ __forceinline__ RunPos run_pos() { return {(int)(threadIdx.x % kRuns) * kRun, (int)threadIdx.x / kRuns}; }

// The kRun values of a thread into a plane at row gy, columns gx .. gx + 7:
// two 16-byte stores. gx is a multiple of 8 below the width and rows are
// padded to a multiple of 64 floats, so the run stays inside its row; values
// past the width land in the row padding, which no kernel reads.
__device__ __forceinline__ void store_run(float* plane, int pitch, int gx, int gy, const float (&v)[kRun]) {
    float4* o = reinterpret_cast<float4*>(&at(plane, gy * pitch + gx));
    o[0] = make_float4(v[0], v[1], v[2], v[3]);
    o[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// staging registers of a blur tile's region
template <int R>
using BlurRegs = StageRegs<BlurLds<R>::SW, BlurLds<R>::SWP, BlurLds<R>::SH>;

// Blur of the tile region whose staging loads S holds (stage_load of the
// plane at ox0 - R, oy0 - R) into out[i], the thread's run_pos() pixels; ctr[i]
// receives the unblurred value at the same pixel. Once S is in LDS, the
// staging loads of plane next (if not null) are issued into S, so that they
// are in flight during this blur's passes. The LDS is free for reuse after
// the caller's next __syncthreads (the next blur_tile begins with one).
template <int R>
__device__ __forceinline__ void blur_tile(BlurLds<R>& L, BlurRegs<R>& S, const float* __restrict__ next,
                                          const Taps* __restrict__ T, int w, int h, int pitch, int ox0, int oy0,
                                          float (&out)[kRun], float (&ctr)[kRun]) {
    constexpr int SW = BlurLds<R>::SW, SWP = BlurLds<R>::SWP;
    static_assert(SWP % 4 == 0 && (kRuns - 1) * kRun + 4 * run_vectors<R>() <= SWP,
                  "row-job loads stay inside the padded row");
    __syncthreads();
    stage_store(S, L.in, w, h, ox0 - R, oy0 - R);
    if (next) stage_load(S, next, w, h, pitch, ox0 - R, oy0 - R);
    __syncthreads();
    // column pass: job (rb, c) sums rows kRun rb .. kRun rb + kRun - 1 of staged
    // column c from staged rows kRun rb .. kRun rb + kRun - 1 + 2R
    for (int job = threadIdx.x; job < SW * (kTH / kRun); job += kThreads) {
        const int rb = job / SW, c = job - rb * SW;
        float o[kRun];
        vpass_block<R, SWP, kRun>(L.in, T, rb * kRun, c, o);
#pragma unroll
        for (int j = 0; j < kRun; j++) L.mid[(rb * kRun + j) * SWP + c] = o[j] * L.rowf[rb * kRun + j];
    }
    const RunPos p = run_pos();
#pragma unroll
    for (int i = 0; i < kRun; i++) ctr[i] = L.in[(p.ty + R) * SWP + p.tx + R + i];
    __syncthreads();
    // row pass: outputs tx .. tx + 7 of row ty from mid columns tx .. tx + 7 + 2R
    hrun<R>(L.mid + p.ty * SWP + p.tx, T, out);
#pragma unroll
    for (int i = 0; i < kRun; i++) out[i] *= L.colf[p.tx + i];
}

// sigma 1.2 blur of linear RGB + opsin dynamics (colors.hpp:50-91) -> XYB
__global__ __launch_bounds__(kThreads) void butter_opsin_kernel(Scale S, const Taps* __restrict__ taps,
                                                                float intensity) {
    constexpr int R = kRad[SG_OPSIN];
    __shared__ BlurLds<R> L;
    const Taps* T = taps + SG_OPSIN;
    const TileIdx ti = tile_of_block();
    const int pair = ti.z >> 1, im = ti.z & 1;
    const int ox0 = ti.x * kTW, oy0 = ti.y * kTH;
    edge_factors<R, kTW, kTH>(L.colf, L.rowf, T, S.w, S.h, ox0, oy0);
    float bl[3][kRun], ctr[3][kRun];
    BlurRegs<R> SR;
    stage_load(SR, plane_u(S, pair, P_LIN + im * 3), S.w, S.h, S.pitch, ox0 - R, oy0 - R);
#pragma unroll
    for (int c = 0; c < 3; c++)
        blur_tile<R>(L, SR, c < 2 ? plane_u(S, pair, P_LIN + im * 3 + c + 1) : nullptr, T, S.w, S.h, S.pitch, ox0,
                     oy0, bl[c], ctr[c]);
    const RunPos p = run_pos();
    const int gx = ox0 + p.tx, gy = oy0 + p.ty;
    if (gx >= S.w || gy >= S.h) return;
    float x[kRun], y[kRun], z[kRun];
#pragma unroll
    for (int i = 0; i < kRun; i++) {
        float b0, b1, b2;
        opsin_absorbance(bl[0][i] * intensity, bl[1][i] * intensity, bl[2][i] * intensity, b0, b1, b2);
        b0 = fmaxf(fmaxf(b0, 1.7557483643287353f), 1e-4f);
        b1 = fmaxf(fmaxf(b1, 1.7557483643287353f), 1e-4f);
        b2 = fmaxf(fmaxf(b2, 12.226454707163354f), 1e-4f);
        // b >= 1.75: the hardware reciprocal (1 ulp) for the reference's division
        const float s0 = fmaxf(opsin_gamma(b0) * __builtin_amdgcn_rcpf(b0), 1e-4f);
        const float s1 = fmaxf(opsin_gamma(b1) * __builtin_amdgcn_rcpf(b1), 1e-4f);
        const float s2 = fmaxf(opsin_gamma(b2) * __builtin_amdgcn_rcpf(b2), 1e-4f);
        float o0, o1, o2;
        opsin_absorbance(ctr[0][i] * intensity, ctr[1][i] * intensity, ctr[2][i] * intensity, o0, o1, o2);
        o0 = fmaxf(o0 * s0, 1.7557483643287353f);
        o1 = fmaxf(o1 * s1, 1.7557483643287353f);
        o2 = fmaxf(o2 * s2, 12.226454707163354f);
        x[i] = o0 - o1;
        y[i] = o0 + o1;
        z[i] = o2;
    }
    store_run(plane_u(S, pair, P_XYB + im * 3 + 0), S.pitch, gx, gy, x);
    store_run(plane_u(S, pair, P_XYB + im * 3 + 1), S.pitch, gx, gy, y);
    store_run(plane_u(S, pair, P_XYB + im * 3 + 2), S.pitch, gx, gy, z);
}

// sigma 7.156 blur: low band (stored as values, XybLowFreqToVals) and the
// residual XYB - low (separatefrequencies.hpp:142-170)
__global__ __launch_bounds__(kThreads) void butter_low_kernel(Scale S, const Taps* __restrict__ taps) {
    constexpr int R = kRad[SG_LOW];
    __shared__ BlurLds<R> L;
    const Taps* T = taps + SG_LOW;
    const TileIdx ti = tile_of_block();
    const int pair = ti.z >> 1, im = ti.z & 1;
    const int ox0 = ti.x * kTW, oy0 = ti.y * kTH;
    const RunPos p = run_pos();
    const int gx = ox0 + p.tx, gy = oy0 + p.ty;
    const bool ok = gx < S.w && gy < S.h;
    edge_factors<R, kTW, kTH>(L.colf, L.rowf, T, S.w, S.h, ox0, oy0);
    float lf[3][kRun];
    BlurRegs<R> SR;
    stage_load(SR, plane_u(S, pair, P_XYB + im * 3), S.w, S.h, S.pitch, ox0 - R, oy0 - R);
#pragma unroll
    for (int c = 0; c < 3; c++) {
        float ctr[kRun], t[kRun];
        blur_tile<R>(L, SR, c < 2 ? plane_u(S, pair, P_XYB + im * 3 + c + 1) : nullptr, T, S.w, S.h, S.pitch, ox0,
                     oy0, lf[c], ctr);
#pragma unroll
        for (int i = 0; i < kRun; i++) t[i] = ctr[i] - lf[c][i];
        if (ok) store_run(plane_u(S, pair, P_T + im * 3 + c), S.pitch, gx, gy, t);
    }
    if (!ok) return;
    float v[3][kRun];
#pragma unroll
    for (int i = 0; i < kRun; i++) {
        float b = lf[2][i] + -0.362267051518f * lf[1][i];
        v[0][i] = lf[0][i] * 33.832837186260f;
        v[1][i] = lf[1][i] * 14.458268100570f;
        v[2][i] = b * 49.87984651440f;
    }
#pragma unroll
    for (int c = 0; c < 3; c++) store_run(plane_u(S, pair, P_LFV + im * 3 + c), S.pitch, gx, gy, v[c]);
}

// sigma 3.225 blur of the residual: middle band after its range operation,
// high residual = residual - middle band, X's residual suppressed by Y's
// (separatefrequencies.hpp:160-189)
__global__ __launch_bounds__(kThreads) void butter_mid_kernel(Scale S, const Taps* __restrict__ taps) {
    constexpr int R = kRad[SG_MID];
    __shared__ BlurLds<R> L;
    const Taps* T = taps + SG_MID;
    const TileIdx ti = tile_of_block();
    const int pair = ti.z >> 1, im = ti.z & 1;
    const int ox0 = ti.x * kTW, oy0 = ti.y * kTH;
    const RunPos p = run_pos();
    const int gx = ox0 + p.tx, gy = oy0 + p.ty;
    const bool ok = gx < S.w && gy < S.h;
    edge_factors<R, kTW, kTH>(L.colf, L.rowf, T, S.w, S.h, ox0, oy0);
    float hfr[2][kRun];
    constexpr int kOrder[3] = {1, 0, 2};
    BlurRegs<R> SR;
    stage_load(SR, plane_u(S, pair, P_T + im * 3 + kOrder[0]), S.w, S.h, S.pitch, ox0 - R, oy0 - R);
#pragma unroll
    for (int o = 0; o < 3; o++) {
        const int c = kOrder[o];
        float mf[kRun], ctr[kRun];
        const float* next = o < 2 ? plane_u(S, pair, P_T + im * 3 + kOrder[(o + 1) % 3]) : nullptr;
        blur_tile<R>(L, SR, next, T, S.w, S.h, S.pitch, ox0, oy0, mf, ctr);
#pragma unroll
        for (int i = 0; i < kRun; i++) {
            const float m = mf[i];
            if (c < 2) {
                hfr[c][i] = ctr[i] - m;
                mf[i] = c == 0 ? remove_range(m, 0.29f) : amplify_range(m, 0.1f);
            }
        }
        if (ok) store_run(plane_u(S, pair, P_MF + im * 3 + c), S.pitch, gx, gy, mf);
    }
    if (!ok) return;
#pragma unroll
    for (int i = 0; i < kRun; i++) {
        const float s = 0.653020556257f, yw = 46.0f, yv = hfr[1][i];
        hfr[0][i] = hfr[0][i] * (s + (yw * (1.0f - s)) / (yw + yv * yv));
    }
    store_run(plane_u(S, pair, P_HFR + im * 2 + 0), S.pitch, gx, gy, hfr[0]);
    store_run(plane_u(S, pair, P_HFR + im * 2 + 1), S.pitch, gx, gy, hfr[1]);
}

// sigma 1.564 blur of the high residual: high and ultra-high bands
// (separatefrequencies.hpp:117-140, 190-200)
__global__ __launch_bounds__(kThreads) void butter_high_kernel(Scale S, const Taps* __restrict__ taps) {
    constexpr int R = kRad[SG_HIGH];
    __shared__ BlurLds<R> L;
    const Taps* T = taps + SG_HIGH;
    const TileIdx ti = tile_of_block();
    const int pair = ti.z >> 1, im = ti.z & 1;
    const int ox0 = ti.x * kTW, oy0 = ti.y * kTH;
    const RunPos p = run_pos();
    const int gx = ox0 + p.tx, gy = oy0 + p.ty;
    const bool ok = gx < S.w && gy < S.h;
    edge_factors<R, kTW, kTH>(L.colf, L.rowf, T, S.w, S.h, ox0, oy0);
    BlurRegs<R> SR;
    stage_load(SR, plane_u(S, pair, P_HFR + im * 2), S.w, S.h, S.pitch, ox0 - R, oy0 - R);
#pragma unroll
    for (int c = 0; c < 2; c++) {
        float hf[kRun], uhf[kRun];
        blur_tile<R>(L, SR, c < 1 ? plane_u(S, pair, P_HFR + im * 2 + c + 1) : nullptr, T, S.w, S.h, S.pitch, ox0,
                     oy0, hf, uhf);
#pragma unroll
        for (int i = 0; i < kRun; i++) {
            float hv = hf[i], uv = uhf[i];
            if (c == 0) {
                uv -= hv;
                hv = remove_range(hv, 1.5f);
                uv = remove_range(uv, 0.04f);
            } else {
                hv = maximum_clamp(hv, 28.4691806922f);
                uv -= hv;
                uv = maximum_clamp(uv, 5.19175294647f);
                uv *= 2.69313763794f;
                hv *= 2.155f;
                hv = amplify_range(hv, 0.132f);
            }
            hf[i] = hv;
            uhf[i] = uv;
        }
        if (ok) {
            store_run(plane_u(S, pair, P_HF + im * 2 + c), S.pitch, gx, gy, hf);
            store_run(plane_u(S, pair, P_UHF + im * 2 + c), S.pitch, gx, gy, uhf);
        }
    }
}

// ---------------------------------------------------------------------------
// diff: Malta line terms, L2 terms, mask inputs (main.hpp:54-110,
// maltaDiff.hpp, simplerdiff.hpp:38-101, maskPsycho.hpp:24-40)
// ---------------------------------------------------------------------------
constexpr int kDW = 64, kDH = 16, kDHalo = 4;
constexpr int kDSW = kDW + 2 * kDHalo, kDSH = kDH + 2 * kDHalo;
constexpr int kDRows = kDH / (kThreads / kDW);

// Offsets (dx, dy) of the 16 lines of each Malta filter, the same line
// geometry as MaltaUnit / MaltaUnitLF (maltaDiff.hpp:5-512). Lines of fewer
// than 9 taps are padded with a count.
struct MaltaLines {
    int n[16];
    int d[16][9][2];
};

constexpr MaltaLines kMaltaHF = {
    {9, 9, 7, 7, 9, 9, 9, 9, 7, 7, 7, 7, 8, 8, 8, 8},
    {{{-4, 0}, {-3, 0}, {-2, 0}, {-1, 0}, {0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0}},
     {{0, -4}, {0, -3}, {0, -2}, {0, -1}, {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}},
     {{-3, -3}, {-2, -2}, {-1, -1}, {0, 0}, {1, 1}, {2, 2}, {3, 3}},
     {{3, -3}, {2, -2}, {1, -1}, {0, 0}, {-1, 1}, {-2, 2}, {-3, 3}},
     {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}},
     {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}},
     {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}},
     {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}},
     {{-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {1, 1}, {1, 2}, {2, 3}},
     {{2, -3}, {1, -2}, {1, -1}, {0, 0}, {-1, 1}, {-1, 2}, {-2, 3}},
     {{-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 1}, {2, 1}, {3, 2}},
     {{3, -2}, {2, -1}, {1, -1}, {0, 0}, {-1, 1}, {-2, 1}, {-3, 2}},
     {{-4, 2}, {-3, 2}, {-2, 1}, {-1, 1}, {0, 0}, {1, 0}, {2, -1}, {3, -1}},
     {{-4, -2}, {-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 0}, {2, 1}, {3, 1}},
     {{-2, -4}, {-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}},
     {{2, -4}, {2, -3}, {1, -2}, {1, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}}},
};

constexpr MaltaLines kMaltaLF = {
    {5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5},
    {{{-4, 0}, {-2, 0}, {0, 0}, {2, 0}, {4, 0}},
     {{0, -4}, {0, -2}, {0, 0}, {0, 2}, {0, 4}},
     {{-3, -3}, {-2, -2}, {0, 0}, {2, 2}, {3, 3}},
     {{3, -3}, {2, -2}, {0, 0}, {-2, 2}, {-3, 3}},
     {{1, -4}, {1, -2}, {0, 0}, {-1, 2}, {-1, 4}},
     {{-1, -4}, {-1, -2}, {0, 0}, {1, 2}, {1, 4}},
     {{-4, -1}, {-2, -1}, {0, 0}, {2, 1}, {4, 1}},
     {{-4, 1}, {-2, 1}, {0, 0}, {2, -1}, {4, -1}},
     {{-2, -3}, {-1, -2}, {0, 0}, {1, 2}, {2, 3}},
     {{2, -3}, {1, -2}, {0, 0}, {-1, 2}, {-2, 3}},
     {{-3, -2}, {-2, -1}, {0, 0}, {2, 1}, {3, 2}},
     {{3, -2}, {2, -1}, {0, 0}, {-2, 1}, {-3, 2}},
     {{-4, 2}, {-2, 1}, {0, 0}, {2, -1}, {4, -2}},
     {{-4, -2}, {-2, -1}, {0, 0}, {2, 1}, {4, 2}},
     {{-2, -4}, {-1, -2}, {0, 0}, {1, 2}, {2, 4}},
     {{2, -4}, {1, -2}, {0, 0}, {-1, 2}, {-2, 4}}},
};

// The Malta terms of kDRows vertically adjacent pixels: a thread loads the
// (kDRows + 8) x 9 window of scaled differences around its column once and
// sums every line of every pixel from registers (the lines' taps lie within
// 4 pixels of the centre).
constexpr int kWinR = kDRows + 2 * kDHalo, kWinC = 2 * kDHalo + 1;

template <bool HF>
__device__ __forceinline__ void malta_block(const float* __restrict__ win_origin, float (&ac)[kDRows]) {
    constexpr const MaltaLines& M = HF ? kMaltaHF : kMaltaLF;
    float win[kWinR][kWinC];
#pragma unroll
    for (int r = 0; r < kWinR; r++)
#pragma unroll
        for (int c = 0; c < kWinC; c++) win[r][c] = win_origin[r * kDSW + c];
#pragma unroll
    for (int j = 0; j < kDRows; j++) {
        float res = 0.0f;
#pragma unroll
        for (int l = 0; l < 16; l++) {
            float sum = 0.0f;
#pragma unroll
            for (int k = 0; k < 9; k++)
                if (k < M.n[l]) sum += win[kDHalo + j + M.d[l][k][1]][kDHalo + M.d[l][k][0]];
            res = fmaf(sum, sum, res);
        }
        ac[j] += res;
    }
}

// One Malta term: the scaled differences of the two planes.
struct MaltaTerm {
    float n2gt, n2lt, norm1;
};

// The two weights share one hardware reciprocal (1 ulp) of
// norm1 + |a|/2 + |b|/2; the reference divides twice, and the results differ
// by rounding only.
//
// The impact term is formed without branches: with b mirrored to bm = -b when
// a < 0, the reference's two cases (maltaDiff.hpp:553-572) are the one case
// a >= 0, and too_small - bm, bm - too_big are exactly its b + too_small,
// -b - too_big (negation is exact), so the value is bit-identical.
__device__ __forceinline__ float malta_scaled_diff(float a, float b, const MaltaTerm& M) {
    const float fabs0 = fabsf(a);
    const float absval = 0.5f * fabs0 + 0.5f * fabsf(b);
    const float diff = a - b;
    const float inv = __builtin_amdgcn_rcpf(M.norm1 + absval);
    const float v = (M.n2gt * inv) * diff;
    const float scaler2 = M.n2lt * inv;
    const float too_small = 0.55f * fabs0, too_big = 1.05f * fabs0;
    const float bm = a < 0.0f ? -b : b;
    float impact = bm < too_small ? too_small - bm : (bm > too_big ? bm - too_big : 0.0f);
    impact *= scaler2;
    return diff < 0.0f ? v - impact : v + impact;
}

struct DiffParams {
    MaltaTerm malta[6];  // uhf Y, uhf X, hf Y, hf X, mf Y, mf X
    float asym_gt[2], asym_lt[2];
};

// asymmetric L2 term (simplerdiff.hpp:60-90), its cases folded as in
// malta_scaled_diff (bit-identical)
__device__ __forceinline__ float l2_asym(float a, float b, float wgt, float wlt) {
    const float diff = a - b;
    float r = wgt * diff * diff;
    const float fabs0 = fabsf(a);
    // 0.4 * |a| is a double product rounded to float in simplerdiff.hpp:69
    const float too_small = (float)(0.4 * (double)fabs0), too_big = fabs0;
    const float bm = a < 0.0f ? -b : b;
    const float v = bm < too_small ? too_small - bm : (bm > too_big ? bm - too_big : 0.0f);
    return r + wlt * v * v;
}

__global__ __launch_bounds__(kThreads) void butter_diff_kernel(Scale S, DiffParams D) {
    __shared__ float dt[kDSH * kDSW];
    const TileIdx ti = tile_of_block();
    const int pair = ti.z;
    const int ox0 = ti.x * kDW, oy0 = ti.y * kDH;
    const int tx = threadIdx.x & (kDW - 1), ty = threadIdx.x / kDW;
    const int w = S.w, h = S.h;
    float ac[kDRows];
#pragma unroll
    for (int j = 0; j < kDRows; j++) ac[j] = 0.0f;
    // (plane of image 0, of image 1, HF lines) per term
    const int src0[6] = {P_UHF + 1, P_UHF + 0, P_HF + 1, P_HF + 0, P_MF + 1, P_MF + 0};
    const int src1[6] = {P_UHF + 3, P_UHF + 2, P_HF + 3, P_HF + 2, P_MF + 4, P_MF + 3};
    // not unrolled: each term's loads stay in its own iteration (hoisted into
    // the previous term they would hold registers across its line sums)
#pragma unroll 1
    for (int t = 0; t < 6; t++) {
        const float* a = plane_u(S, pair, src0[t]);
        const float* b = plane_u(S, pair, src1[t]);
        // stage the scaled differences of the tile and its halo: all of this
        // thread's loads are issued before the first difference is formed
        // (addresses clamped into the image, so that every load is issued
        // unconditionally; values outside it are replaced by 0; the last
        // round's lanes past the region read its last row, whose lines the
        // region loads anyway)
        constexpr int NE = (kDSW * kDSH + kThreads - 1) / kThreads;
        float va[NE], vb[NE];
#pragma unroll
        for (int k = 0; k < NE; k++) {
            const int i = threadIdx.x + k * kThreads;
            const int r = i / kDSW, c = i - r * kDSW;
            const int gx = ox0 - kDHalo + c, gy = oy0 - kDHalo + min(r, kDSH - 1);
            const bool ok = i < kDSW * kDSH && gx >= 0 && gx < w && gy >= 0 && gy < h;
            const int g = min(max(gy, 0), h - 1) * S.pitch + min(max(gx, 0), w - 1);
            const float x = at(a, g), y = at(b, g);
            va[k] = ok ? x : 0.0f;
            vb[k] = ok ? y : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NE; k++) {
            const int i = threadIdx.x + k * kThreads;
            if (i < kDSW * kDSH) dt[i] = malta_scaled_diff(va[k], vb[k], D.malta[t]);
        }
        __syncthreads();
        const float* origin = dt + (kDRows * ty) * kDSW + tx;
        if (t < 2)
            malta_block<true>(origin, ac);
        else
            malta_block<false>(origin, ac);
    }
    const int gx = ox0 + tx;
    if (gx >= w) return;
#pragma unroll
    for (int j = 0; j < kDRows; j++) {
        const int gy = oy0 + kDRows * ty + j;
        if (gy >= h) continue;
        const int g = gy * S.pitch + gx;
        float acc = ac[j];
        float hf[2][2], uhf[2][2];
#pragma unroll
        for (int im = 0; im < 2; im++)
#pragma unroll
            for (int c = 0; c < 2; c++) {
                hf[im][c] = at(plane_u(S, pair, P_HF + im * 2 + c), g);
                uhf[im][c] = at(plane_u(S, pair, P_UHF + im * 2 + c), g);
            }
        acc += l2_asym(hf[0][0], hf[1][0], D.asym_gt[0], D.asym_lt[0]);
        acc += l2_asym(hf[0][1], hf[1][1], D.asym_gt[1], D.asym_lt[1]);
        const float wmf[3] = {2150.0f, 10.6195433239f, 16.2176043152f};
        const float wlf[3] = {29.2353797994f, 0.844626970982f, 0.703646627719f};
        float dc = 0.0f;
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const float dm = at(plane_u(S, pair, P_MF + c), g) - at(plane_u(S, pair, P_MF + 3 + c), g);
            acc += wmf[c] * dm * dm;
            const float dl = at(plane_u(S, pair, P_LFV + c), g) - at(plane_u(S, pair, P_LFV + 3 + c), g);
            dc += wlf[c] * dl * dl;
        }
        at(plane_u(S, pair, P_ACS), g) = acc;
        at(plane_u(S, pair, P_DCS), g) = dc;
        at(plane_u(S, pair, P_M0), g) = mask_input(hf[0][0], uhf[0][0], hf[0][1], uhf[0][1]);
        at(plane_u(S, pair, P_M1), g) = mask_input(hf[1][0], uhf[1][0], hf[1][1], uhf[1][1]);
    }
}

// ---------------------------------------------------------------------------
// mask: sigma 2.7 blur of both mask inputs, fuzzy erosion of the reference's,
// L2 term of the two, final map (maskPsycho.hpp:57-132, combineMasks.hpp);
// at full scale the blend with the half-scale map (downupsample.hpp:33-56)
// and the norm partial sums (diffnorms.hpp:48-83).
// 32 x 32 outputs; the eroded mask needs the blurred reference mask on a
// 3-pixel halo.
// ---------------------------------------------------------------------------
constexpr int kMT = 32, kMErode = 3;
// the blurred reference mask is kept on a 40 x 40 region (the tile and a
// 4-pixel frame, of which the erosion reads 3) so that both passes run in
// jobs of 8
constexpr int kMF = 4, kMB = kMT + 2 * kMF;

struct alignas(16) MaskLds {
    static constexpr int R = kRad[SG_MASK];
    static constexpr int SW = kMB + 2 * R;               // 56: staged reference mask
    static constexpr int SWP = (SW + 3) / 8 * 8 + 4;     // 60
    static constexpr int MW = kMB + 4;                   // 44
    static constexpr int SW1 = kMT + 2 * R;              // 48: staged distorted mask
    static constexpr int SWP1 = (SW1 + 3) / 8 * 8 + 4;   // 52
    static constexpr int MW1 = kMT + 4;                  // 36
    float in[SW * SWP];
    float mid[SW * MW];
    float b0[kMB * kMB];
    float colf[kMB];
    float rowf[kMB];
    double red[3][kThreads / 64];
};
static_assert(MaskLds::SW1 * MaskLds::SWP1 <= MaskLds::SW * MaskLds::SWP, "staging reuse");
static_assert(MaskLds::SW1 * MaskLds::MW1 <= MaskLds::SW * MaskLds::MW, "row-pass reuse");

struct NormArgs {
    int qnorm;
    double* partials;  // [pair][tile][3]
    int tiles;
    const PairDesc* pairs;
    const float* half_map_base;  // half-scale P_MAP of pair 0
    int64_t half_pair;           // floats between pairs in the half scale
    int half_pitch;              // floats between rows in the half scale
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

template <bool FULL>
__global__ __launch_bounds__(kThreads) void butter_mask_kernel(Scale S, const Taps* __restrict__ taps, NormArgs N) {
    using M = MaskLds;
    constexpr int R = M::R;
    __shared__ MaskLds L;
    const Taps* T = taps + SG_MASK;
    const TileIdx ti = tile_of_block();
    const int pair = ti.z;
    const int ox0 = ti.x * kMT, oy0 = ti.y * kMT;
    const int w = S.w, h = S.h;
    edge_factors<R, kMB, kMB>(L.colf, L.rowf, T, w, h, ox0 - kMF, oy0 - kMF);
    // thread: column tx, rows 4 ty .. 4 ty + 3 of the tile; the per-pixel
    // inputs of the end are loaded first, to arrive during the blurs
    constexpr int kMRows = kMT * kMT / kThreads;
    const int tx = threadIdx.x & (kMT - 1), ty = threadIdx.x / kMT;
    const int gx = ox0 + tx;
    float acs_in[kMRows], dcs_in[kMRows], half_in[kMRows];
#pragma unroll
    for (int j = 0; j < kMRows; j++) {
        const int gy = oy0 + kMRows * ty + j;
        const bool ok = gx < w && gy < h;
        const int g = gy * S.pitch + gx;
        acs_in[j] = ok ? at(plane_u(S, pair, P_ACS), g) : 0.0f;
        dcs_in[j] = ok ? at(plane_u(S, pair, P_DCS), g) : 0.0f;
        half_in[j] = (FULL && ok) ? N.half_map_base[(int64_t)pair * N.half_pair + (int64_t)(gy / 2) * N.half_pitch + gx / 2]
                                  : 0.0f;
    }
    // blurred reference mask on the 40 x 40 region
    stage_padded<M::SW, M::SWP, M::SW>(L.in, plane_u(S, pair, P_M0), w, h, S.pitch, ox0 - kMF - R, oy0 - kMF - R);
    __syncthreads();
    hpass_blocked<R, kMB, M::SWP, M::MW, M::SW>(L.in, L.mid, L.colf, T);
    // the distorted mask's loads, in flight during the column pass
    StageRegs<M::SW1, M::SWP1, M::SW1> m1_regs;
    stage_load(m1_regs, plane_u(S, pair, P_M1), w, h, S.pitch, ox0 - R, oy0 - R);
    __syncthreads();
    for (int job = threadIdx.x; job < kMB * (kMB / 8); job += kThreads) {
        const int c = job % kMB, r0 = (job / kMB) * 8;
        float o[8];
        vpass_block<R, M::MW, 8>(L.mid, T, r0, c, o);
#pragma unroll
        for (int j = 0; j < 8; j++) L.b0[(r0 + j) * kMB + c] = mul_rounded(o[j], L.rowf[r0 + j]);
    }
    __syncthreads();
    // blurred distorted mask on the 32 x 32 tile
    stage_store(m1_regs, L.in, w, h, ox0 - R, oy0 - R);
    __syncthreads();
    hpass_blocked<R, kMT, M::SWP1, M::MW1, M::SW1>(L.in, L.mid, L.colf + kMF, T);
    __syncthreads();
    float b1v[kMRows];
    vpass_block<R, M::MW1, kMRows>(L.mid, T, kMRows * ty, tx, b1v);
    double sq = 0.0, s3 = 0.0, mx = 0.0;
#pragma unroll
    for (int j = 0; j < kMRows; j++) {
        const int r = kMRows * ty + j;
        const int gy = oy0 + r;
        // rounded like b0 (stored in LDS): a contraction of the product into
        // b0 - b1 would make identical masks differ near the top and bottom
        // edges, where the row factor is not 1
        const float b1 = mul_rounded(b1v[j], L.rowf[r + kMF]);
        if (gx >= w || gy >= h) continue;
        const float* bc = L.b0 + (r + kMF) * kMB + tx + kMF;
        const float b0 = bc[0];
        float m0 = b0, m1 = 2.0f * b0, m2 = m1;
        if (gx >= kMErode) {
            store_min3(bc[-kMErode], m0, m1, m2);
            if (gy >= kMErode) store_min3(bc[-kMErode * kMB - kMErode], m0, m1, m2);
            if (gy < h - kMErode) store_min3(bc[kMErode * kMB - kMErode], m0, m1, m2);
        }
        if (gx < w - kMErode) {
            store_min3(bc[kMErode], m0, m1, m2);
            if (gy >= kMErode) store_min3(bc[-kMErode * kMB + kMErode], m0, m1, m2);
            if (gy < h - kMErode) store_min3(bc[kMErode * kMB + kMErode], m0, m1, m2);
        }
        if (gy >= kMErode) store_min3(bc[-kMErode * kMB], m0, m1, m2);
        if (gy < h - kMErode) store_min3(bc[kMErode * kMB], m0, m1, m2);
        const float mask = 0.45f * m0 + 0.3f * m1 + 0.25f * m2;
        const int g = gy * S.pitch + gx;
        const float db = b0 - b1;
        const float acs = acs_in[j] + 10.0f * db * db;
        const float dcs = dcs_in[j];
        float v = __builtin_amdgcn_sqrtf(mask_y(mask) * acs + mask_dc_y(mask) * dcs);  // of a sum >= 0, 1 ulp
        if (!FULL) {
            at(plane_u(S, pair, P_MAP), g) = v;
        } else {
            v = (float)((double)v * (1.0 - 0.3 * 0.5));
            v += 0.5f * half_in[j];
            const PairDesc& P = N.pairs[pair];
            if (P.map) ((float*)((uint8_t*)P.map + (int64_t)gy * P.map_stride))[gx] = v;
            const double dv = v;
            double pq = 1.0;
            for (int k = 0; k < N.qnorm; k++) pq *= dv;
            for (int k = 0; k < -N.qnorm; k++) pq /= dv;
            sq += pq;
            s3 += fabs(dv) * dv * dv;
            mx = fmax(mx, fabs(dv));
        }
    }
    if (FULL) {
        sq = wave_sum(sq);
        s3 = wave_sum(s3);
        mx = wave_max(mx);
        const int wv = threadIdx.x / 64;
        if ((threadIdx.x & 63) == 0) {
            L.red[0][wv] = sq;
            L.red[1][wv] = s3;
            L.red[2][wv] = mx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double a = 0.0, b = 0.0, m = 0.0;
            for (int k = 0; k < kThreads / 64; k++) {
                a += L.red[0][k];
                b += L.red[1][k];
                m = fmax(m, L.red[2][k]);
            }
            double* out = N.partials + ((int64_t)pair * N.tiles + (int64_t)ti.y * gridDim.x + ti.x) * 3;
            out[0] = a;
            out[1] = b;
            out[2] = m;
        }
    }
}

// norms of each pair from its tile partials (diffnorms.hpp:85-127)
__global__ __launch_bounds__(kThreads) void butter_norm_kernel(const double* __restrict__ partials, int tiles,
                                                               int qnorm, int64_t npix, double* __restrict__ norms) {
    __shared__ double red[3][kThreads / 64];
    const int pair = blockIdx.x;
    const double* p = partials + (int64_t)pair * tiles * 3;
    double a = 0.0, b = 0.0, m = 0.0;
    for (int i = threadIdx.x; i < tiles; i += kThreads) {
        a += p[3 * i];
        b += p[3 * i + 1];
        m = fmax(m, p[3 * i + 2]);
    }
    a = wave_sum(a);
    b = wave_sum(b);
    m = wave_max(m);
    const int wv = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 0) {
        red[0][wv] = a;
        red[1][wv] = b;
        red[2][wv] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0, sm = 0.0;
        for (int k = 0; k < kThreads / 64; k++) {
            sa += red[0][k];
            sb += red[1][k];
            sm = fmax(sm, red[2][k]);
        }
        norms[3 * pair + 0] = pow(sa / (double)npix, 1.0 / (double)qnorm);
        norms[3 * pair + 1] = pow(sb / (double)npix, 1.0 / 3.0);
        norms[3 * pair + 2] = sm;
    }
}

int cdiv(int a, int b) { return (a + b - 1) / b; }

int row_pitch(int w) { return (w + 63) / 64 * 64; }
int64_t plane_floats(int w, int h) { return (int64_t)row_pitch(w) * h; }

// Malta and asymmetric L2 weights as the reference's getdiffmap passes them
// (main.hpp:57-89) and its launchers scale them (maltaDiff.hpp:663-683,
// simplerdiff.hpp:103-107).
DiffParams diff_params() {
    DiffParams D;
    const float asym = 0.8f, sasym = std::sqrt(asym);
    const float len2 = 3.75f * 2.0f + 1.0f;
    auto term = [&](bool hf, float wgt, float wlt, float norm1) {
        const float mulli = hf ? 0.39905817637f : 0.611612573796f;
        MaltaTerm t;
        t.n2gt = mulli * std::sqrt(0.5f * wgt) / len2 * norm1;
        t.n2lt = mulli * std::sqrt(0.33f * wlt) / len2 * norm1;
        t.norm1 = norm1;
        return t;
    };
    D.malta[0] = term(true, 1.10039032555f * asym, 1.10039032555f / asym, 71.7800275169f);
    D.malta[1] = term(true, 173.5f * asym, 173.5f / asym, 5.0f);
    D.malta[2] = term(false, 18.7237414387f * sasym, 18.7237414387f / sasym, 4498534.45232f);
    D.malta[3] = term(false, 6923.99476109f * sasym, 6923.99476109f / sasym, 8051.15833247f);
    D.malta[4] = term(false, 37.0819870399f, 37.0819870399f, 130262059.556f);
    D.malta[5] = term(false, 8246.75321353f, 8246.75321353f, 1009002.70582f);
    const float wmul[2] = {400.0f, 1.50815703118f};
    for (int c = 0; c < 2; c++) {
        D.asym_gt[c] = (float)((double)(wmul[c] * asym) * 0.8);
        D.asym_lt[c] = (float)((double)(wmul[c] / asym) * 0.8);
    }
    return D;
}

}  // namespace

uint64_t Engine::pair_bytes(int width, int height) {
    const int w2 = (width - 1) / 2 + 1, h2 = (height - 1) / 2 + 1;
    return (uint64_t)P_COUNT * (plane_floats(width, height) + plane_floats(w2, h2)) * sizeof(float);
}

Engine::Engine(int width, int height, int capacity, int qnorm, float intensity, int device)
    : w_(width), h_(height), cap_(capacity), q_(qnorm), intensity_(intensity), device_(device) {
    if (width <= 0 || height <= 0 || capacity <= 0 || capacity > 4096) VSHIP_THROW(E_BadPointer);
    w2_ = (w_ - 1) / 2 + 1;
    h2_ = (h_ - 1) / 2 + 1;
    plane0_ = plane_floats(w_, h_);
    plane1_ = plane_floats(w2_, h2_);
    // the kernels address a plane with 32-bit byte offsets
    if (plane0_ >= (int64_t)1 << 30) VSHIP_THROW(E_BadPointer);
    HIP_OR_THROW(hipSetDevice(device_), E_BadDeviceArgument);
    try {
        HIP_OR_THROW(hipStreamCreate(&stream_), E_BadDeviceArgument);
        const size_t arena = (size_t)cap_ * P_COUNT * (plane0_ + plane1_) * sizeof(float);
        HIP_OR_THROW(hipMalloc(&arena_, arena), E_OutOfVRAM);
        mask_tiles_ = cdiv(w_, kMT) * cdiv(h_, kMT);
        HIP_OR_THROW(hipMalloc(&partials_d_, sizeof(double) * 3 * mask_tiles_ * cap_), E_OutOfVRAM);
        HIP_OR_THROW(hipMalloc(&norms_d_, sizeof(double) * 3 * cap_), E_OutOfVRAM);
        HIP_OR_THROW(hipHostMalloc(&norms_h_, sizeof(double) * 3 * cap_), E_OutOfRAM);
        HIP_OR_THROW(hipMalloc(&desc_d_, sizeof(PairDesc) * cap_), E_OutOfVRAM);
        HIP_OR_THROW(hipHostMalloc(&desc_h_, sizeof(PairDesc) * cap_), E_OutOfRAM);
        HIP_OR_THROW(hipMalloc(&taps_d_, sizeof(Taps) * kSigmaCount), E_OutOfVRAM);
        Taps th[kSigmaCount];
        std::memset(th, 0, sizeof(th));
        for (int i = 0; i < kSigmaCount; i++) {
            const int R = kRad[i];
            double g[43], tot = 0.0;
            for (int k = -R; k <= R; k++) tot += g[k + R] = std::exp(-(double)k * k / (2.0 * kSigma[i] * kSigma[i]));
            double run = 0.0;
            for (int k = 0; k <= 2 * R; k++) {
                th[i].cum[k] = (float)(run / tot);
                th[i].g[k] = (float)(g[k] / tot);
                run += g[k];
            }
            th[i].cum[2 * R + 1] = (float)(run / tot);
        }
        HIP_OR_THROW(hipMemcpy(taps_d_, th, sizeof(th), hipMemcpyHostToDevice), E_BadDeviceArgument);
        for (auto& e : ev_) HIP_OR_THROW(hipEventCreate(&e), E_BadDeviceArgument);
    } catch (...) {
        release();
        throw;
    }
}

Engine::~Engine() { release(); }

void Engine::release() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (arena_) (void)hipFree(arena_), arena_ = nullptr;
    if (partials_d_) (void)hipFree(partials_d_), partials_d_ = nullptr;
    if (norms_d_) (void)hipFree(norms_d_), norms_d_ = nullptr;
    if (norms_h_) (void)hipHostFree(norms_h_), norms_h_ = nullptr;
    if (desc_d_) (void)hipFree(desc_d_), desc_d_ = nullptr;
    if (desc_h_) (void)hipHostFree(desc_h_), desc_h_ = nullptr;
    if (taps_d_) (void)hipFree(taps_d_), taps_d_ = nullptr;
    if (stream_) (void)hipStreamDestroy(stream_), stream_ = nullptr;
}

void Engine::submit(int fmt, const PairDesc* pairs, int n) {
    if (n <= 0 || n > cap_ || !pairs) VSHIP_THROW(E_BadPointer);
    if (fmt < 0 || fmt >= IN_NUM_FORMATS) VSHIP_THROW(E_BadPointer);
    // bytes per sample and plane widths of the format: every row the kernel
    // reads must lie within the plane's stride
    const int bps = fmt_yuv(fmt) ? (fmt_wide(fmt) ? 2 : 1) : (fmt == IN_RGB_U8 ? 1 : fmt == IN_RGB_F32 ? 4 : 2);
    int64_t row[3];
    int rows[3];
    for (int c = 0; c < 3; c++) {
        const bool sub = c > 0 && fmt_yuv(fmt);
        const int pw = sub ? (w_ + (1 << fmt_csw(fmt)) - 1) >> fmt_csw(fmt) : w_;
        rows[c] = sub ? (h_ + (1 << fmt_csh(fmt)) - 1) >> fmt_csh(fmt) : h_;
        row[c] = (int64_t)pw * bps;
    }
    // every source plane and the distance map must lie inside a device
    // allocation: a host pointer (the reference's ABI writes its distance map
    // to host memory) is an error here, before any kernel runs
    for (int i = 0; i < n; i++) {
        for (int im = 0; im < 2; im++)
            for (int c = 0; c < 3; c++)
                require_device_rows(pairs[i].src[im][c], pairs[i].stride[im][c], rows[c], row[c]);
        if (pairs[i].map) require_device_rows(pairs[i].map, pairs[i].map_stride, h_, (int64_t)w_ * 4);
        if (fmt_yuv(fmt)) {
            const int bits = pairs[i].bits;
            if (fmt_wide(fmt) ? (bits < 9 || bits > 16) : bits != 8) VSHIP_THROW(E_BadPointer);
            if (pairs[i].cloc < 0 || pairs[i].cloc > 3) VSHIP_THROW(E_BadPointer);
        }
    }
    HIP_OR_THROW(hipSetDevice(device_), E_BadDeviceArgument);
    // the previous batch may still read the pinned descriptors
    HIP_OR_THROW(hipStreamSynchronize(stream_), E_BadDeviceArgument);
    // 4:2:0 batches whose pairs all have left-located chroma take the staged
    // input kernel
    staged420_ = fmt == IN_YUV420_U8 || fmt == IN_YUV420_U16;
    for (int i = 0; i < n && staged420_; i++) staged420_ = pairs[i].cloc == CLOC_LEFT;
    std::memcpy(desc_h_, pairs, sizeof(PairDesc) * n);
    HIP_OR_THROW(hipMemcpyAsync(desc_d_, desc_h_, sizeof(PairDesc) * n, hipMemcpyHostToDevice, stream_),
                 E_BadDeviceArgument);
    launch(fmt, n);
    last_fmt_ = fmt;
    HIP_OR_THROW(hipMemcpyAsync(norms_h_, norms_d_, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, stream_),
                 E_BadDeviceArgument);
    last_n_ = n;
}

void Engine::run(int fmt, const PairDesc* pairs, int n, double* norms) {
    submit(fmt, pairs, n);
    HIP_OR_THROW(hipStreamSynchronize(stream_), E_BadDeviceArgument);
    if (norms) std::memcpy(norms, norms_h_, sizeof(double) * 3 * n);
}

void Engine::launch(int fmt, int n) {
    const Scale S0{arena_, w_, h_, row_pitch(w_), plane0_, (int64_t)P_COUNT * plane0_};
    const Scale S1{arena_ + (int64_t)cap_ * P_COUNT * plane0_, w2_, h2_, row_pitch(w2_), plane1_,
                   (int64_t)P_COUNT * plane1_};
    const DiffParams D = diff_params();
    const bool prof = profile_;
    auto mark = [&](int st) {
        if (prof) (void)hipEventRecord(ev_[st], stream_);
    };
    // stage marks: ev_[s] before the stage's first launch; the half and full
    // scale launches of a stage are interleaved with others, so per-stage time
    // is taken by a second, stage-ordered pass only when profiling
    auto scale_pass = [&](const Scale& S, bool full, int only) {
        const dim3 blk(kThreads);
        const dim3 gt(cdiv(S.w, kTW), cdiv(S.h, kTH), 2 * n);
        if (only < 0 || only == ST_OPSIN) butter_opsin_kernel<<<gt, blk, 0, stream_>>>(S, taps_d_, intensity_);
        if (only < 0 || only == ST_LOW) butter_low_kernel<<<gt, blk, 0, stream_>>>(S, taps_d_);
        if (only < 0 || only == ST_MID) butter_mid_kernel<<<gt, blk, 0, stream_>>>(S, taps_d_);
        if (only < 0 || only == ST_HIGH) butter_high_kernel<<<gt, blk, 0, stream_>>>(S, taps_d_);
        if (only < 0 || only == ST_DIFF)
            butter_diff_kernel<<<dim3(cdiv(S.w, kDW), cdiv(S.h, kDH), n), blk, 0, stream_>>>(S, D);
        if (only < 0 || only == ST_MASK) {
            NormArgs N;
            N.qnorm = q_;
            N.partials = partials_d_;
            N.tiles = mask_tiles_;
            N.pairs = desc_d_;
            N.half_map_base = S1.base + (int64_t)P_MAP * S1.plane;
            N.half_pair = S1.pair;
            N.half_pitch = S1.pitch;
            const dim3 gm(cdiv(S.w, kMT), cdiv(S.h, kMT), n);
            if (full)
                butter_mask_kernel<true><<<gm, blk, 0, stream_>>>(S, taps_d_, N);
            else
                butter_mask_kernel<false><<<gm, blk, 0, stream_>>>(S, taps_d_, N);
        }
    };
    const dim3 gi(cdiv(w2_, 16), cdiv(h2_, 16), n);
    auto input = [&] {
        if (staged420_) {
            const dim3 g4(cdiv(w_, 64), cdiv(h_, 64), n);
            if (fmt == IN_YUV420_U16)
                butter_input420_kernel<true><<<g4, dim3(kThreads), 0, stream_>>>(desc_d_, S0, S1);
            else
                butter_input420_kernel<false><<<g4, dim3(kThreads), 0, stream_>>>(desc_d_, S0, S1);
            return;
        }
        launch_input_fmt(fmt, gi, stream_, desc_d_, S0, S1, std::make_integer_sequence<int, IN_NUM_FORMATS>{});
    };
    if (!prof) {
        input();
        scale_pass(S1, false, -1);
        scale_pass(S0, true, -1);
        butter_norm_kernel<<<dim3(n), dim3(kThreads), 0, stream_>>>(partials_d_, mask_tiles_, q_, (int64_t)w_ * h_,
                                                                    norms_d_);
        timed_ = false;
    } else {
        // stage-ordered: every stage's half- and full-scale launches back to
        // back between two events; the data dependences hold because each
        // stage of the half scale only reads half-scale planes
        mark(ST_INPUT);
        input();
        for (int st = ST_OPSIN; st <= ST_MASK; st++) {
            mark(st);
            scale_pass(S1, false, st);
            scale_pass(S0, true, st);
        }
        mark(ST_NORM);
        butter_norm_kernel<<<dim3(n), dim3(kThreads), 0, stream_>>>(partials_d_, mask_tiles_, q_, (int64_t)w_ * h_,
                                                                    norms_d_);
        mark(ST_COUNT);
        timed_ = true;
    }
    HIP_OR_THROW(hipGetLastError(), E_BadDeviceCode);
}

bool Engine::stage_ms(float* out) const {
    if (!timed_) return false;
    for (int s = 0; s < ST_COUNT; s++)
        if (hipEventElapsedTime(&out[s], ev_[s], ev_[s + 1]) != hipSuccess) return false;
    return true;
}

uint64_t Engine::stage_bytes(int stage) const {
    // planes read and written per pixel and image pair, at each scale
    // (inputs: the sampled bytes; everything else FP32)
    const double px = (double)w_ * h_ + (double)w2_ * h2_;
    const double full = (double)w_ * h_;
    double b = 0.0;
    switch (stage) {
        case ST_INPUT: {
            // samples read (chroma planes at their subsampled size), linear
            // planes written at both scales
            const int f = last_fmt_;
            const double bps = fmt_yuv(f) ? (fmt_wide(f) ? 2 : 1) : (f == IN_RGB_U8 ? 1 : f == IN_RGB_F32 ? 4 : 2);
            const double chroma = fmt_yuv(f) ? 1.0 / (double)(1 << (fmt_csw(f) + fmt_csh(f))) : 1.0;
            b = full * 2 * (1 + 2 * chroma) * bps + full * 6 * 4 + (double)w2_ * h2_ * 6 * 4;
            break;
        }
        case ST_OPSIN: b = px * (6 + 6) * 4; break;
        case ST_LOW: b = px * (6 + 12) * 4; break;
        case ST_MID: b = px * (6 + 10) * 4; break;
        case ST_HIGH: b = px * (4 + 8) * 4; break;
        case ST_DIFF: b = px * (20 + 4) * 4; break;
        case ST_MASK: b = px * (4 + 1) * 4 + (double)w2_ * h2_ * 4; break;
        default: b = 0.0;
    }
    return (uint64_t)b;
}

}  // namespace bt
}  // namespace vship
