/*
 * SSIMULACRA2 oracle (plain C, CPU). TEST INFRASTRUCTURE ONLY: see
 * ssimu2_oracle.h for scope, stage citations and what pins each function.
 */
#include "ssimu2_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------------- */
/* helpers                                                                */
/* ---------------------------------------------------------------------- */

/* IEEE binary16 -> double (input format HALF, src/util/float3operations.hpp:210-213). */
static double so_half_to_double(uint16_t u) {
    const int sign = (u >> 15) & 1, ex = (u >> 10) & 0x1f, man = u & 0x3ff;
    double v;
    if (ex == 0) v = ldexp((double)man, -24);
    else if (ex == 31) v = man ? NAN : INFINITY;
    else v = ldexp((double)(man | 0x400), ex - 25);
    return sign ? -v : v;
}

int64_t so_total_scale_size(int w, int h) {
    /* src/ssimu2/main.hpp:34-42 */
    int64_t t = 0, ww = w, hh = h;
    for (int s = 0; s < 6; s++) {
        t += ww * hh;
        ww = (ww - 1) / 2 + 1;
        hh = (hh - 1) / 2 + 1;
    }
    return t;
}

void so_scale_dims(int w, int h, int dims[12]) {
    for (int s = 0; s < 6; s++) {
        dims[2 * s] = w;
        dims[2 * s + 1] = h;
        w = (w - 1) / 2 + 1;
        h = (h - 1) / 2 + 1;
    }
}

/* g_k = exp(-k^2 / (2 sigma^2)) / sqrt(2 pi sigma^2), sigma = 1.5, k = -8..8
 * (src/ssimu2/gaussianblur.hpp:7-18, src/util/preprocessor.hpp:89-92).
 * The constant factor cancels in the renormalised blur. */
void so_gaussian_taps(double g[17]) {
    const double sigma = 1.5;
    const double norm = sqrt(2.0 * M_PI * sigma * sigma);
    for (int i = 0; i < 17; i++) {
        const double k = (double)(i - 8);
        g[i] = exp(-k * k / (2.0 * sigma * sigma)) / norm;
    }
}

/* sign(v) * |v|^2.4, src/ssimu2/makeXYB.hpp:52-64 */
double so_linearize(double v) { return v < 0 ? -pow(-v, 2.4) : pow(v, 2.4); }

static int g_threads = 0;
void so_set_threads(int n) {
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int so_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---------------------------------------------------------------------- */
/* the two precision builds of the pipeline                               */
/* ---------------------------------------------------------------------- */
#define REAL double
#define FN(x) x##_f64
#define CBRT(v) cbrt(v)
#define FABS(v) fabs(v)
#define FN_PAR ((int64_t)w * h >= 65536)
#include "ssimu2_oracle_impl.h"
#undef REAL
#undef FN
#undef CBRT
#undef FABS
#undef FN_PAR

#define REAL float
#define FN(x) x##_f32
#define CBRT(v) cbrtf(v)
#define FABS(v) fabsf(v)
#define FN_PAR ((int64_t)w * h >= 65536)
#include "ssimu2_oracle_impl.h"
#undef REAL
#undef FN
#undef CBRT
#undef FABS
#undef FN_PAR

/* ---------------------------------------------------------------------- */
/* double-precision stage entry points                                    */
/* ---------------------------------------------------------------------- */

void so_xyb(const double rgb[3], double out[3]) { xyb_px_f64(rgb[0], rgb[1], rgb[2], &out[0], &out[1], &out[2]); }

void so_downsample(const double* src, int w, int h, double* dst) { downsample_f64(src, w, h, dst); }

void so_blur(const double* src, int w, int h, double* dst) {
    double* tmp = (double*)malloc(sizeof(double) * (size_t)w * h);
    blur_f64(src, w, h, dst, tmp);
    free(tmp);
}

void so_scale_features(const double* xyb1, const double* xyb2, int w, int h, double out18[18]) {
    scale_features_f64(xyb1, xyb2, w, h, out18);
}

void so_channel_maps(const double* i1, const double* i2, int w, int h, double* dssim, double* art, double* det) {
    const int64_t n = (int64_t)w * h;
    double* buf = (double*)malloc(sizeof(double) * n * 7);
    double *m1 = buf, *m2 = buf + n, *s11 = buf + 2 * n, *s22 = buf + 3 * n, *s12 = buf + 4 * n;
    double *prod = buf + 5 * n, *tmp = buf + 6 * n;
    blur_f64(i1, w, h, m1, tmp);
    blur_f64(i2, w, h, m2, tmp);
    for (int64_t i = 0; i < n; i++) prod[i] = i1[i] * i1[i];
    blur_f64(prod, w, h, s11, tmp);
    for (int64_t i = 0; i < n; i++) prod[i] = i2[i] * i2[i];
    blur_f64(prod, w, h, s22, tmp);
    for (int64_t i = 0; i < n; i++) prod[i] = i1[i] * i2[i];
    blur_f64(prod, w, h, s12, tmp);
    for (int64_t i = 0; i < n; i++) {
        const double md = m1[i] - m2[i];
        const double num_m = 1.0 - md * md;
        const double num_s = 2.0 * (s12[i] - m1[i] * m2[i]) + 0.0009;
        const double den_s = (s11[i] - m1[i] * m1[i]) + (s22[i] - m2[i] * m2[i]) + 0.0009;
        double d = 1.0 - num_m * num_s / den_s;
        dssim[i] = d > 0 ? d : 0;
        const double v = (fabs(i2[i] - m2[i]) + 1.0) / (fabs(i1[i] - m1[i]) + 1.0) - 1.0;
        art[i] = v > 0 ? v : 0;
        det[i] = -v > 0 ? -v : 0;
    }
    free(buf);
}

int so_score_linear_f64(const double* lin1, const double* lin2, int w, int h, double* score, double* feat108) {
    return score_from_linear_f64(lin1, lin2, w, h, score, feat108);
}

/* ---------------------------------------------------------------------- */
/* final score: the 108 published SSIMULACRA2 weights and the output      */
/* polynomial, src/ssimu2/score.hpp:272-401                               */
/* ---------------------------------------------------------------------- */
static const double so_weights[108] = {
    /* X channel: per scale [1-norm ssim, artifact, detail, 4-norm ssim, artifact, detail] */
    0.0, 0.0007376606707406586, 0.0, 0.0, 0.0007793481682867309, 0.0,
    0.0, 0.0004371155730107379, 0.0, 1.1041726426657346, 0.00066284834129271, 0.00015231632783718752,
    0.0, 0.0016406437456599754, 0.0, 1.8422455520539298, 11.441172603757666, 0.0,
    0.0007989109436015163, 0.000176816438078653, 0.0, 1.8787594979546387, 10.94906990605142, 0.0,
    0.0007289346991508072, 0.9677937080626833, 0.0, 0.00014003424285435884, 0.9981766977854967, 0.00031949755934435053,
    0.0004550992113792063, 0.0, 0.0, 0.0013648766163243398, 0.0, 0.0,
    /* Y channel: per scale [1-norm ssim, artifact, detail, 4-norm ssim, artifact, detail] */
    0.0, 0.0, 0.0, 7.466890328078848, 0.0, 17.445833984131262,
    0.0006235601634041466, 0.0, 0.0, 6.683678146179332, 0.00037724407979611296, 1.027889937768264,
    225.20515300849274, 0.0, 0.0, 19.213238186143016, 0.0011401524586618361, 0.001237755635509985,
    176.39317598450694, 0.0, 0.0, 24.43300999870476, 0.28520802612117757, 0.0004485436923833408,
    0.0, 0.0, 0.0, 34.77906344483772, 44.835625328877896, 0.0,
    0.0, 0.0, 0.0, 0.0, 0.0, 0.0,
    /* B channel: per scale [1-norm ssim, artifact, detail, 4-norm ssim, artifact, detail] */
    0.0, 0.0008680556573291698, 0.0, 0.0, 0.0, 0.0,
    0.0, 0.0005313191874358747, 0.0, 0.00016533814161379112, 0.0, 0.0,
    0.0, 0.0, 0.0, 0.0004179171803251336, 0.0017290828234722833, 0.0,
    0.0020827005846636437, 0.0, 0.0, 8.826982764996862, 23.19243343998926, 0.0,
    95.1080498811086, 0.9863978034400682, 0.9834382792465353, 0.0012286405048278493, 171.2667255897307, 0.9807858872435379,
    0.0, 0.0, 0.0, 0.0005130064588990679, 0.0, 0.00010854057858411537,
};

double so_final_score(const double f[108]) {
    double s = 0.0;
    for (int i = 0; i < 108; i++) s += so_weights[i] * f[i];
    s *= 0.9562382616834844;
    s = (6.248496625763138e-5 * s * s) * s + 2.326765642916932 * s - 0.020884521182843837 * s * s;
    if (s > 0.0) return 100.0 - 10.0 * pow(s, 0.6276336467831387);
    return 100.0;
}

/* ---------------------------------------------------------------------- */
/* YUV 4:2:0 front end (definition shared with the GPU front end)         */
/*                                                                        */
/*  chroma planes are (w+1)/2 x (h+1)/2, samples u8 (bits == 8) or        */
/*  little-endian u16 (bits 9..16). Upsampling is Catmull-Rom cubic       */
/*  (cubic Hermite with central-difference tangents) with clamp-to-edge   */
/*  sample indices, as in the reference's unused GPU helper               */
/*  src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp:5-68:         */
/*   horizontal, left-sited: even x -> C[x/2];                            */
/*     odd x -> (-C[k-1] + 9C[k] + 9C[k+1] - C[k+2]) / 16, k = (x-1)/2    */
/*   vertical, centre-sited: odd y -> t=0.25 from row i=(y-1)/2:           */
/*     (-9C[i-1] + 111C[i] + 29C[i+1] - 3C[i+2]) / 128                    */
/*     even y -> t=0.75 from row i=y/2-1:                                  */
/*     (-3C[i-1] + 29C[i] + 111C[i+1] - 9C[i+2]) / 128                    */
/*  limited range: Y' = (Y - 16s)/(219s), C' = (C - 128s)/(224s),         */
/*  s = 2^(bits-8); matrix: R = Y' + 2(1-Kr)Cr', B = Y' + 2(1-Kb)Cb',      */
/*  G = (Y' - Kr R - Kb B)/(1 - Kr - Kb); no clamping of the result.      */
/*  16-bit-container samples above 2^bits - 1 are clamped to it.          */
/* ---------------------------------------------------------------------- */
static double so_sample(const uint8_t* p, int64_t stride, int bits, int x, int y) {
    if (bits <= 8) return (double)p[(int64_t)y * stride + x];
    uint16_t u;
    memcpy(&u, p + (int64_t)y * stride + 2 * (int64_t)x, 2);
    const unsigned maxcode = (1u << bits) - 1u;
    return (double)(u > maxcode ? maxcode : u);
}

static void so_matrix_coefs(int matrix, double* kr, double* kb) {
    switch (matrix) {
    case SO_MAT_BT601: *kr = 0.299; *kb = 0.114; break;
    case SO_MAT_BT2020NCL: *kr = 0.2627; *kb = 0.0593; break;
    default: *kr = 0.2126; *kb = 0.0722; break;
    }
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---------------------------------------------------------------------- */
/* Any planar YUV layout (definition shared with the GPU front end)       */
/*                                                                        */
/*  chroma planes are ceil(w / 2^sw) x ceil(h / 2^sh); chroma sample i of */
/*  an axis subsampled by f = 2^s sits at luma position f*i (co-sited) or */
/*  f*i + (f-1)/2 (centred), as the reference's upsampling kernels place  */
/*  it (src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp:44-128:   */
/*  left/top write t = 0 and 1/2 (x2) or k/4 (x4) at luma f*i + k,        */
/*  centre writes t = 1/4, 3/4 (x2) or 1/8 .. 7/8 (x4) across [i, i+1]).  */
/*  Luma position x is thus at chroma coordinate u; the value is the      */
/*  Catmull-Rom interpolant of the four samples floor(u)-1 .. floor(u)+2  */
/*  (clamped to the plane), horizontally then vertically (:131-205).      */
/*  Locations (AVChromaLocation): left = co-sited horizontally, centred   */
/*  vertically; center = centred both ways; topleft = co-sited both ways; */
/*  top = centred horizontally, co-sited vertically.                       */
/* ---------------------------------------------------------------------- */
static void so_chroma_pos(int x, int s, int centred, int* c, double* t) {
    const double f = (double)(1 << s);
    const double u = centred ? ((double)x - (f - 1.0) / 2.0) / f : (double)x / f;
    const double fl = floor(u);
    *c = (int)fl;
    *t = u - fl;
}

/* Catmull-Rom weights from the cubic Hermite basis: value at t of the
   segment [p1, p2] with tangents (p2 - p0)/2 and (p3 - p1)/2. */
static void so_cr_weights(double t, double w[4]) {
    const double t2 = t * t, t3 = t2 * t;
    const double h00 = 2.0 * t3 - 3.0 * t2 + 1.0, h10 = t3 - 2.0 * t2 + t;
    const double h01 = -2.0 * t3 + 3.0 * t2, h11 = t3 - t2;
    w[0] = -0.5 * h10;
    w[1] = h00 - 0.5 * h11;
    w[2] = h01 + 0.5 * h10;
    w[3] = 0.5 * h11;
}

double so_chroma_sample(const uint8_t* plane, int64_t stride, int w, int h, int bits, int sw, int sh, int loc,
                        int x, int y) {
    const int cw = (w + (1 << sw) - 1) >> sw, ch = (h + (1 << sh) - 1) >> sh;
    const int hc = loc == SO_LOC_CENTER || loc == SO_LOC_TOP;
    const int vc = loc == SO_LOC_LEFT || loc == SO_LOC_CENTER;
    int cx, cy;
    double tx, ty, wx[4], wy[4];
    so_chroma_pos(x, sw, hc, &cx, &tx);
    so_chroma_pos(y, sh, vc, &cy, &ty);
    so_cr_weights(tx, wx);
    so_cr_weights(ty, wy);
    double acc = 0.0;
    for (int a = 0; a < 4; a++) {
        const int yy = clampi(cy - 1 + a, 0, ch - 1);
        double hv = 0.0;
        for (int b = 0; b < 4; b++) hv += wx[b] * so_sample(plane, stride, bits, clampi(cx - 1 + b, 0, cw - 1), yy);
        acc += wy[a] * hv;
    }
    return acc;
}

int so_yuv_to_rgb_f64(const uint8_t* const p[3], const int64_t stride[3], int w, int h, int bits, int matrix,
                      int sw, int sh, int loc, double* rgb) {
    if (w <= 0 || h <= 0 || bits < 8 || bits > 16 || sw < 0 || sw > 2 || sh < 0 || sh > 1 || loc < 0 || loc > 3)
        return -1;
    const double sc = (double)(1 << (bits - 8));
    double kr, kb;
    so_matrix_coefs(matrix, &kr, &kb);
    const double kg = 1.0 - kr - kb;
    const int64_t n = (int64_t)w * h;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const double cbv = so_chroma_sample(p[1], stride[1], w, h, bits, sw, sh, loc, x, y);
            const double crv = so_chroma_sample(p[2], stride[2], w, h, bits, sw, sh, loc, x, y);
            const double Y = (so_sample(p[0], stride[0], bits, x, y) - 16.0 * sc) / (219.0 * sc);
            const double cb = (cbv - 128.0 * sc) / (224.0 * sc);
            const double cr = (crv - 128.0 * sc) / (224.0 * sc);
            const double R = Y + 2.0 * (1.0 - kr) * cr;
            const double B = Y + 2.0 * (1.0 - kb) * cb;
            const double G = (Y - kr * R - kb * B) / kg;
            const int64_t o = (int64_t)y * w + x;
            rgb[o] = R;
            rgb[n + o] = G;
            rgb[2 * n + o] = B;
        }
    return 0;
}

int so_yuv420_to_rgb_f64(const uint8_t* const p[3], const int64_t stride[3], int w, int h, int bits, int matrix,
                         double* rgb) {
    if (w <= 0 || h <= 0 || bits < 8 || bits > 16) return -1;
    const int cw = (w + 1) / 2, ch = (h + 1) / 2;
    const double sc = (double)(1 << (bits - 8));
    double kr, kb;
    so_matrix_coefs(matrix, &kr, &kb);
    const double kg = 1.0 - kr - kb;
    static const double wv_odd[4] = {-9.0 / 128, 111.0 / 128, 29.0 / 128, -3.0 / 128};
    static const double wv_even[4] = {-3.0 / 128, 29.0 / 128, 111.0 / 128, -9.0 / 128};
    static const double wh_odd[4] = {-1.0 / 16, 9.0 / 16, 9.0 / 16, -1.0 / 16};
    const int64_t n = (int64_t)w * h;
    for (int y = 0; y < h; y++) {
        const int i0 = (y & 1) ? (y - 1) / 2 : y / 2 - 1;
        const double* wv = (y & 1) ? wv_odd : wv_even;
        for (int x = 0; x < w; x++) {
            double cc[2];
            for (int pl = 1; pl <= 2; pl++) {
                double acc = 0.0;
                for (int a = 0; a < 4; a++) {
                    const int yy = clampi(i0 - 1 + a, 0, ch - 1);
                    double hv;
                    if ((x & 1) == 0) {
                        hv = so_sample(p[pl], stride[pl], bits, x / 2, yy);
                    } else {
                        const int k = (x - 1) / 2;
                        hv = 0.0;
                        for (int b = 0; b < 4; b++)
                            hv += wh_odd[b] * so_sample(p[pl], stride[pl], bits, clampi(k - 1 + b, 0, cw - 1), yy);
                    }
                    acc += wv[a] * hv;
                }
                cc[pl - 1] = acc;
            }
            const double Y = (so_sample(p[0], stride[0], bits, x, y) - 16.0 * sc) / (219.0 * sc);
            const double cb = (cc[0] - 128.0 * sc) / (224.0 * sc);
            const double cr = (cc[1] - 128.0 * sc) / (224.0 * sc);
            const double R = Y + 2.0 * (1.0 - kr) * cr;
            const double B = Y + 2.0 * (1.0 - kb) * cb;
            const double G = (Y - kr * R - kb * B) / kg;
            const int64_t o = (int64_t)y * w + x;
            rgb[o] = R;
            rgb[n + o] = G;
            rgb[2 * n + o] = B;
        }
    }
    return 0;
}
