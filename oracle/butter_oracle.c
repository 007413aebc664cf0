/*
 * CPU oracle for Butteraugli, FP64. TEST INFRASTRUCTURE ONLY (see
 * butter_oracle.h): a restatement of the reference's src/butter/ stages in
 * double arithmetic, each function citing the reference lines it follows.
 */
#include "butter_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* blur: the reference's five Gaussians (src/butter/gaussianblur.hpp:27-42):
   window = max(floor(3 sigma), 8), taps exp(-k^2 / (2 sigma^2)) for
   k = -window..window, every output divided by the sum of its in-image taps
   (horizontal, then vertical) (:55-85, and the tiled form :87-204 for
   window 8, which computes the same renormalised sums). */
static const double kSigmas[5] = {1.2, 1.56416327805, 2.7, 3.22489901262, 7.15593339443};

static int blur_window(int i) {
    const int w3 = (int)(kSigmas[i] * 3.0);
    return w3 > 8 ? w3 : 8;
}

void bo_blur(const double* src, int w, int h, int si, double* dst) {
    const int R = blur_window(si);
    const double s = kSigmas[si];
    double* g = (double*)malloc(sizeof(double) * (2 * R + 1));
    double* tmp = (double*)malloc(sizeof(double) * (size_t)w * h);
    for (int k = -R; k <= R; k++) g[k + R] = exp(-(double)k * k / (2.0 * s * s));
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double acc = 0.0, wsum = 0.0;
            for (int k = -R; k <= R; k++) {
                const int xx = x + k;
                if (xx < 0 || xx >= w) continue;
                acc += g[k + R] * src[(int64_t)y * w + xx];
                wsum += g[k + R];
            }
            tmp[(int64_t)y * w + x] = acc / wsum;
        }
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double acc = 0.0, wsum = 0.0;
            for (int k = -R; k <= R; k++) {
                const int yy = y + k;
                if (yy < 0 || yy >= h) continue;
                acc += g[k + R] * tmp[(int64_t)yy * w + x];
                wsum += g[k + R];
            }
            dst[(int64_t)y * w + x] = acc / wsum;
        }
    free(g);
    free(tmp);
}

/* ---------------------------------------------------------------------- */
/* opsin dynamics (src/butter/colors.hpp:3-5 gamma, :26-48 absorbance,
   :50-79 per-pixel step, :81-90 the sigma-1.2 blur of the input) */
static const double kOpsin[3][4] = {
    {0.29956550340058319, 0.63373087833825936, 0.077705617820981968, 1.7557483643287353},
    {0.22158691104574774, 0.69391388044116142, 0.0987313588422, 1.7557483643287353},
    {0.02, 0.02, 0.20480129041026129, 12.226454707163354},
};

static double bgamma(double v) { return 19.245013259874995 * log(v + 9.9710635769299145) - 23.16046239805755; }

static void opsin_dynamics(const double* lin, int w, int h, double intensity, double* xyb) {
    const int64_t n = (int64_t)w * h;
    double* bl = (double*)malloc(sizeof(double) * 3 * n);
    for (int c = 0; c < 3; c++) bo_blur(lin + c * n, w, h, 0, bl + c * n);
    for (int64_t i = 0; i < n; i++) {
        double s[3], b[3], sens[3], o[3];
        for (int c = 0; c < 3; c++) {
            s[c] = lin[c * n + i] * intensity;
            b[c] = bl[c * n + i] * intensity;
        }
        for (int r = 0; r < 3; r++) {
            double v = kOpsin[r][0] * b[0] + kOpsin[r][1] * b[1] + kOpsin[r][2] * b[2] + kOpsin[r][3];
            v = v > kOpsin[r][3] ? v : kOpsin[r][3]; /* clamp to the bias */
            v = v > 1e-4 ? v : 1e-4;
            sens[r] = bgamma(v) / v;
            sens[r] = sens[r] > 1e-4 ? sens[r] : 1e-4;
        }
        for (int r = 0; r < 3; r++) {
            o[r] = (kOpsin[r][0] * s[0] + kOpsin[r][1] * s[1] + kOpsin[r][2] * s[2] + kOpsin[r][3]) * sens[r];
            o[r] = o[r] > kOpsin[r][3] ? o[r] : kOpsin[r][3];
        }
        xyb[i] = o[0] - o[1];
        xyb[n + i] = o[0] + o[1];
        xyb[2 * n + i] = o[2];
    }
    free(bl);
}

/* ---------------------------------------------------------------------- */
/* frequency separation (src/butter/separatefrequencies.hpp:3-55 point
   operations, :120-175 the stages) */
static double remove_range(double x, double wv) { return x > wv ? x - wv : (x < -wv ? x + wv : 0.0); }
static double amplify_range(double x, double wv) { return x > wv ? x + wv : (x < -wv ? x - wv : 2.0 * x); }
static double max_clamp(double v, double maxval) {
    const double mul = 0.688059627878;
    if (v >= maxval) return (v - maxval) * mul + maxval;
    if (v < -maxval) return (v + maxval) * mul - maxval;
    return v;
}

/* lf[3], mf[3], hf[2], uhf[2] planes of w*h */
static void separate(const double* xyb, int w, int h, double* lf, double* mf, double* hf, double* uhf) {
    const int64_t n = (int64_t)w * h;
    double* t = (double*)malloc(sizeof(double) * n);
    for (int c = 0; c < 3; c++) {
        bo_blur(xyb + c * n, w, h, 4, lf + c * n);
        for (int64_t i = 0; i < n; i++) t[i] = xyb[c * n + i] - lf[c * n + i];
        bo_blur(t, w, h, 3, mf + c * n);
        if (c == 2) break;
        /* high band before its own blur: (xyb - lf) - mf, from the mf before
           its range operation (:57-93) */
        for (int64_t i = 0; i < n; i++) {
            uhf[c * n + i] = t[i] - mf[c * n + i];
            mf[c * n + i] = c == 0 ? remove_range(mf[c * n + i], 0.29) : amplify_range(mf[c * n + i], 0.1);
        }
    }
    /* X's high band suppressed by Y's (:26-30, :160) */
    for (int64_t i = 0; i < n; i++) {
        const double s = 0.653020556257, yw = 46.0, yv = uhf[n + i];
        uhf[i] *= s + (yw * (1.0 - s)) / (yw + yv * yv);
    }
    for (int c = 0; c < 2; c++) {
        bo_blur(uhf + c * n, w, h, 1, hf + c * n);
        for (int64_t i = 0; i < n; i++) {
            double hv = hf[c * n + i], uv = uhf[c * n + i];
            if (c == 0) {
                uv -= hv;
                hv = remove_range(hv, 1.5);
                uv = remove_range(uv, 0.04);
            } else {
                hv = max_clamp(hv, 28.4691806922);
                uv -= hv;
                uv = max_clamp(uv, 5.19175294647);
                uv *= 2.69313763794;
                hv *= 2.155;
                hv = amplify_range(hv, 0.132);
            }
            hf[c * n + i] = hv;
            uhf[c * n + i] = uv;
        }
    }
    /* low band to values (:3-13) */
    for (int64_t i = 0; i < n; i++) {
        double x = lf[i], y = lf[n + i], b = lf[2 * n + i];
        b = (b + -0.362267051518 * y) * 49.87984651440;
        lf[i] = x * 33.832837186260;
        lf[n + i] = y * 14.458268100570;
        lf[2 * n + i] = b;
    }
    free(t);
}

/* ---------------------------------------------------------------------- */
/* Malta line filters (src/butter/maltaDiff.hpp): 16 lines around the pixel,
   each the square of the sum of the scaled differences on it. Offsets
   (dx, dy) of the 5-tap (LF) and the 7- to 9-tap (HF) lines. */
typedef struct { int n; int off[9][2]; } MaltaLine;

static const MaltaLine kMaltaLF[16] = {
    {5, {{-4, 0}, {-2, 0}, {0, 0}, {2, 0}, {4, 0}}},
    {5, {{0, -4}, {0, -2}, {0, 0}, {0, 2}, {0, 4}}},
    {5, {{-3, -3}, {-2, -2}, {0, 0}, {2, 2}, {3, 3}}},
    {5, {{3, -3}, {2, -2}, {0, 0}, {-2, 2}, {-3, 3}}},
    {5, {{1, -4}, {1, -2}, {0, 0}, {-1, 2}, {-1, 4}}},
    {5, {{-1, -4}, {-1, -2}, {0, 0}, {1, 2}, {1, 4}}},
    {5, {{-4, -1}, {-2, -1}, {0, 0}, {2, 1}, {4, 1}}},
    {5, {{-4, 1}, {-2, 1}, {0, 0}, {2, -1}, {4, -1}}},
    {5, {{-2, -3}, {-1, -2}, {0, 0}, {1, 2}, {2, 3}}},
    {5, {{2, -3}, {1, -2}, {0, 0}, {-1, 2}, {-2, 3}}},
    {5, {{-3, -2}, {-2, -1}, {0, 0}, {2, 1}, {3, 2}}},
    {5, {{3, -2}, {2, -1}, {0, 0}, {-2, 1}, {-3, 2}}},
    {5, {{-4, 2}, {-2, 1}, {0, 0}, {2, -1}, {4, -2}}},
    {5, {{-4, -2}, {-2, -1}, {0, 0}, {2, 1}, {4, 2}}},
    {5, {{-2, -4}, {-1, -2}, {0, 0}, {1, 2}, {2, 4}}},
    {5, {{2, -4}, {1, -2}, {0, 0}, {-1, 2}, {-2, 4}}},
};

static const MaltaLine kMaltaHF[16] = {
    {9, {{-4, 0}, {-3, 0}, {-2, 0}, {-1, 0}, {0, 0}, {1, 0}, {2, 0}, {3, 0}, {4, 0}}},
    {9, {{0, -4}, {0, -3}, {0, -2}, {0, -1}, {0, 0}, {0, 1}, {0, 2}, {0, 3}, {0, 4}}},
    {7, {{-3, -3}, {-2, -2}, {-1, -1}, {0, 0}, {1, 1}, {2, 2}, {3, 3}}},
    {7, {{3, -3}, {2, -2}, {1, -1}, {0, 0}, {-1, 1}, {-2, 2}, {-3, 3}}},
    {9, {{1, -4}, {1, -3}, {1, -2}, {0, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}, {-1, 4}}},
    {9, {{-1, -4}, {-1, -3}, {-1, -2}, {0, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}, {1, 4}}},
    {9, {{-4, -1}, {-3, -1}, {-2, -1}, {-1, 0}, {0, 0}, {1, 0}, {2, 1}, {3, 1}, {4, 1}}},
    {9, {{-4, 1}, {-3, 1}, {-2, 1}, {-1, 0}, {0, 0}, {1, 0}, {2, -1}, {3, -1}, {4, -1}}},
    {7, {{-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {1, 1}, {1, 2}, {2, 3}}},
    {7, {{2, -3}, {1, -2}, {1, -1}, {0, 0}, {-1, 1}, {-1, 2}, {-2, 3}}},
    {7, {{-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 1}, {2, 1}, {3, 2}}},
    {7, {{3, -2}, {2, -1}, {1, -1}, {0, 0}, {-1, 1}, {-2, 1}, {-3, 2}}},
    {8, {{-4, 2}, {-3, 2}, {-2, 1}, {-1, 1}, {0, 0}, {1, 0}, {2, -1}, {3, -1}}},
    {8, {{-4, -2}, {-3, -2}, {-2, -1}, {-1, -1}, {0, 0}, {1, 0}, {2, 1}, {3, 1}}},
    {8, {{-2, -4}, {-2, -3}, {-1, -2}, {-1, -1}, {0, 0}, {0, 1}, {1, 2}, {1, 3}}},
    {8, {{2, -4}, {2, -3}, {1, -2}, {1, -1}, {0, 0}, {0, 1}, {-1, 2}, {-1, 3}}},
};

/* the scaled difference the lines sum (MaltaDiffMap*_Kernel): a symmetric
   term and a half-open term that penalises l1 outside [0.55, 1.05] |l0| on
   l0's side; 0 outside the image */
void bo_malta(const double* l0, const double* l1, int w, int h, int hf, double w_0gt1, double w_0lt1,
              double norm1, double* acc) {
    const int64_t n = (int64_t)w * h;
    const double len = 3.75, mulli = hf ? 0.39905817637 : 0.611612573796;
    const double len2 = len * 2.0 + 1.0;
    const double n2gt = mulli * sqrt(0.5 * w_0gt1) / len2 * norm1;
    const double n2lt = mulli * sqrt(0.33 * w_0lt1) / len2 * norm1;
    double* d = (double*)malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; i++) {
        const double a = l0[i], b = l1[i];
        const double absval = 0.5 * fabs(a) + 0.5 * fabs(b);
        const double diff = a - b;
        double v = n2gt / (norm1 + absval) * diff;
        const double small = 0.55 * fabs(a), big = 1.05 * fabs(a);
        double impact = 0.0;
        if (a < 0.0) {
            if (b > -small) impact = b + small;
            else if (b < -big) impact = -b - big;
        } else {
            if (b < small) impact = -b + small;
            else if (b > big) impact = b - big;
        }
        impact *= n2lt / (norm1 + absval);
        v += diff < 0.0 ? -impact : impact;
        d[i] = v;
    }
    const MaltaLine* lines = hf ? kMaltaHF : kMaltaLF;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double r = 0.0;
            for (int l = 0; l < 16; l++) {
                double s = 0.0;
                for (int k = 0; k < lines[l].n; k++) {
                    const int xx = x + lines[l].off[k][0], yy = y + lines[l].off[k][1];
                    if (xx >= 0 && xx < w && yy >= 0 && yy < h) s += d[(int64_t)yy * w + xx];
                }
                r += s * s;
            }
            acc[(int64_t)y * w + x] += r;
        }
    free(d);
}

/* ---------------------------------------------------------------------- */
/* L2 terms (src/butter/simplerdiff.hpp:38-111); the asymmetric one's
   weights are multiplied by 0.8 once more in its launcher (:104-107) */
static void l2diff(const double* a, const double* b, int64_t n, double wv, double* acc) {
    if (wv == 0.0) return;
    for (int64_t i = 0; i < n; i++) {
        const double d = a[i] - b[i];
        acc[i] += wv * d * d;
    }
}

static void l2asym(const double* a, const double* b, int64_t n, double wgt, double wlt, double* acc) {
    wgt *= 0.8;
    wlt *= 0.8;
    for (int64_t i = 0; i < n; i++) {
        const double d = a[i] - b[i];
        acc[i] += wgt * d * d;
        const double f = fabs(a[i]), small = 0.4 * f, big = f;
        double v = 0.0;
        if (a[i] < 0.0) {
            if (b[i] > -small) v = b[i] + small;
            else if (b[i] < -big) v = -b[i] - big;
        } else {
            if (b[i] < small) v = small - b[i];
            else if (b[i] > big) v = b[i] - big;
        }
        acc[i] += wlt * v * v;
    }
}

/* ---------------------------------------------------------------------- */
/* psycho mask (src/butter/maskPsycho.hpp): mask amplitude of the high
   bands, compressed, blurred at sigma 2.7, fuzzy erosion of the reference's
   (the mask) and an L2 term between the two blurred masks (into ac[1]) */
static void mask_xyb(const double* hf, const double* uhf, int64_t n, double* out) {
    for (int64_t i = 0; i < n; i++) {
        const double xd = (uhf[i] + hf[i]) * 2.5;
        const double yd = uhf[n + i] * 0.4 + hf[n + i] * 0.4;
        const double m = sqrt(xd * xd + yd * yd);
        const double mul = 6.19424080439, bias = mul * 12.61050594197;
        out[i] = sqrt(mul * fabs(m) + bias) - sqrt(bias);
    }
}

static void store_min3(double v, double* m0, double* m1, double* m2) {
    if (v < *m2) {
        if (v < *m0) {
            *m2 = *m1;
            *m1 = *m0;
            *m0 = v;
        } else if (v < *m1) {
            *m2 = *m1;
            *m1 = v;
        } else {
            *m2 = v;
        }
    }
}

static void fuzzy_erosion(const double* s, int w, int h, double* out) {
    const int K = 3;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            double m0 = s[(int64_t)y * w + x], m1 = 2.0 * m0, m2 = m1;
            if (x >= K) {
                store_min3(s[(int64_t)y * w + x - K], &m0, &m1, &m2);
                if (y >= K) store_min3(s[(int64_t)(y - K) * w + x - K], &m0, &m1, &m2);
                if (y < h - K) store_min3(s[(int64_t)(y + K) * w + x - K], &m0, &m1, &m2);
            }
            if (x < w - K) {
                store_min3(s[(int64_t)y * w + x + K], &m0, &m1, &m2);
                if (y >= K) store_min3(s[(int64_t)(y - K) * w + x + K], &m0, &m1, &m2);
                if (y < h - K) store_min3(s[(int64_t)(y + K) * w + x + K], &m0, &m1, &m2);
            }
            if (y >= K) store_min3(s[(int64_t)(y - K) * w + x], &m0, &m1, &m2);
            if (y < h - K) store_min3(s[(int64_t)(y + K) * w + x], &m0, &m1, &m2);
            out[(int64_t)y * w + x] = 0.45 * m0 + 0.3 * m1 + 0.25 * m2;
        }
}

/* mask weights of the AC and DC sums (src/butter/combineMasks.hpp) */
static double mask_y(double d) {
    const double c = 2.5485944793 / (0.451936922203 * d + 0.829591754942);
    const double r = (1.0 + c) * (1.0 / (0.79079917404 * 17.83));
    return r * r;
}
static double mask_dc_y(double d) {
    const double c = 0.505054525019 / (3.87449418804 * d + 0.20025578522);
    const double r = (1.0 + c) * (1.0 / (0.79079917404 * 17.83));
    return r * r;
}

/* ---------------------------------------------------------------------- */
/* one scale (src/butter/main.hpp:22-110) */
int bo_diffmap_scale(const double* lin1, const double* lin2, int w, int h, double intensity, double* diffmap) {
    const int64_t n = (int64_t)w * h;
    /* planes: xyb 2x3, lf/mf 2x3 each, hf/uhf 2x2 each, ac and dc 3 each, and
       the two masks, their blurs and the eroded mask */
    enum { kPlanes = 6 + 12 + 8 + 6 + 5 };
    double* mem = (double*)calloc((size_t)n * kPlanes, sizeof(double));
    if (!mem) return -1;
    double *x1 = mem, *x2 = x1 + 3 * n;
    double *lf1 = x2 + 3 * n, *mf1 = lf1 + 3 * n, *hf1 = mf1 + 3 * n, *uhf1 = hf1 + 2 * n;
    double *lf2 = uhf1 + 2 * n, *mf2 = lf2 + 3 * n, *hf2 = mf2 + 3 * n, *uhf2 = hf2 + 2 * n;
    double *ac = uhf2 + 2 * n, *dc = ac + 3 * n;
    double *m0 = dc + 3 * n, *m1 = m0 + n, *b0 = m1 + n, *b1 = b0 + n, *mask = b1 + n;
    opsin_dynamics(lin1, w, h, intensity, x1);
    opsin_dynamics(lin2, w, h, intensity, x2);
    separate(x1, w, h, lf1, mf1, hf1, uhf1);
    separate(x2, w, h, lf2, mf2, hf2, uhf2);

    const double asym = 0.8, sa = sqrt(asym);
    bo_malta(uhf1 + n, uhf2 + n, w, h, 1, 1.10039032555 * asym, 1.10039032555 / asym, 71.7800275169, ac + n);
    bo_malta(uhf1, uhf2, w, h, 1, 173.5 * asym, 173.5 / asym, 5.0, ac);
    bo_malta(hf1 + n, hf2 + n, w, h, 0, 18.7237414387 * sa, 18.7237414387 / sa, 4498534.45232, ac + n);
    bo_malta(hf1, hf2, w, h, 0, 6923.99476109 * sa, 6923.99476109 / sa, 8051.15833247, ac);
    bo_malta(mf1 + n, mf2 + n, w, h, 0, 37.0819870399, 37.0819870399, 130262059.556, ac + n);
    bo_malta(mf1, mf2, w, h, 0, 8246.75321353, 8246.75321353, 1009002.70582, ac);

    static const double wmul[9] = {400.0,         1.50815703118,  0.0,
                                   2150.0,        10.6195433239,  16.2176043152,
                                   29.2353797994, 0.844626970982, 0.703646627719};
    for (int c = 0; c < 3; c++) {
        if (c < 2) l2asym(hf1 + c * n, hf2 + c * n, n, wmul[c] * asym, wmul[c] / asym, ac + c * n);
        l2diff(mf1 + c * n, mf2 + c * n, n, wmul[3 + c], ac + c * n);
        l2diff(lf1 + c * n, lf2 + c * n, n, wmul[6 + c], dc + c * n);
    }

    mask_xyb(hf1, uhf1, n, m0);
    mask_xyb(hf2, uhf2, n, m1);
    bo_blur(m0, w, h, 2, b0);
    bo_blur(m1, w, h, 2, b1);
    fuzzy_erosion(b0, w, h, mask);
    l2diff(b0, b1, n, 10.0, ac + n);

    for (int64_t i = 0; i < n; i++) {
        const double v = mask[i];
        diffmap[i] = sqrt(mask_y(v) * (ac[i] + ac[n + i] + ac[2 * n + i]) +
                          mask_dc_y(v) * (dc[i] + dc[n + i] + dc[2 * n + i]));
    }
    free(mem);
    return 0;
}

/* ---------------------------------------------------------------------- */
/* full metric (src/butter/main.hpp:112-195): linearise, map at full and
   half scale, blend, norms (src/butter/diffnorms.hpp:78-120) */
static double linearize(double v) { return v < 0.0 ? -pow(-v, 2.4) : pow(v, 2.4); }

static double sample(int fmt, const uint8_t* plane, int64_t stride, int x, int y) {
    const uint8_t* row = plane + (int64_t)y * stride;
    if (fmt == BO_U16) {
        uint16_t u;
        memcpy(&u, row + 2 * (int64_t)x, 2);
        return u / 65535.0;
    }
    float f;
    memcpy(&f, row + 4 * (int64_t)x, 4);
    return f;
}

int bo_butteraugli(int fmt, const uint8_t* const p1[3], int64_t stride1, const uint8_t* const p2[3],
                   int64_t stride2, int w, int h, int qnorm, double intensity, double norms3[3], double* diffmap) {
    if (w < 1 || h < 1 || qnorm < 1 || (fmt != BO_FLOAT && fmt != BO_U16)) return -1;
    const int64_t n = (int64_t)w * h;
    const int w2 = (w - 1) / 2 + 1, h2 = (h - 1) / 2 + 1;
    const int64_t n2 = (int64_t)w2 * h2;
    double* a = (double*)malloc(sizeof(double) * 3 * n);
    double* b = (double*)malloc(sizeof(double) * 3 * n);
    double* a2 = (double*)malloc(sizeof(double) * 3 * n2);
    double* b2 = (double*)malloc(sizeof(double) * 3 * n2);
    double* map = (double*)malloc(sizeof(double) * n);
    double* map2 = (double*)malloc(sizeof(double) * n2);
    if (!a || !b || !a2 || !b2 || !map || !map2) return -1;
    for (int c = 0; c < 3; c++)
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                a[c * n + (int64_t)y * w + x] = linearize(sample(fmt, p1[c], stride1, x, y));
                b[c * n + (int64_t)y * w + x] = linearize(sample(fmt, p2[c], stride2, x, y));
            }
    /* 2x2 mean with the second row/column clamped (downupsample.hpp:5-20) */
    for (int c = 0; c < 3; c++)
        for (int y = 0; y < h2; y++)
            for (int x = 0; x < w2; x++) {
                const int y0 = 2 * y < h - 1 ? 2 * y : h - 1, y1 = 2 * y + 1 < h - 1 ? 2 * y + 1 : h - 1;
                const int x0 = 2 * x < w - 1 ? 2 * x : w - 1, x1 = 2 * x + 1 < w - 1 ? 2 * x + 1 : w - 1;
                const double* pa = a + c * n;
                const double* pb = b + c * n;
                a2[c * n2 + (int64_t)y * w2 + x] =
                    0.25 * (pa[(int64_t)y0 * w + x0] + pa[(int64_t)y1 * w + x0] + pa[(int64_t)y0 * w + x1] +
                            pa[(int64_t)y1 * w + x1]);
                b2[c * n2 + (int64_t)y * w2 + x] =
                    0.25 * (pb[(int64_t)y0 * w + x0] + pb[(int64_t)y1 * w + x0] + pb[(int64_t)y0 * w + x1] +
                            pb[(int64_t)y1 * w + x1]);
            }
    int rc = bo_diffmap_scale(a, b, w, h, intensity, map);
    if (rc == 0) rc = bo_diffmap_scale(a2, b2, w2, h2, intensity, map2);
    if (rc == 0) {
        /* blend: full * (1 - 0.3 w) + w * half, w = 0.5 (downupsample.hpp:37-48) */
        double sq = 0.0, s3 = 0.0, mx = 0.0;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) {
                const double v = map[(int64_t)y * w + x] * (1.0 - 0.3 * 0.5) + 0.5 * map2[(int64_t)(y / 2) * w2 + x / 2];
                if (diffmap) diffmap[(int64_t)y * w + x] = v;
                sq += pow(v, (double)qnorm);
                s3 += fabs(v) * v * v;
                mx = fabs(v) > mx ? fabs(v) : mx;
            }
        norms3[0] = pow(sq / (double)n, 1.0 / qnorm);
        norms3[1] = pow(s3 / (double)n, 1.0 / 3.0);
        norms3[2] = mx;
    }
    free(a); free(b); free(a2); free(b2); free(map); free(map2);
    return rc;
}
