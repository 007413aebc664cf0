/*
 * Body of the oracle pipeline, included twice by ssimu2_oracle.c:
 *   REAL=double, FN(x)=x##_f64   (oracle)
 *   REAL=float,  FN(x)=x##_f32   (timed CPU baseline)
 * Sums over pixels are always accumulated in double.
 *
 * TEST INFRASTRUCTURE ONLY (see ssimu2_oracle.h).
 */

/* 2x2 mean with clamp to the last row/column, src/ssimu2/downsample.hpp:3-19.
 * The four samples are added in the reference's order, then scaled by 1/4. */
static void FN(downsample)(const REAL* src, int w, int h, REAL* dst) {
    const int nw = (w - 1) / 2 + 1, nh = (h - 1) / 2 + 1;
    for (int y = 0; y < nh; y++) {
        const int y0 = 2 * y < h - 1 ? 2 * y : h - 1;
        const int y1 = 2 * y + 1 < h - 1 ? 2 * y + 1 : h - 1;
        for (int x = 0; x < nw; x++) {
            const int x0 = 2 * x < w - 1 ? 2 * x : w - 1;
            const int x1 = 2 * x + 1 < w - 1 ? 2 * x + 1 : w - 1;
            REAL v = src[(int64_t)y0 * w + x0];
            v += src[(int64_t)y1 * w + x0];
            v += src[(int64_t)y0 * w + x1];
            v += src[(int64_t)y1 * w + x1];
            dst[(int64_t)y * nw + x] = v * (REAL)0.25;
        }
    }
}

/* Separable 17-tap FIR, each pass renormalised by the in-bounds taps
 * (src/ssimu2/gaussianblur.hpp:55-92; out-of-image samples read as 0,
 * gaussianblur.hpp:24-50). tmp must hold w*h values. */
static void FN(blur)(const REAL* src, int w, int h, REAL* dst, REAL* tmp) {
    REAL g[17];
    double gd[17];
    so_gaussian_taps(gd);
    for (int i = 0; i < 17; i++) g[i] = (REAL)gd[i];
    #pragma omp parallel for schedule(static) if (FN_PAR)
    for (int y = 0; y < h; y++) {
        const REAL* row = src + (int64_t)y * w;
        REAL* out = tmp + (int64_t)y * w;
        for (int x = 0; x < w; x++) {
            const int k0 = x - 8 < 0 ? 8 - x : 0;
            const int k1 = x + 8 >= w ? w - 1 - x + 8 : 16;
            REAL s = 0, t = 0;
            for (int k = k0; k <= k1; k++) {
                s += g[k] * row[x + k - 8];
                t += g[k];
            }
            out[x] = s / t;
        }
    }
    #pragma omp parallel for schedule(static) if (FN_PAR)
    for (int y = 0; y < h; y++) {
        const int k0 = y - 8 < 0 ? 8 - y : 0;
        const int k1 = y + 8 >= h ? h - 1 - y + 8 : 16;
        REAL t = 0;
        for (int k = k0; k <= k1; k++) t += g[k];
        REAL* out = dst + (int64_t)y * w;
        for (int x = 0; x < w; x++) {
            REAL s = 0;
            for (int k = k0; k <= k1; k++) s += g[k] * tmp[(int64_t)(y + k - 8) * w + x];
            out[x] = s / t;
        }
    }
}

/* Positive XYB of one linear-RGB pixel, src/ssimu2/makeXYB.hpp:4-50. */
static inline void FN(xyb_px)(REAL r, REAL g, REAL b, REAL* ox, REAL* oy, REAL* ob) {
    const REAL bias = (REAL)0.0037930734;
    const REAL abs_bias = (REAL)-0.1559542025327239;
    REAL l = (REAL)0.30 * r + (REAL)0.622 * g + (REAL)0.078 * b + bias;
    REAL m = (REAL)0.23 * r + (REAL)0.692 * g + (REAL)0.078 * b + bias;
    REAL s = (REAL)0.24342269 * r + (REAL)0.20476745 * g + (REAL)0.55180986 * b + bias;
    l = CBRT(l > 0 ? l : 0) + abs_bias;
    m = CBRT(m > 0 ? m : 0) + abs_bias;
    s = CBRT(s > 0 ? s : 0) + abs_bias;
    const REAL X = (REAL)0.5 * (l - m);
    const REAL Y = (REAL)0.5 * (l + m);
    *ob = (s - Y) + (REAL)0.55;
    *ox = X * (REAL)14.0 + (REAL)0.42;
    *oy = Y + (REAL)0.01;
}

/* SSIM, artifact and detail-loss sums of one scale,
 * src/ssimu2/score.hpp:113-196 (maps) and :244-267 (norms).
 * xyb*: 3 planes of w*h. out18[ch*6 + map*2 + norm]. */
static void FN(scale_features)(const REAL* xyb1, const REAL* xyb2, int w, int h, double out18[18]) {
    const int64_t n = (int64_t)w * h;
    REAL* buf = (REAL*)malloc(sizeof(REAL) * n * 7);
    REAL *m1 = buf, *m2 = buf + n, *s11 = buf + 2 * n, *s22 = buf + 3 * n, *s12 = buf + 4 * n;
    REAL *prod = buf + 5 * n, *tmp = buf + 6 * n;
    for (int c = 0; c < 3; c++) {
        const REAL* i1 = xyb1 + c * n;
        const REAL* i2 = xyb2 + c * n;
        FN(blur)(i1, w, h, m1, tmp);
        FN(blur)(i2, w, h, m2, tmp);
        for (int64_t i = 0; i < n; i++) prod[i] = i1[i] * i1[i];
        FN(blur)(prod, w, h, s11, tmp);
        for (int64_t i = 0; i < n; i++) prod[i] = i2[i] * i2[i];
        FN(blur)(prod, w, h, s22, tmp);
        for (int64_t i = 0; i < n; i++) prod[i] = i1[i] * i2[i];
        FN(blur)(prod, w, h, s12, tmp);
        double a[6] = {0, 0, 0, 0, 0, 0};
        for (int64_t i = 0; i < n; i++) {
            const REAL mu1 = m1[i], mu2 = m2[i];
            const REAL md = mu1 - mu2;
            const REAL num_m = (REAL)1.0 - md * md;
            const REAL num_s = (REAL)2.0 * (s12[i] - mu1 * mu2) + (REAL)0.0009;
            const REAL den_s = (s11[i] - mu1 * mu1) + (s22[i] - mu2 * mu2) + (REAL)0.0009;
            REAL d = (REAL)1.0 - (num_m * num_s) / den_s;
            if (!(d > 0)) d = 0;
            const REAL v = (FABS(i2[i] - mu2) + (REAL)1.0) / (FABS(i1[i] - mu1) + (REAL)1.0) - (REAL)1.0;
            const REAL art = v > 0 ? v : 0;
            const REAL det = -v > 0 ? -v : 0;
            const double dd = (double)d, da = (double)art, de = (double)det;
            a[0] += dd; a[1] += (dd * dd) * (dd * dd);
            a[2] += da; a[3] += (da * da) * (da * da);
            a[4] += de; a[5] += (de * de) * (de * de);
        }
        for (int k = 0; k < 3; k++) {
            out18[c * 6 + 2 * k] = a[2 * k] / (double)n;
            out18[c * 6 + 2 * k + 1] = sqrt(sqrt(a[2 * k + 1] / (double)n));
        }
    }
    free(buf);
}

/* Pyramid, XYB and features from linear RGB at scale 0.
 * lin1/lin2: 3 planes of w*h (planar). feat108 in the reference's order
 * f[c*36 + s*6 + n*3 + m] (src/ssimu2/main.hpp:82-95). */
static int FN(score_from_linear)(const REAL* lin1, const REAL* lin2, int w, int h,
                                 double* score, double* feat108) {
    int dims[12];
    so_scale_dims(w, h, dims);
    int64_t off[7];
    off[0] = 0;
    for (int s = 0; s < 6; s++) off[s + 1] = off[s] + (int64_t)dims[2 * s] * dims[2 * s + 1];
    const int64_t T = off[6];
    REAL* L1 = (REAL*)malloc(sizeof(REAL) * T * 3);
    REAL* L2 = (REAL*)malloc(sizeof(REAL) * T * 3);
    REAL* X1 = (REAL*)malloc(sizeof(REAL) * T * 3);
    REAL* X2 = (REAL*)malloc(sizeof(REAL) * T * 3);
    if (!L1 || !L2 || !X1 || !X2) { free(L1); free(L2); free(X1); free(X2); return -1; }
    const int64_t n0 = (int64_t)w * h;
    for (int c = 0; c < 3; c++) {
        memcpy(L1 + c * T, lin1 + c * n0, sizeof(REAL) * n0);
        memcpy(L2 + c * T, lin2 + c * n0, sizeof(REAL) * n0);
    }
    /* pyramid in linear RGB, src/ssimu2/main.hpp:48-66 */
    for (int s = 1; s < 6; s++)
        for (int c = 0; c < 3; c++) {
            FN(downsample)(L1 + c * T + off[s - 1], dims[2 * s - 2], dims[2 * s - 1], L1 + c * T + off[s]);
            FN(downsample)(L2 + c * T + off[s - 1], dims[2 * s - 2], dims[2 * s - 1], L2 + c * T + off[s]);
        }
    /* positive XYB on every scale, src/ssimu2/main.hpp:68-70 */
    for (int64_t i = 0; i < T; i++) {
        FN(xyb_px)(L1[i], L1[T + i], L1[2 * T + i], &X1[i], &X1[T + i], &X1[2 * T + i]);
        FN(xyb_px)(L2[i], L2[T + i], L2[2 * T + i], &X2[i], &X2[T + i], &X2[2 * T + i]);
    }
    double f[108];
    for (int s = 0; s < 6; s++) {
        const int sw = dims[2 * s], sh = dims[2 * s + 1];
        const int64_t ns = (int64_t)sw * sh;
        REAL* a = (REAL*)malloc(sizeof(REAL) * ns * 6);
        for (int c = 0; c < 3; c++) {
            memcpy(a + c * ns, X1 + c * T + off[s], sizeof(REAL) * ns);
            memcpy(a + (3 + c) * ns, X2 + c * T + off[s], sizeof(REAL) * ns);
        }
        double r[18];
        FN(scale_features)(a, a + 3 * ns, sw, sh, r);
        free(a);
        for (int c = 0; c < 3; c++)
            for (int m = 0; m < 3; m++)
                for (int nrm = 0; nrm < 2; nrm++)
                    f[c * 36 + s * 6 + nrm * 3 + m] = r[c * 6 + m * 2 + nrm];
    }
    free(L1); free(L2); free(X1); free(X2);
    if (feat108) memcpy(feat108, f, sizeof(f));
    if (score) *score = so_final_score(f);
    return 0;
}

/* Planar RGB (gamma-encoded) -> linear RGB, src/util/float3operations.hpp:197-218
 * then src/ssimu2/makeXYB.hpp:52-64. */
static void FN(unpack_linear)(int fmt, const uint8_t* const p[3], int64_t stride, int w, int h, REAL* out) {
    const int64_t n = (int64_t)w * h;
    for (int c = 0; c < 3; c++)
        for (int y = 0; y < h; y++) {
            const uint8_t* row = p[c] + (int64_t)y * stride;
            for (int x = 0; x < w; x++) {
                double v;
                switch (fmt) {
                case SO_RGB_U8: v = (double)(float)((float)row[x] / 255.0f); break;
                case SO_RGB_U16: { uint16_t u; memcpy(&u, row + 2 * x, 2); v = (double)(float)((float)u / 65535.0f); } break;
                case SO_RGB_F16: { uint16_t u; memcpy(&u, row + 2 * x, 2); v = so_half_to_double(u); } break;
                default: { float f; memcpy(&f, row + 4 * x, 4); v = f; } break;
                }
                out[c * n + (int64_t)y * w + x] = (REAL)so_linearize(v);
            }
        }
}

int FN(so_score_rgb)(int fmt, const uint8_t* const p1[3], int64_t stride1,
                     const uint8_t* const p2[3], int64_t stride2, int w, int h,
                     double* score, double* feat108) {
    if (w <= 0 || h <= 0 || fmt < 0 || fmt > 3) return -1;
    const int64_t n = (int64_t)w * h;
    REAL* a = (REAL*)malloc(sizeof(REAL) * n * 3);
    REAL* b = (REAL*)malloc(sizeof(REAL) * n * 3);
    if (!a || !b) { free(a); free(b); return -1; }
    FN(unpack_linear)(fmt, p1, stride1, w, h, a);
    FN(unpack_linear)(fmt, p2, stride2, w, h, b);
    const int rc = FN(score_from_linear)(a, b, w, h, score, feat108);
    free(a); free(b);
    return rc;
}

int FN(so_score_yuv)(const uint8_t* const p1[3], const int64_t s1[3],
                     const uint8_t* const p2[3], const int64_t s2[3], int w, int h,
                     int bits, int matrix, int sw, int sh, int loc, double* score, double* feat108) {
    if (w <= 0 || h <= 0) return -1;
    const int64_t n = (int64_t)w * h;
    double* rgb = (double*)malloc(sizeof(double) * n * 3);
    REAL* a = (REAL*)malloc(sizeof(REAL) * n * 3);
    REAL* b = (REAL*)malloc(sizeof(REAL) * n * 3);
    int rc = -1;
    if (!rgb || !a || !b) goto done;
    if (so_yuv_to_rgb_f64(p1, s1, w, h, bits, matrix, sw, sh, loc, rgb)) goto done;
    for (int64_t i = 0; i < 3 * n; i++) a[i] = (REAL)so_linearize(rgb[i]);
    if (so_yuv_to_rgb_f64(p2, s2, w, h, bits, matrix, sw, sh, loc, rgb)) goto done;
    for (int64_t i = 0; i < 3 * n; i++) b[i] = (REAL)so_linearize(rgb[i]);
    rc = FN(score_from_linear)(a, b, w, h, score, feat108);
done:
    free(rgb); free(a); free(b);
    return rc;
}

int FN(so_score_yuv420)(const uint8_t* const p1[3], const int64_t s1[3],
                        const uint8_t* const p2[3], const int64_t s2[3], int w, int h,
                        int bits, int matrix, double* score, double* feat108) {
    if (w <= 0 || h <= 0) return -1;
    const int64_t n = (int64_t)w * h;
    double* rgb = (double*)malloc(sizeof(double) * n * 3);
    REAL* a = (REAL*)malloc(sizeof(REAL) * n * 3);
    REAL* b = (REAL*)malloc(sizeof(REAL) * n * 3);
    int rc = -1;
    if (!rgb || !a || !b) goto done;
    if (so_yuv420_to_rgb_f64(p1, s1, w, h, bits, matrix, rgb)) goto done;
    for (int64_t i = 0; i < 3 * n; i++) a[i] = (REAL)so_linearize(rgb[i]);
    if (so_yuv420_to_rgb_f64(p2, s2, w, h, bits, matrix, rgb)) goto done;
    for (int64_t i = 0; i < 3 * n; i++) b[i] = (REAL)so_linearize(rgb[i]);
    rc = FN(score_from_linear)(a, b, w, h, score, feat108);
done:
    free(rgb); free(a); free(b);
    return rc;
}
