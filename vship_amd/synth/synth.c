/*
 * Seeded synthetic frame-pair generator for the SSIMULACRA2 workloads.
 *
 * Reference frames: a field of 32 Gaussian blobs (centre uniform over the
 * field, sigma U[8,64] luma px, one amplitude per plane), composed as
 *     F = 2 * sum_i A_i g_i / (1 + sum_i g_i)
 * so the picture neither saturates nor goes dark whatever the resolution,
 * plus uniform noise U[-8,8] (scaled by 2^(bits-8)), rounded and clipped.
 * Video clips translate the field by `shift` luma px per frame.
 *
 *   RGB (4:4:4, full range, 8-bit): A ~ U[0,255], value = F + noise, clip [0,255]
 *   YUV 4:2:0 limited range:        luma A ~ U[0,255], Y = 16s + (219/255) s F
 *                                   chroma A ~ U[-127.5,127.5] on a half-res
 *                                   field, C = 128s + (224/255) s F
 *                                   s = 2^(bits-8); clip to [16s,235s] / [16s,240s]
 *
 * Distorted frames:
 *   RGB: reference + N(0, 4^2), rounded, clipped.
 *   YUV: per plane, 8x8 orthonormal DCT-II (edge-replicated partial blocks),
 *        coef <- round(coef/q)*q with q = 16 s (luma) / 24 s (chroma),
 *        inverse DCT, rounded, clipped; then on a seeded 50% of each plane's
 *        8x8 blocks the pixels are replaced by the rounded 3x3 box mean of the
 *        quantised plane (edge-replicated at the frame border).
 *
 * All randomness is SplitMix64 keyed by (seed, frame, plane, row/purpose),
 * so any frame can be produced independently and in parallel.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NBLOBS 32

typedef struct {
    int w, h, kind, bits, nframes, shift;
    uint64_t seed;
    int pw[3], ph[3];     /* plane sizes */
    int fw[3];            /* extended field widths */
    float* field[3];      /* ph x fw fields (F, before offset/scale) */
} synth_clip;

static uint64_t sm64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t mix(uint64_t a, uint64_t b) {
    uint64_t s = a ^ (b * 0xD1B54A32D192ED03ull);
    return sm64(&s);
}
static double u01(uint64_t* s) { return (double)(sm64(s) >> 11) * (1.0 / 9007199254740992.0); }

static double rnd(double v) { return floor(v + 0.5); }
static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

void* synth_new(int w, int h, int kind, int bits, int nframes, int shift, uint64_t seed) {
    if (w <= 0 || h <= 0 || nframes <= 0 || bits < 8 || bits > 16) return NULL;
    synth_clip* c = (synth_clip*)calloc(1, sizeof(synth_clip));
    c->w = w; c->h = h; c->kind = kind; c->bits = bits; c->nframes = nframes; c->shift = shift; c->seed = seed;
    for (int p = 0; p < 3; p++) {
        const int sub = (kind == 1 && p > 0) ? 1 : 0;
        c->pw[p] = sub ? (w + 1) / 2 : w;
        c->ph[p] = sub ? (h + 1) / 2 : h;
        const int pshift = sub ? shift / 2 : shift;
        c->fw[p] = c->pw[p] + pshift * (nframes - 1);
    }
    /* blobs in luma coordinates over the extended luma field */
    double cx[NBLOBS], cy[NBLOBS], sg[NBLOBS], amp[3][NBLOBS];
    uint64_t st = mix(seed, 0xB10B);
    for (int i = 0; i < NBLOBS; i++) {
        cx[i] = u01(&st) * c->fw[0];
        cy[i] = u01(&st) * h;
        sg[i] = 8.0 + 56.0 * u01(&st);
        for (int p = 0; p < 3; p++) {
            const double a = u01(&st);
            amp[p][i] = (kind == 1 && p > 0) ? 255.0 * a - 127.5 : 255.0 * a;
        }
    }
    for (int p = 0; p < 3; p++) {
        const int fw = c->fw[p], fh = c->ph[p];
        const double sc = (kind == 1 && p > 0) ? 2.0 : 1.0; /* plane px -> luma px */
        float* F = (float*)malloc(sizeof(float) * (size_t)fw * fh);
        double* gx = (double*)malloc(sizeof(double) * (size_t)NBLOBS * fw);
        double* gy = (double*)malloc(sizeof(double) * (size_t)NBLOBS * fh);
        for (int i = 0; i < NBLOBS; i++) {
            const double inv = 1.0 / (2.0 * sg[i] * sg[i]);
            for (int x = 0; x < fw; x++) {
                const double d = (x + 0.5) * sc - cx[i];
                gx[(size_t)i * fw + x] = exp(-d * d * inv);
            }
            for (int y = 0; y < fh; y++) {
                const double d = (y + 0.5) * sc - cy[i];
                gy[(size_t)i * fh + y] = exp(-d * d * inv);
            }
        }
        #pragma omp parallel for schedule(static)
        for (int y = 0; y < fh; y++) {
            for (int x = 0; x < fw; x++) {
                double num = 0.0, den = 0.0;
                for (int i = 0; i < NBLOBS; i++) {
                    const double g = gx[(size_t)i * fw + x] * gy[(size_t)i * fh + y];
                    num += amp[p][i] * g;
                    den += g;
                }
                F[(size_t)y * fw + x] = (float)(2.0 * num / (1.0 + den));
            }
        }
        free(gx);
        free(gy);
        c->field[p] = F;
    }
    return c;
}

void synth_free(void* h) {
    synth_clip* c = (synth_clip*)h;
    if (!c) return;
    for (int p = 0; p < 3; p++) free(c->field[p]);
    free(c);
}

void synth_plane_dims(void* h, int dims[6]) {
    synth_clip* c = (synth_clip*)h;
    for (int p = 0; p < 3; p++) { dims[2 * p] = c->pw[p]; dims[2 * p + 1] = c->ph[p]; }
}

static void put(uint8_t* plane, int bpp, int64_t idx, double v) {
    if (bpp == 1) plane[idx] = (uint8_t)v;
    else { const uint16_t u = (uint16_t)v; memcpy(plane + 2 * idx, &u, 2); }
}

/* reference plane p of frame f as doubles (rounded, clipped code values) */
static void ref_plane(const synth_clip* c, int f, int p, double* out) {
    const int pw = c->pw[p], ph = c->ph[p], fw = c->fw[p];
    const int sub = (c->kind == 1 && p > 0);
    const int off = (sub ? c->shift / 2 : c->shift) * f;
    const double s = (double)(1 << (c->bits - 8));
    double base, gain, lo, hi;
    if (c->kind == 1) {
        base = sub ? 128.0 * s : 16.0 * s;
        gain = (sub ? 224.0 : 219.0) / 255.0 * s;
        lo = 16.0 * s;
        hi = (sub ? 240.0 : 235.0) * s;
    } else {
        base = 0.0; gain = 1.0; lo = 0.0; hi = 255.0;
    }
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < ph; y++) {
        uint64_t st = mix(mix(mix(c->seed, 0x4E01 + (uint64_t)f), (uint64_t)p), (uint64_t)y);
        const float* F = c->field[p] + (size_t)y * fw + off;
        for (int x = 0; x < pw; x++) {
            const double noise = (16.0 * u01(&st) - 8.0) * s;
            out[(size_t)y * pw + x] = clampd(rnd(base + gain * F[x] + noise), lo, hi);
        }
    }
}

static void dct_matrix(double D[8][8]) {
    for (int k = 0; k < 8; k++)
        for (int n = 0; n < 8; n++)
            D[k][n] = (k == 0 ? sqrt(1.0 / 8.0) : 0.5) * cos((2 * n + 1) * k * M_PI / 16.0);
}

static void distort_plane(const synth_clip* c, int f, int p, const double* ref, double* out) {
    const int pw = c->pw[p], ph = c->ph[p];
    const double s = (double)(1 << (c->bits - 8));
    const int sub = (c->kind == 1 && p > 0);
    const double q = (sub ? 24.0 : 16.0) * s;
    const double lo = 16.0 * s, hi = (sub ? 240.0 : 235.0) * s;
    const int bw = (pw + 7) / 8, bh = (ph + 7) / 8;
    double D[8][8];
    dct_matrix(D);
    double* qz = (double*)malloc(sizeof(double) * (size_t)pw * ph);
    #pragma omp parallel for schedule(static)
    for (int by = 0; by < bh; by++) {
        for (int bx = 0; bx < bw; bx++) {
            double B[8][8], T[8][8], C[8][8];
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) {
                    int yy = by * 8 + i, xx = bx * 8 + j;
                    if (yy >= ph) yy = ph - 1;
                    if (xx >= pw) xx = pw - 1;
                    B[i][j] = ref[(size_t)yy * pw + xx];
                }
            for (int k = 0; k < 8; k++)
                for (int j = 0; j < 8; j++) {
                    double a = 0;
                    for (int n = 0; n < 8; n++) a += D[k][n] * B[n][j];
                    T[k][j] = a;
                }
            for (int k = 0; k < 8; k++)
                for (int l = 0; l < 8; l++) {
                    double a = 0;
                    for (int n = 0; n < 8; n++) a += T[k][n] * D[l][n];
                    C[k][l] = rnd(a / q) * q;
                }
            for (int n = 0; n < 8; n++)
                for (int l = 0; l < 8; l++) {
                    double a = 0;
                    for (int k = 0; k < 8; k++) a += D[k][n] * C[k][l];
                    T[n][l] = a;
                }
            for (int n = 0; n < 8; n++)
                for (int m = 0; m < 8; m++) {
                    double a = 0;
                    for (int l = 0; l < 8; l++) a += T[n][l] * D[l][m];
                    const int yy = by * 8 + n, xx = bx * 8 + m;
                    if (yy < ph && xx < pw) qz[(size_t)yy * pw + xx] = clampd(rnd(a), lo, hi);
                }
        }
    }
    #pragma omp parallel for schedule(static)
    for (int by = 0; by < bh; by++) {
        for (int bx = 0; bx < bw; bx++) {
            uint64_t st = mix(mix(mix(c->seed, 0xB1A5 + (uint64_t)f), (uint64_t)p), (uint64_t)by * bw + bx);
            const int blur = (sm64(&st) >> 63) != 0;
            for (int i = 0; i < 8; i++)
                for (int j = 0; j < 8; j++) {
                    const int yy = by * 8 + i, xx = bx * 8 + j;
                    if (yy >= ph || xx >= pw) continue;
                    double v = qz[(size_t)yy * pw + xx];
                    if (blur) {
                        double a = 0;
                        for (int dy = -1; dy <= 1; dy++)
                            for (int dx = -1; dx <= 1; dx++) {
                                int y2 = yy + dy, x2 = xx + dx;
                                y2 = y2 < 0 ? 0 : (y2 >= ph ? ph - 1 : y2);
                                x2 = x2 < 0 ? 0 : (x2 >= pw ? pw - 1 : x2);
                                a += qz[(size_t)y2 * pw + x2];
                            }
                        v = clampd(rnd(a / 9.0), lo, hi);
                    }
                    out[(size_t)yy * pw + xx] = v;
                }
        }
    }
    free(qz);
}

static void noise_plane(const synth_clip* c, int f, int p, const double* ref, double* out) {
    const int pw = c->pw[p], ph = c->ph[p];
    #pragma omp parallel for schedule(static)
    for (int y = 0; y < ph; y++) {
        uint64_t st = mix(mix(mix(c->seed, 0x6A55 + (uint64_t)f), (uint64_t)p), (uint64_t)y);
        for (int x = 0; x < pw; x++) {
            const double u1 = u01(&st), u2 = u01(&st);
            const double g = sqrt(-2.0 * log(1.0 - u1)) * cos(2.0 * M_PI * u2);
            out[(size_t)y * pw + x] = clampd(rnd(ref[(size_t)y * pw + x] + 4.0 * g), 0.0, 255.0);
        }
    }
}

/* Writes frame f's planes into tightly packed buffers (row stride = plane
 * width * bytes per sample; 1 byte for 8-bit, 2 bytes little-endian above).
 * Either output set may be NULL. */
int synth_frame(void* h, int f, uint8_t* ref[3], uint8_t* dist[3]) {
    synth_clip* c = (synth_clip*)h;
    if (!c || f < 0 || f >= c->nframes) return -1;
    const int bpp = c->bits > 8 ? 2 : 1;
    for (int p = 0; p < 3; p++) {
        const size_t n = (size_t)c->pw[p] * c->ph[p];
        double* r = (double*)malloc(sizeof(double) * n);
        double* d = (double*)malloc(sizeof(double) * n);
        ref_plane(c, f, p, r);
        if (dist) {
            if (c->kind == 1) distort_plane(c, f, p, r, d);
            else noise_plane(c, f, p, r, d);
        }
        for (size_t i = 0; i < n; i++) {
            if (ref && ref[p]) put(ref[p], bpp, (int64_t)i, r[i]);
            if (dist && dist[p]) put(dist[p], bpp, (int64_t)i, d[i]);
        }
        free(r);
        free(d);
    }
    return 0;
}
