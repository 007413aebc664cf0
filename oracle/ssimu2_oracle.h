/*
 * SSIMULACRA2 oracle: a plain-C CPU restatement of Vship's SSIMULACRA2 math.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links or calls
 * this code; only the test suite, the smoke check and the CPU-baseline leg of
 * the benchmark use it, and only as the checker (or as the timed CPU
 * baseline). It is built into liboracle_ssimu2.so by oracle/Makefile.
 *
 * Two builds of the same source exist:
 *   *_f64  every stage in double  -> the correctness oracle
 *   *_f32  blur/maps in float, sums in double -> the timed CPU baseline
 *
 * Stage-by-stage correspondence with the reference (pkp24/Vship):
 *   input unpack        src/util/float3operations.hpp:197-218
 *   linearisation       src/ssimu2/makeXYB.hpp:52-64
 *   pyramid sizes       src/ssimu2/main.hpp:34-42
 *   2x2 downsample      src/ssimu2/downsample.hpp:3-19
 *   positive XYB        src/ssimu2/makeXYB.hpp:4-50
 *   Gaussian taps       src/ssimu2/gaussianblur.hpp:7-18, src/util/preprocessor.hpp:89-92
 *   separable FIR blur  src/ssimu2/gaussianblur.hpp:24-92
 *   SSIM / edge maps    src/ssimu2/score.hpp:113-196
 *   norms               src/ssimu2/score.hpp:244-267
 *   feature order       src/ssimu2/main.hpp:82-95
 *   weights + score     src/ssimu2/score.hpp:272-401
 *
 * The YUV front end has no counterpart inside the reference's SSIMULACRA2
 * path (the reference receives R'G'B' from VapourSynth resize / zimg); the
 * definition used here is stated in so_yuv420_to_rgb_* and so_yuv_to_rgb_*
 * and is shared with the GPU front end. Its chroma positions and
 * Catmull-Rom weights follow the reference's FFVship upsampling kernels
 * (src/ffvship_utility/gpuColorToLinear/chromaUpsample.hpp).
 *
 * Pinning: every function is checked by tests/test_oracle.py against
 * closed forms, brute force and the reference's own known answer
 * (identical inputs, any row padding => exactly 100,
 * test/basic_metric_math_test.cpp:40-58). The YUV front end is pinned only
 * by its own closed forms (constant planes, exact Catmull-Rom weights,
 * reproduction of linear chroma ramps at the stated sample positions for
 * every layout and location); its equality with zimg / VapourSynth resize
 * is "parity unpinned".
 */
#ifndef SSIMU2_ORACLE_H
#define SSIMU2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RGB sample formats of the reference's InputMemType (src/util/preprocessor.hpp:94) */
enum { SO_RGB_U8 = 0, SO_RGB_U16 = 1, SO_RGB_F16 = 2, SO_RGB_F32 = 3 };
/* YUV matrices of the front end */
enum { SO_MAT_BT709 = 0, SO_MAT_BT601 = 1, SO_MAT_BT2020NCL = 2 };
/* chroma sample locations (AVChromaLocation subset) */
enum { SO_LOC_LEFT = 0, SO_LOC_CENTER = 1, SO_LOC_TOPLEFT = 2, SO_LOC_TOP = 3 };

/* ---- stage functions (double precision) ------------------------------ */
int64_t so_total_scale_size(int w, int h);
void so_scale_dims(int w, int h, int dims[12]); /* w0,h0,...,w5,h5 */
void so_gaussian_taps(double g[17]);
double so_linearize(double v);
void so_xyb(const double rgb[3], double out[3]);
void so_downsample(const double* src, int w, int h, double* dst);
void so_blur(const double* src, int w, int h, double* dst);
/* out18[ch*6 + map*2 + norm], ch in X,Y,B; map in ssim,artifact,detail; norm in 1,4 */
void so_scale_features(const double* xyb1, const double* xyb2, int w, int h, double out18[18]);
double so_final_score(const double f[108]);
/* Per-pixel maps of one channel, for map-level checks: d_ssim, artifact, detail */
void so_channel_maps(const double* i1, const double* i2, int w, int h,
                     double* dssim, double* art, double* det);

/* ---- full pipeline, planar RGB input --------------------------------- */
int so_score_rgb_f64(int fmt, const uint8_t* const p1[3], int64_t stride1,
                     const uint8_t* const p2[3], int64_t stride2, int w, int h,
                     double* score, double* feat108);
int so_score_rgb_f32(int fmt, const uint8_t* const p1[3], int64_t stride1,
                     const uint8_t* const p2[3], int64_t stride2, int w, int h,
                     double* score, double* feat108);

/* ---- YUV 4:2:0 front end and full pipeline --------------------------- */
int so_yuv420_to_rgb_f64(const uint8_t* const p[3], const int64_t stride[3], int w, int h,
                         int bits, int matrix, double* rgb /* 3*w*h planar */);
int so_score_yuv420_f64(const uint8_t* const p1[3], const int64_t s1[3],
                        const uint8_t* const p2[3], const int64_t s2[3], int w, int h,
                        int bits, int matrix, double* score, double* feat108);
int so_score_yuv420_f32(const uint8_t* const p1[3], const int64_t s1[3],
                        const uint8_t* const p2[3], const int64_t s2[3], int w, int h,
                        int bits, int matrix, double* score, double* feat108);

/* ---- any planar YUV layout: log2 chroma subsampling sw (0..2), sh (0..1),
   chroma location loc; 4:2:0 with SO_LOC_LEFT is so_yuv420_* ------------- */
double so_chroma_sample(const uint8_t* plane, int64_t stride, int w, int h, int bits, int sw, int sh, int loc,
                        int x, int y);
int so_yuv_to_rgb_f64(const uint8_t* const p[3], const int64_t stride[3], int w, int h, int bits, int matrix,
                      int sw, int sh, int loc, double* rgb /* 3*w*h planar */);
int so_score_yuv_f64(const uint8_t* const p1[3], const int64_t s1[3], const uint8_t* const p2[3],
                     const int64_t s2[3], int w, int h, int bits, int matrix, int sw, int sh, int loc,
                     double* score, double* feat108);
int so_score_yuv_f32(const uint8_t* const p1[3], const int64_t s1[3], const uint8_t* const p2[3],
                     const int64_t s2[3], int w, int h, int bits, int matrix, int sw, int sh, int loc,
                     double* score, double* feat108);

/* Score from already linear-RGB planes at scale 0 (3*w*h each, planar). */
int so_score_linear_f64(const double* lin1, const double* lin2, int w, int h,
                        double* score, double* feat108);

/* Number of OpenMP threads the f32 build may use (0 = runtime default). */
void so_set_threads(int n);
int so_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
