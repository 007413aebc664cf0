/*
 * CPU oracle for Butteraugli (the reference's src/butter/), FP64.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * CPU baseline may load it, as a checker; nothing in vship_amd/ uses it.
 *
 * Restates, stage by stage, with double arithmetic:
 *   linearisation           src/butter/colors.hpp:7-13
 *   2x2 downsample          src/butter/downupsample.hpp:5-20
 *   renormalised Gaussians  src/butter/gaussianblur.hpp:3-50,55-85,206-241
 *   opsin dynamics          src/butter/colors.hpp:3-5,26-84
 *   frequency separation    src/butter/separatefrequencies.hpp:3-175
 *   Malta line filters      src/butter/maltaDiff.hpp (16 lines each, HF and LF)
 *   L2 / asymmetric L2      src/butter/simplerdiff.hpp:38-111
 *   psycho mask             src/butter/maskPsycho.hpp
 *   diff map                src/butter/combineMasks.hpp, src/butter/main.hpp:22-137
 *   norms                   src/butter/diffnorms.hpp:78-120
 *
 * Pinning: tests/test_butter_oracle.py checks it against closed forms
 * (identical images give an all-zero map; blurs of constants are the
 * constant; the Malta and L2 terms against brute-force sums) and against the
 * reference's own HIP code run on the GPU on the same inputs
 * (tests/golden/reference_butter.json, made by tests/golden/make_golden.py
 * through oracle/ref_butter_shim.hip).
 */
#ifndef BUTTER_ORACLE_H
#define BUTTER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* input formats: 0 float RGB in [0,1], 1 uint16 RGB (full range) */
enum { BO_FLOAT = 0, BO_U16 = 1 };

/* renormalised separable Gaussian of the reference's table: sigma index
   0..4 = 1.2, 1.56416327805, 2.7, 3.22489901262, 7.15593339443 */
void bo_blur(const double* src, int w, int h, int sigma_index, double* dst);

/* one scale's distortion map from linear RGB planes (3*w*h each) */
int bo_diffmap_scale(const double* lin1, const double* lin2, int w, int h, double intensity, double* diffmap);

/* full metric: the two-scale map (w*h, may be NULL) and the three norms
   (Q-norm, 3-norm, inf-norm) */
int bo_butteraugli(int fmt, const uint8_t* const p1[3], int64_t stride1, const uint8_t* const p2[3],
                   int64_t stride2, int w, int h, int qnorm, double intensity, double norms3[3], double* diffmap);

/* the Malta term of one plane pair (hf: 1 for the 9-tap lines, 0 for the
   5-tap ones), added to acc (w*h) */
void bo_malta(const double* l0, const double* l1, int w, int h, int hf, double w_0gt1, double w_0lt1,
              double norm1, double* acc);

#ifdef __cplusplus
}
#endif

#endif
