"""ctypes binding of the CPU oracle (oracle/liboracle_ssimu2.so).

Test infrastructure: only tests, the smoke check and the benchmark's CPU
baseline leg load this module.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_ssimu2.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libvship_ref.so")

_lib = None

D = ctypes.POINTER(ctypes.c_double)
F = ctypes.POINTER(ctypes.c_float)
I64 = ctypes.c_int64
P3 = ctypes.c_void_p * 3
S3 = ctypes.c_int64 * 3


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.so_total_scale_size.restype = ctypes.c_int64
        L.so_total_scale_size.argtypes = [ctypes.c_int, ctypes.c_int]
        L.so_linearize.restype = ctypes.c_double
        L.so_linearize.argtypes = [ctypes.c_double]
        L.so_final_score.restype = ctypes.c_double
        L.so_final_score.argtypes = [D]
        for name in ("so_downsample", "so_blur"):
            getattr(L, name).argtypes = [D, ctypes.c_int, ctypes.c_int, D]
        L.so_xyb.argtypes = [D, D]
        L.so_gaussian_taps.argtypes = [D]
        L.so_scale_features.argtypes = [D, D, ctypes.c_int, ctypes.c_int, D]
        L.so_channel_maps.argtypes = [D, D, ctypes.c_int, ctypes.c_int, D, D, D]
        for name in ("so_score_rgb_f64", "so_score_rgb_f32"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_int, P3, I64, P3, I64, ctypes.c_int, ctypes.c_int, D, D]
        for name in ("so_score_yuv420_f64", "so_score_yuv420_f32"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [P3, S3, P3, S3, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, D, D]
        for name in ("so_score_yuv_f64", "so_score_yuv_f32"):
            fn = getattr(L, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [P3, S3, P3, S3] + [ctypes.c_int] * 7 + [D, D]
        L.so_yuv_to_rgb_f64.restype = ctypes.c_int
        L.so_yuv_to_rgb_f64.argtypes = [P3, S3] + [ctypes.c_int] * 7 + [D]
        L.so_chroma_sample.restype = ctypes.c_double
        L.so_chroma_sample.argtypes = [ctypes.c_void_p, ctypes.c_int64] + [ctypes.c_int] * 8
        L.so_yuv420_to_rgb_f64.restype = ctypes.c_int
        L.so_yuv420_to_rgb_f64.argtypes = [P3, S3, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, D]
        L.so_score_linear_f64.restype = ctypes.c_int
        L.so_score_linear_f64.argtypes = [D, D, ctypes.c_int, ctypes.c_int, D, D]
        L.so_set_threads.argtypes = [ctypes.c_int]
        L.so_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(D)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def total_scale_size(w, h):
    return lib().so_total_scale_size(w, h)


def scale_dims(w, h):
    d = (ctypes.c_int * 12)()
    lib().so_scale_dims(w, h, d)
    return [(d[2 * s], d[2 * s + 1]) for s in range(6)]


def linearize(v):
    return lib().so_linearize(float(v))


def gaussian_taps():
    g = np.zeros(17)
    lib().so_gaussian_taps(_dp(g))
    return g


def xyb(rgb):
    a = c64(rgb)
    out = np.zeros(3)
    lib().so_xyb(_dp(a), _dp(out))
    return out


def downsample(img):
    a = c64(img)
    h, w = a.shape
    out = np.zeros(((h - 1) // 2 + 1, (w - 1) // 2 + 1))
    lib().so_downsample(_dp(a), w, h, _dp(out))
    return out


def blur(img):
    a = c64(img)
    h, w = a.shape
    out = np.zeros_like(a)
    lib().so_blur(_dp(a), w, h, _dp(out))
    return out


def channel_maps(i1, i2):
    a, b = c64(i1), c64(i2)
    h, w = a.shape
    d, art, det = np.zeros_like(a), np.zeros_like(a), np.zeros_like(a)
    lib().so_channel_maps(_dp(a), _dp(b), w, h, _dp(d), _dp(art), _dp(det))
    return d, art, det


def scale_features(xyb1, xyb2):
    """xyb*: (3, h, w). Returns (3 ch, 3 maps, 2 norms)."""
    a, b = c64(xyb1), c64(xyb2)
    _, h, w = a.shape
    out = np.zeros(18)
    lib().so_scale_features(_dp(a), _dp(b), w, h, _dp(out))
    return out.reshape(3, 3, 2)


def final_score(f108):
    f = c64(f108)
    assert f.shape == (108,)
    return lib().so_final_score(_dp(f))


def _planes(arrs):
    return P3(*[a.ctypes.data for a in arrs])


def score_rgb(fmt, img1, img2, precision="f64"):
    """img*: list of 3 2-D arrays (u8/u16/f16/f32) or (3, h, w) arrays; row stride from the arrays."""
    p1 = [np.ascontiguousarray(p) for p in img1]
    p2 = [np.ascontiguousarray(p) for p in img2]
    h, w = p1[0].shape
    s1, s2 = p1[0].strides[0], p2[0].strides[0]
    score = ctypes.c_double()
    feat = np.zeros(108)
    fn = lib().so_score_rgb_f64 if precision == "f64" else lib().so_score_rgb_f32
    rc = fn(fmt, _planes(p1), s1, _planes(p2), s2, w, h, ctypes.byref(score), _dp(feat))
    assert rc == 0
    return score.value, feat


def score_rgb_strided(fmt, planes1, stride1, planes2, stride2, w, h):
    """planes*: 3 byte buffers (numpy uint8 1-D) holding rows of `stride` bytes."""
    score = ctypes.c_double()
    feat = np.zeros(108)
    rc = lib().so_score_rgb_f64(fmt, _planes(planes1), stride1, _planes(planes2), stride2, w, h,
                                ctypes.byref(score), _dp(feat))
    assert rc == 0
    return score.value, feat


def yuv_to_rgb(planes, bits, matrix):
    p = [np.ascontiguousarray(x) for x in planes]
    h, w = p[0].shape
    out = np.zeros((3, h, w))
    rc = lib().so_yuv420_to_rgb_f64(_planes(p), S3(*[x.strides[0] for x in p]), w, h, bits, matrix, _dp(out))
    assert rc == 0
    return out


def score_yuv420(ref, dist, bits, matrix, precision="f64"):
    p1 = [np.ascontiguousarray(x) for x in ref]
    p2 = [np.ascontiguousarray(x) for x in dist]
    h, w = p1[0].shape
    score = ctypes.c_double()
    feat = np.zeros(108)
    fn = lib().so_score_yuv420_f64 if precision == "f64" else lib().so_score_yuv420_f32
    rc = fn(_planes(p1), S3(*[x.strides[0] for x in p1]), _planes(p2), S3(*[x.strides[0] for x in p2]),
            w, h, bits, matrix, ctypes.byref(score), _dp(feat))
    assert rc == 0
    return score.value, feat


# YUV layouts: name -> (log2 horizontal, log2 vertical chroma subsampling),
# and chroma locations (include/VshipSSIMU2Batch.h)
LAYOUTS = {"420": (1, 1), "422": (1, 0), "444": (0, 0), "440": (0, 1), "411": (2, 0), "410": (2, 1)}
SO_LOC_LEFT, SO_LOC_CENTER, SO_LOC_TOPLEFT, SO_LOC_TOP = range(4)


def chroma_sample(plane, w, h, bits, sw, sh, loc, x, y):
    p = np.ascontiguousarray(plane)
    return lib().so_chroma_sample(p.ctypes.data, p.strides[0], w, h, bits, sw, sh, loc, x, y)


def yuv_to_rgb_any(planes, bits, matrix, sw, sh, loc):
    p = [np.ascontiguousarray(x) for x in planes]
    h, w = p[0].shape
    out = np.zeros((3, h, w))
    rc = lib().so_yuv_to_rgb_f64(_planes(p), S3(*[x.strides[0] for x in p]), w, h, bits, matrix, sw, sh, loc,
                                 _dp(out))
    assert rc == 0
    return out


def score_yuv(ref, dist, bits, matrix, sw, sh, loc, precision="f64"):
    p1 = [np.ascontiguousarray(x) for x in ref]
    p2 = [np.ascontiguousarray(x) for x in dist]
    h, w = p1[0].shape
    score = ctypes.c_double()
    feat = np.zeros(108)
    fn = lib().so_score_yuv_f64 if precision == "f64" else lib().so_score_yuv_f32
    rc = fn(_planes(p1), S3(*[x.strides[0] for x in p1]), _planes(p2), S3(*[x.strides[0] for x in p2]),
            w, h, bits, matrix, sw, sh, loc, ctypes.byref(score), _dp(feat))
    assert rc == 0
    return score.value, feat


def score_linear(lin1, lin2):
    a, b = c64(lin1), c64(lin2)
    _, h, w = a.shape
    score = ctypes.c_double()
    feat = np.zeros(108)
    assert lib().so_score_linear_f64(_dp(a), _dp(b), w, h, ctypes.byref(score), _dp(feat)) == 0
    return score.value, feat


def set_threads(n):
    lib().so_set_threads(int(n))


def get_threads():
    return lib().so_get_threads()


SO_RGB_U8, SO_RGB_U16, SO_RGB_F16, SO_RGB_F32 = 0, 1, 2, 3
SO_MAT_BT709, SO_MAT_BT601, SO_MAT_BT2020NCL = 0, 1, 2


# ---------------------------------------------------------------------------
# the reference's own HIP implementation, built from its sources by
# `make -C oracle ref` (GPU only; absent where the sources were not mounted)
# ---------------------------------------------------------------------------
_ref = None


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        R.vref_ssimu2.restype = ctypes.c_int
        R.vref_ssimu2.argtypes = [ctypes.c_int, P3, P3, I64, I64, ctypes.c_int, ctypes.c_int, D]
        R.vref_ssimu2_features.restype = ctypes.c_int
        R.vref_ssimu2_features.argtypes = [ctypes.c_int, P3, P3, I64, I64, ctypes.c_int, ctypes.c_int, F]
        R.vref_ssimu2_time.restype = ctypes.c_int
        R.vref_ssimu2_time.argtypes = [ctypes.c_int, P3, P3, I64, I64, ctypes.c_int, ctypes.c_int, ctypes.c_int, D,
                                       D]
        _ref = R
    return _ref


def ref_score_rgb(fmt, img1, img2):
    """Score and features of the reference's HIP code on planar RGB (same formats as score_rgb)."""
    p1 = [np.ascontiguousarray(p) for p in img1]
    p2 = [np.ascontiguousarray(p) for p in img2]
    h, w = p1[0].shape
    s = ctypes.c_double()
    R = ref_lib()
    rc = R.vref_ssimu2(fmt, _planes(p1), _planes(p2), p1[0].strides[0], p2[0].strides[0], w, h, ctypes.byref(s))
    assert rc == 0, rc
    f = np.zeros(108, np.float32)
    rc = R.vref_ssimu2_features(fmt, _planes(p1), _planes(p2), p1[0].strides[0], p2[0].strides[0], w, h,
                                f.ctypes.data_as(F))
    assert rc == 0, rc
    return s.value, f.astype(np.float64)


def ref_time_rgb(fmt, img1, img2, iters):
    """(score, ms per call) of the reference's per-frame path: one handle, host planes."""
    p1 = [np.ascontiguousarray(p) for p in img1]
    p2 = [np.ascontiguousarray(p) for p in img2]
    h, w = p1[0].shape
    s, ms = ctypes.c_double(), ctypes.c_double()
    rc = ref_lib().vref_ssimu2_time(fmt, _planes(p1), _planes(p2), p1[0].strides[0], p2[0].strides[0], w, h,
                                    int(iters), ctypes.byref(s), ctypes.byref(ms))
    assert rc == 0, rc
    return s.value, ms.value / iters


def reference_feature_mismatch(f_ref, f_oracle):
    """Indices where the reference's FP32 HIP features (f_ref) leave the band
    around the FP64 oracle's (f_oracle) that FP32 rounding explains: the
    artifact and detail-loss features within 1e-3 relative at every scale;
    the SSIM features within 5e-4 relative at scales 0-1 and 3e-5 absolute at
    every scale (the reference's SSIM map takes sigma^2 = E[x^2] - mu^2 in
    FP32, whose cancellation at the coarse scales, where d is tiny, reaches
    1e-5 absolute). Feature index c*36 + s*6 + n*3 + m (m = 0 SSIM)."""
    f_ref = np.asarray(f_ref, np.float64)
    f_oracle = np.asarray(f_oracle, np.float64)
    idx = np.arange(108)
    scale = (idx % 36) // 6
    ssim = idx % 3 == 0
    d = np.abs(f_ref - f_oracle)
    a = np.abs(f_oracle)
    ok = np.where(ssim, d <= 3e-5, d <= 1e-3 * a + 1e-7)
    ok &= ~(ssim & (scale < 2)) | (d <= 5e-4 * a + 1e-7)
    return np.nonzero(~ok)[0].tolist()
