"""ctypes wrappers of the Butteraugli oracle (oracle/liboracle_butter.so) and
of the reference's own Butteraugli HIP code (oracle/_ref/libvship_ref_butter.so,
GPU). Test infrastructure."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle_butter.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libvship_ref_butter.so")
BO_FLOAT, BO_U16 = 0, 1
D = ctypes.POINTER(ctypes.c_double)
P3 = ctypes.c_void_p * 3

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.bo_blur.argtypes = [D, ctypes.c_int, ctypes.c_int, ctypes.c_int, D]
        L.bo_diffmap_scale.argtypes = [D, D, ctypes.c_int, ctypes.c_int, ctypes.c_double, D]
        L.bo_butteraugli.argtypes = [ctypes.c_int, P3, ctypes.c_int64, P3, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, D, D]
        L.bo_malta.argtypes = [D, D, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, D]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(D)


def _planes(ps):
    return P3(*[p.ctypes.data for p in ps])


def fmt_of(planes):
    return BO_U16 if planes[0].dtype == np.uint16 else BO_FLOAT


def butteraugli(ref, dist, qnorm=2, intensity=203.0, diffmap=False):
    """Norms (Q, 3, inf) of the oracle and, optionally, its distortion map."""
    p1 = [np.ascontiguousarray(p) for p in ref]
    p2 = [np.ascontiguousarray(p) for p in dist]
    h, w = p1[0].shape
    norms = np.zeros(3)
    dm = np.zeros((h, w)) if diffmap else None
    rc = lib().bo_butteraugli(fmt_of(p1), _planes(p1), p1[0].strides[0], _planes(p2), p2[0].strides[0], w, h, qnorm,
                              intensity, _dp(norms), _dp(dm) if diffmap else None)
    assert rc == 0
    return (norms, dm) if diffmap else norms


def blur(src, sigma_index):
    a = np.ascontiguousarray(src, np.float64)
    out = np.zeros_like(a)
    lib().bo_blur(_dp(a), a.shape[1], a.shape[0], sigma_index, _dp(out))
    return out


def malta(l0, l1, hf, w_gt, w_lt, norm1):
    a, b = np.ascontiguousarray(l0, np.float64), np.ascontiguousarray(l1, np.float64)
    acc = np.zeros_like(a)
    lib().bo_malta(_dp(a), _dp(b), a.shape[1], a.shape[0], int(hf), w_gt, w_lt, norm1, _dp(acc))
    return acc


def ref_available():
    return os.path.exists(REF_PATH)


def ref_lib():
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        F = ctypes.POINTER(ctypes.c_float)
        R.vref_butter.argtypes = [ctypes.c_int, P3, P3, ctypes.c_int64, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_float, F, F]
        R.vref_butter_size_ok.argtypes = [ctypes.c_int, ctypes.c_int]
        _ref = R
    return _ref


def ref_size_ok(w, h):
    """False for sizes where the reference's TiledGaussianBlur_Kernel reads its
    sigma 1.2 integral table before the table's buffer (src/butter/
    gaussianblur.hpp:127-130, 154-157): width or height, at full or half
    scale, with a remainder of 1..5 modulo 32."""
    def ok(n):
        return n % 32 == 0 or n % 32 > 5
    return w > 0 and h > 0 and ok(w) and ok(h) and ok((w + 1) // 2) and ok((h + 1) // 2)


def ref_butteraugli(ref, dist, qnorm=2, intensity=203.0):
    """Norms and distortion map of the reference's own HIP code (GPU)."""
    p1 = [np.ascontiguousarray(p) for p in ref]
    p2 = [np.ascontiguousarray(p) for p in dist]
    h, w = p1[0].shape
    if not ref_size_ok(w, h):
        raise ValueError(f"{w}x{h}: the reference's tiled blur reads out of bounds at this size")
    norms = np.zeros(3, np.float32)
    dm = np.zeros((h, w), np.float32)
    F = ctypes.POINTER(ctypes.c_float)
    rc = ref_lib().vref_butter(fmt_of(p1), _planes(p1), _planes(p2), p1[0].strides[0], p2[0].strides[0], w, h,
                               qnorm, intensity, norms.ctypes.data_as(F), dm.ctypes.data_as(F))
    assert rc == 0
    return norms.astype(np.float64), dm.astype(np.float64)


def sample_positions(w, h, n=64, seed=7):
    """Fixed pixel positions, the corners and edges included, for map records."""
    rng = np.random.default_rng(seed)
    pts = [(0, 0), (w - 1, 0), (0, h - 1), (w - 1, h - 1), (w // 2, 0), (0, h // 2)]
    while len(pts) < n:
        pts.append((int(rng.integers(0, w)), int(rng.integers(0, h))))
    return pts
