"""Seeded synthetic frame-pair workloads (ctypes binding of synth.c).

The generator is C (OpenMP) so the benchmark can build 240 1080p pairs in
seconds. Planes come back as numpy arrays: uint8 for 8-bit, uint16 above.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvship_synth.so")
_lib = None

RGB, YUV420 = 0, 1
RESOLUTIONS = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libvship_synth.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.synth_new.restype = ctypes.c_void_p
        L.synth_new.argtypes = [ctypes.c_int] * 6 + [ctypes.c_uint64]
        L.synth_free.argtypes = [ctypes.c_void_p]
        L.synth_plane_dims.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.synth_frame.restype = ctypes.c_int
        L.synth_frame.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _lib = L
    return _lib


class Clip:
    """A clip of `nframes` reference/distorted pairs (see synth.c for the definition)."""

    def __init__(self, w, h, kind=YUV420, bits=8, nframes=1, shift=2, seed=0):
        L = _load()
        self.w, self.h, self.kind, self.bits, self.nframes = w, h, kind, bits, nframes
        self._h = L.synth_new(w, h, kind, bits, nframes, shift, seed)
        if not self._h:
            raise ValueError("bad synthetic clip parameters")
        dims = (ctypes.c_int * 6)()
        L.synth_plane_dims(self._h, dims)
        self.plane_dims = [(dims[2 * p], dims[2 * p + 1]) for p in range(3)]
        self.dtype = np.uint8 if bits <= 8 else np.uint16

    def frame(self, f, want_ref=True, want_dist=True):
        ref = [np.empty((ph, pw), self.dtype) for pw, ph in self.plane_dims] if want_ref else None
        dist = [np.empty((ph, pw), self.dtype) for pw, ph in self.plane_dims] if want_dist else None
        P3 = ctypes.c_void_p * 3
        rp = P3(*[a.ctypes.data for a in ref]) if ref else None
        dp = P3(*[a.ctypes.data for a in dist]) if dist else None
        rc = _load().synth_frame(self._h, f, rp, dp)
        if rc != 0:
            raise IndexError(f"frame {f} out of range")
        return ref, dist

    def close(self):
        if self._h:
            _load().synth_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def config0(seed=0):
    """256x256 planar RGB8 full range, one pair (reference + Gaussian noise sigma 4)."""
    c = Clip(256, 256, RGB, 8, 1, 0, seed)
    r, d = c.frame(0)
    c.close()
    return r, d


def video_clip(w, h, bits, nframes, seed):
    """YUV420 limited-range clip translating 2 px/frame, DCT-quantised + block-blurred distortion."""
    return Clip(w, h, YUV420, bits, nframes, 2, seed)


def batch_resolutions(n, seed):
    """Resolutions of an uneven batch, drawn uniformly from RESOLUTIONS by seed."""
    rng = np.random.default_rng(seed)
    return [RESOLUTIONS[i] for i in rng.integers(0, len(RESOLUTIONS), n)]
