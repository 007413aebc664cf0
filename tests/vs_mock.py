"""ctypes wrappers of the mock VapourSynth host (tests/vs_mock_host.cpp,
built into tests/_build/libvs_mock_host.so by `python -m vship_amd.build`),
which loads libvship.so as a VapourSynth plugin and requests frames of its
filters from several threads."""
import ctypes
import os

import numpy as np

from vship_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "_build", "libvs_mock_host.so")
VS_HEADER_DIR = os.environ.get("VSHIP_VS_INCLUDE", "/root/reference/include")


def mock():
    L = ctypes.CDLL(MOCK)
    L.vsmock_registration.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    PP = ctypes.POINTER(ctypes.c_void_p)
    ssimu2_args = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [PP, PP] + [ctypes.c_int] * 7 + [
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    L.vsmock_ssimu2.argtypes = ssimu2_args
    L.vsmock_ssimu2_timed.argtypes = ssimu2_args + [ctypes.POINTER(ctypes.c_double)]
    L.vsmock_gpuinfo.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.vsmock_butter.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [PP, PP, ctypes.c_int, ctypes.c_int,
                                                                          ctypes.c_double] + [ctypes.c_int] * 5 + [
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
        ctypes.c_char_p, ctypes.c_int]
    return L


def libpath():
    return api.lib_path().encode()


def run_ssimu2(frames_ref, frames_dist, w, h, stride, w2=None, h2=None, rgbs=(1, 1), num_stream=-1, gpu_id=-1,
               core_threads=4, nthreads=4, timed=False):
    """frames_*: list of 3 (h, stride/4) float32 arrays per frame. Returns
    (rc, scores, output frames equal to the distorted ones, error text), and
    with timed=True also the wall time in ms of the frame requests alone."""
    n = len(frames_ref)
    ptrs_r = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_ref for p in f])
    ptrs_d = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_dist for p in f])
    scores = np.full(n, np.nan)
    matches = ctypes.c_int(0)
    err = ctypes.create_string_buffer(4096)
    loop_ms = ctypes.c_double(0.0)
    rc = mock().vsmock_ssimu2_timed(libpath(), w, h, w2 or w, h2 or h, n, ptrs_r, ptrs_d, stride, rgbs[0], rgbs[1],
                                    num_stream, gpu_id, core_threads, nthreads,
                                    scores.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(matches),
                                    err, len(err), ctypes.byref(loop_ms))
    out = (rc, scores, matches.value, err.value.decode())
    return out + (loop_ms.value,) if timed else out


def run_butter(frames_ref, frames_dist, w, h, stride, w2=None, h2=None, qnorm=-1, intensity=-1.0, distmap=-1,
               num_stream=-1, gpu_id=-1, core_threads=4, nthreads=4):
    n = len(frames_ref)
    ptrs_r = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_ref for p in f])
    ptrs_d = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_dist for p in f])
    norms = np.full((n, 3), np.nan)
    maps = np.zeros((n, h, w), np.float32)
    matches = ctypes.c_int(0)
    err = ctypes.create_string_buffer(4096)
    rc = mock().vsmock_butter(libpath(), w, h, w2 or w, h2 or h, n, ptrs_r, ptrs_d, stride, qnorm, intensity,
                              distmap, num_stream, gpu_id, core_threads, nthreads,
                              norms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                              maps.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(matches), err,
                              len(err))
    return rc, norms, maps, matches.value, err.value.decode()
