"""The VapourSynth plugin (vship.SSIMULACRA2, vship.BUTTERAUGLI, vship.GpuInfo) through a mock
VapourSynth host (tests/vs_mock_host.cpp): registration, argument and error
handling on the CPU; scores, output frames and reference counting on the GPU.
The plugin's own declarations of the VapourSynth API (vship_amd/csrc/vs4.hpp)
are pinned against VapourSynth's VapourSynth4.h where that header is present.
"""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

from vship_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MOCK = os.path.join(ROOT, "tests", "_build", "libvs_mock_host.so")
VS_HEADER_DIR = os.environ.get("VSHIP_VS_INCLUDE", "/root/reference/include")

needs_mock = pytest.mark.skipif(not os.path.exists(MOCK), reason="mock host not built (python -m vship_amd.build)")


def mock():
    L = ctypes.CDLL(MOCK)
    L.vsmock_registration.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    PP = ctypes.POINTER(ctypes.c_void_p)
    L.vsmock_ssimu2.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [PP, PP] + [ctypes.c_int] * 7 + [
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p, ctypes.c_int]
    L.vsmock_gpuinfo.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    L.vsmock_butter.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 5 + [PP, PP, ctypes.c_int, ctypes.c_int,
                                                                          ctypes.c_double] + [ctypes.c_int] * 5 + [
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
        ctypes.c_char_p, ctypes.c_int]
    return L


def libpath():
    return api.lib_path().encode()


def run_ssimu2(frames_ref, frames_dist, w, h, stride, w2=None, h2=None, rgbs=(1, 1), num_stream=-1, gpu_id=-1,
               core_threads=4, nthreads=4):
    """frames_*: list of 3 (h, stride/4) float32 arrays per frame."""
    n = len(frames_ref)
    ptrs_r = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_ref for p in f])
    ptrs_d = (ctypes.c_void_p * (3 * n))(*[p.ctypes.data for f in frames_dist for p in f])
    scores = np.full(n, np.nan)
    matches = ctypes.c_int(0)
    err = ctypes.create_string_buffer(4096)
    rc = mock().vsmock_ssimu2(libpath(), w, h, w2 or w, h2 or h, n, ptrs_r, ptrs_d, stride, rgbs[0], rgbs[1],
                              num_stream, gpu_id, core_threads, nthreads,
                              scores.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(matches), err,
                              len(err))
    return rc, scores, matches.value, err.value.decode()


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


def padded_frames(n, w, h, stride_px, seed):
    """n pairs of planar float RGB in [0, 1] with rows padded to stride_px floats."""
    from tests.gpu_helpers import rgb_pair

    ref, dist = [], []
    for i in range(n):
        r, d = rgb_pair(w, h, api.PF_RGB_F32, seed + i)
        pr, pd = [], []
        for a, b in zip(r, d):
            x = np.zeros((h, stride_px), np.float32)
            y = np.full((h, stride_px), 7.0, np.float32)  # padding that must not be read
            x[:, :w] = a
            y[:, :w] = b
            pr.append(x)
            pd.append(y)
        ref.append(pr)
        dist.append(pd)
    return ref, dist


# ---------------------------------------------------------------- CPU
@pytest.mark.skipif(not os.path.exists(os.path.join(VS_HEADER_DIR, "VapourSynth4.h")) or not shutil.which("g++"),
                    reason="VapourSynth4.h not available")
def test_vs4_declarations_match_vapoursynth_header(tmp_path):
    exe = str(tmp_path / "vs4chk")
    subprocess.run(["g++", "-std=c++17", "-I", VS_HEADER_DIR, os.path.join(ROOT, "tests", "vs4_layout_check.cpp"),
                    "-o", exe], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
    assert "0 mismatches" in r.stdout


@needs_mock
def test_plugin_registers_the_reference_interface():
    buf = ctypes.create_string_buffer(8192)
    assert mock().vsmock_registration(libpath(), buf, len(buf)) == 0
    lines = buf.value.decode().splitlines()
    # src/VshipLib.cpp:59-63 of the reference
    assert lines[0] == f"config|com.lumen.vship|vship|VapourSynth GPU Perceptual Metrics|{(3 << 16) | 3}|{4 << 16}"
    fns = {l.split("|")[1]: l.split("|")[2:] for l in lines[1:]}
    assert fns["SSIMULACRA2"] == ["reference:vnode;distorted:vnode;numStream:int:opt;gpu_id:int:opt;", "clip:vnode;"]
    assert fns["GpuInfo"] == ["gpu_id:int:opt;", "gpu_human_data:data;"]
    # src/butter/vapoursynth.hpp argument names, src/VshipLib.cpp:61
    assert fns["BUTTERAUGLI"] == ["reference:vnode;distorted:vnode;qnorm:int:opt;intensity_multiplier:float:opt;"
                                  "distmap:int:opt;numStream:int:opt;gpu_id:int:opt;", "clip:vnode;"]
    assert set(fns) == {"SSIMULACRA2", "GpuInfo", "BUTTERAUGLI", "CVVDP"}


@needs_mock
def test_differing_inputs_are_rejected_before_any_gpu_work():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4, w2=32)
    assert rc == 1, err
    assert err.startswith("VshipException\nDifferingInputType: Vship received 2 videos with different properties")
    assert "\n - At line " in err and err.endswith("\n")


@needs_mock
def test_butteraugli_differing_inputs_are_rejected():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, _, err = run_butter(ref, dist, 64, 48, 64 * 4, h2=40)
    assert rc == 1, err
    assert err.startswith("VshipException\nDifferingInputType:")


@needs_mock
def test_non_rgbs_input_without_resize_plugin_is_an_error():
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4, rgbs=(0, 1))
    assert rc == 1, err  # 3 would mean a node reference leaked or was freed twice
    assert err.startswith("VshipException\nNonRGBSInput:")


@needs_mock
def test_no_gpu_is_reported_in_the_reference_format():
    import vship_amd

    try:
        if vship_amd.device_count() > 0:
            pytest.skip("a GPU is present")
    except api.VshipError:
        pass  # no device: the case under test
    ref, dist = padded_frames(1, 64, 48, 64, 1)
    rc, _, _, err = run_ssimu2(ref, dist, 64, 48, 64 * 4)
    assert rc == 1
    assert err.startswith("VshipException\nDeviceCountError:") or err.startswith("VshipException\nNoDeviceDetected:")
    buf = ctypes.create_string_buffer(4096)
    assert mock().vsmock_gpuinfo(libpath(), -1, buf, len(buf)) == 1
    assert buf.value.decode().startswith("VshipException\n")


# ---------------------------------------------------------------- GPU
@pytest.mark.gpu
@needs_mock
def test_ssimulacra2_filter_scores_every_frame():
    from tests import oracle_lib as ol
    import vship_amd

    vship_amd.gpu_full_check(0)
    w, h, stride_px, n = 320, 240, 336, 6
    ref, dist = padded_frames(n, w, h, stride_px, 50)
    print(f"case: vship.SSIMULACRA2 via the mock host, {n} frames {w}x{h}, numStream 3, 4 threads", flush=True)
    rc, scores, matches, err = run_ssimu2(ref, dist, w, h, stride_px * 4, num_stream=3, gpu_id=0, nthreads=4)
    assert rc == 0, err
    assert matches == n  # output frame = distorted frame, frame n requests frame n of both clips
    with vship_amd.SSIMU2Handler(w, h) as hd:
        direct = np.array([hd.compute_float([p[:, :w] for p in r], [p[:, :w] for p in d]) for r, d in zip(ref, dist)])
    oracle = np.array([ol.score_rgb(ol.SO_RGB_F32, [p[:, :w] for p in r], [p[:, :w] for p in d], "f64")[0]
                       for r, d in zip(ref, dist)])
    print(f"filter {scores}\noracle {oracle}", flush=True)
    assert np.array_equal(scores, direct)
    assert np.all(np.abs(scores - oracle) <= 1e-3)


@pytest.mark.gpu
@needs_mock
def test_gpuinfo():
    buf = ctypes.create_string_buffer(8192)
    assert mock().vsmock_gpuinfo(libpath(), -1, buf, len(buf)) == 0
    assert buf.value.decode().startswith("GPU 0: ")
    assert mock().vsmock_gpuinfo(libpath(), 0, buf, len(buf)) == 0
    text = buf.value.decode()
    assert "WarpSize: 64" in text and "PassKernelCheck : 1" in text
    assert mock().vsmock_gpuinfo(libpath(), 99, buf, len(buf)) == 1
    assert buf.value.decode().startswith("VshipException\nBadDeviceArgument:")


@pytest.mark.gpu
@needs_mock
def test_butteraugli_filter_scores_every_frame():
    from tests import butter_lib as bl
    import vship_amd

    vship_amd.gpu_full_check(0)
    w, h, stride_px, n = 200, 120, 208, 4
    ref, dist = padded_frames(n, w, h, stride_px, 70)
    print(f"case: vship.BUTTERAUGLI via the mock host, {n} frames {w}x{h}, qnorm 3, distmap", flush=True)
    rc, norms, maps, _, err = run_butter(ref, dist, w, h, stride_px * 4, qnorm=3, distmap=1, num_stream=2, gpu_id=0)
    assert rc == 0, err
    for i in range(n):
        r = [np.ascontiguousarray(p[:, :w]) for p in ref[i]]
        d = [np.ascontiguousarray(p[:, :w]) for p in dist[i]]
        want, wmap = bl.butteraugli(r, d, 3, 203.0, diffmap=True)
        assert np.all(np.abs(norms[i] - want) <= 1e-4 * want), (norms[i], want)
        assert np.all(np.abs(maps[i] - wmap) <= 1e-6 + 1e-4 * np.abs(wmap))
    print("case: vship.BUTTERAUGLI defaults, output = distorted frames", flush=True)
    rc, norms2, _, matches, err = run_butter(ref, dist, w, h, stride_px * 4, gpu_id=0)
    assert rc == 0, err
    assert matches == n
    for i in range(n):
        r = [np.ascontiguousarray(p[:, :w]) for p in ref[i]]
        d = [np.ascontiguousarray(p[:, :w]) for p in dist[i]]
        want = bl.butteraugli(r, d, 2, 203.0)
        assert np.all(np.abs(norms2[i] - want) <= 1e-4 * want)

